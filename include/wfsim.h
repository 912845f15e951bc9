/*
 * wfsim.h -- C ABI of the MI355X (gfx950) propagation library.
 *
 * Drop-in boundary for the WaveFunctionSimPython propagation path.  The
 * reference has no FFI: its boundary is a Python call surface plus the
 * positional argument lists of CuPy RawKernels.  Each entry point below
 * states which reference interface it replaces (paths relative to the
 * reference repository).
 *
 * Conventions
 *  - All data pointers are DEVICE pointers owned by the caller (e.g. torch
 *    tensors).  The library owns only plans (twiddle/kinetic tables and
 *    scratch buffers).
 *  - Grids are C-order (x slowest, z fastest), interleaved complex.  The
 *    state is complex64 (WFS_C64) or complex128 (WFS_C128); the potential V
 *    is always complex64 (sources/cuda_kernels/potential_operator.cu:7).
 *  - Every call is ordered on the given stream (hipStream_t passed as void*;
 *    NULL = the null stream) and returns 0 on success, a hipError_t value on a
 *    HIP failure, or a WFS_ERR_* code; wfs_last_error() describes the last
 *    failure of the calling thread.
 *  - A plan is used by one host thread at a time.
 */
#ifndef WFSIM_H
#define WFSIM_H

#ifdef __cplusplus
extern "C" {
#endif

#define WFS_C64 0
#define WFS_C128 1

#define WFS_METHOD_FOURIER 0
#define WFS_METHOD_POWER_SERIES 1

#define WFS_ERR_INVALID 1000
#define WFS_ERR_UNSUPPORTED 1001
#define WFS_ERR_ALLOC 1002

typedef struct wfs_plan wfs_plan;

/* Sub-box of the grid for an observable: voxel origin and sample counts. */
typedef struct wfs_box {
  int bottom[3];
  int count[3];
} wfs_box;

/* Library version string. */
const char* wfs_version(void);

/* Description of the last error raised on the calling thread ("" if none). */
const char* wfs_last_error(void);

/*
 * Create a propagation plan.
 * Replaces the per-run set-up of SimState.initialize_state
 * (sources/sim_state.py:257-294): kernel JIT, wave numbers
 * (2*pi*fftfreq, :275-278) and ping-pong buffers (:291-294).
 *   n[3]      grid size per axis (8,16,32,64,96,128,192,256,512)
 *   dx[3]     voxel size in Bohr radii
 *   dt        time step in hbar/Hartree
 *   mass      particle mass in electron masses
 *   precision WFS_C64 or WFS_C128 (double_precision_calculation)
 *   method    WFS_METHOD_FOURIER or WFS_METHOD_POWER_SERIES
 *   order     Taylor order p of the power series (order_of_approximation)
 */
int wfs_plan_create(const int n[3], const double dx[3], double dt, double mass, int precision, int method,
                    int order, wfs_plan** out);

/*
 * Plan for `batch` independent grids of identical shape, voxel size, dt and
 * mass, stored back to back (packet b starts at b * n[0]*n[1]*n[2]).  Every
 * entry point then works on the whole stack in one launch per pass; each
 * packet's result is bit-identical to running it alone.  Replaces nothing in
 * the reference (one SimState per process, sources/sim_state.py:21-60).
 * wfs_observables on a batched plan writes out[b][box][7].
 */
int wfs_plan_create_batched(const int n[3], const double dx[3], double dt, double mass, int precision, int method,
                            int order, int batch, wfs_plan** out);

/* Release a plan and its device tables. */
int wfs_plan_destroy(wfs_plan* plan);

/*
 * Advance psi in place by nsteps time steps.
 * Replaces SimState.evolve_state (sources/sim_state.py:754-762) with its two
 * branches _fft_time_evolution (:641-697, kinetic_operator.cu:4-24,
 * potential_operator.cu:4-16) and _power_series_time_evolution (:721-752,
 * power_series_operator.cu:3-65).  psi is in position space on entry and exit;
 * consecutive split-step half kicks are merged internally.
 *   psi  device state, complex64 or complex128 per plan precision
 *   V    device complex64 potential
 */
int wfs_step(wfs_plan* plan, void* psi, const void* V, int nsteps, void* stream);

/*
 * Table-copy schedule of the split-step (no reference counterpart; part of the
 * precision policy).  Complex64 plans hold 8 dithered copies of every per-axis
 * twiddle and kinetic table whose mean is within 1/16 ulp of the exact value,
 * and step k of the run uses copy k mod 8, so that float rounding does not
 * repeat from step to step (complex128 plans hold one copy).  The plan keeps a
 * running index: every wfs_step / wfs_step_observe call advances it by its step
 * count, so the reference's loop of one evolve_state() per step
 * (sources/core_sim.py:42-55) cycles the copies as one long call does.  A
 * captured graph replays the copies of its capture (capture a multiple of 8
 * steps to keep the cycle).  A caller that needs a call to be the same map every
 * time (parareal's fine and coarse propagators) sets the index before the call.
 * wfs_plan_get_step_index returns the index in [0, 8) (-1 for a null plan).
 */
int wfs_plan_set_step_index(wfs_plan* plan, long index);
long wfs_plan_get_step_index(wfs_plan* plan);

/*
 * Store the potential of subsequent wfs_step calls made with V == NULL.
 * For the split-step the plan keeps the phase table cexp_i(-dt V) (the intent
 * of operators.init_potential_operator, sources/operators.py:41-55); callers
 * whose V changes (SimState.update_potential, sources/sim_state.py:466-509)
 * call it again.  wfs_step with V != NULL refreshes the table itself.
 */
int wfs_set_potential(wfs_plan* plan, const void* V, void* stream);

/*
 * Rebuild x planes [x_begin, x_end) of the stored table from V (no reference counterpart: the
 * dynamic-potential mode of SimState.update_potential, sources/sim_state.py:466-509, rebuilds V
 * whole each step; where only the planes a moving wall sweeps change, this refreshes only
 * those).  The caller promises that V differs from the V of the table's last build in those
 * planes only; the result is then the table wfs_set_potential(V) builds, bit for bit.  Without
 * a table of the plan's current kind it builds the whole table.  One packet only
 * (WFS_ERR_UNSUPPORTED for batched plans); 0 <= x_begin <= x_end <= n[0].
 */
int wfs_set_potential_planes(wfs_plan* plan, const void* V, int x_begin, int x_end, void* stream);

/*
 * Separable real potential V(x, y, z) = fx(x) + fy(y) + fz(z) (no reference
 * counterpart; harmonic wells and cos^2 lattices are of this form).  fx, fy,
 * fz are host arrays of batch * n[d] doubles (Hartree, packet-major).  The plan
 * keeps the per-axis factors exp(-i dt f_d) instead of an N^3 phase table, so
 * the potential pass reads 16 instead of 24 B per point.  Split-step plans
 * only; replaced by the next wfs_set_potential or wfs_step with V != NULL.
 * Synchronous (host upload).
 */
int wfs_set_potential_axes(wfs_plan* plan, const double* fx, const double* fy, const double* fz);

/*
 * Unitary 3-D DFT in place: direction -1 = forward (cupy.fft.fftn(norm="ortho"),
 * sources/sim_state.py:646,677), +1 = inverse (ifftn, :665,696).
 */
int wfs_fft3d(wfs_plan* plan, void* data, int direction, void* stream);

/*
 * Simpson-weighted moments of |psi|^2 over each box, accumulated in FP64.
 * Replaces volume_probability_kernel (sources/cuda_kernels/volume_probability.cu:4-47)
 * and expected_location_kernel (expected_location.cu:3-72) as launched by
 * VolumeProbability.calculate / ExpectedLocation.calculate
 * (sources/measurement.py:153-175, 354-384).
 * out (device, 7 doubles per box): P, E[x], E[y], E[z], E[x^2], E[y^2], E[z^2]
 * (each already multiplied by dx*dy*dz; not divided by P).
 */
int wfs_observables(wfs_plan* plan, const void* psi, const wfs_box* boxes, int nboxes, double* out,
                    void* stream);

/*
 * wfs_observables without a plan and without allocating (no counterpart in the
 * reference).  For callers that cannot allocate device memory: code under hipGraph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  Same results
 * and out layout as wfs_observables on a plan of this grid (out[packet][box][7],
 * batch packets of n[0] x n[1] x n[2] stored back to back).  The partial records
 * go to a static 2 MiB device workspace per GPU (262144 doubles): every pass of
 * up to 4 boxes needs ceil(rows / 32) * 7 * boxes * batch records, rows = the x-y
 * extent of the boxes' bounding box; larger requests return WFS_ERR_UNSUPPORTED.
 * The first call on a device looks the workspace up (hipGetSymbolAddress); no
 * call allocates.  Calls that overlap on one GPU share the workspace and race.
 */
int wfs_observables_noalloc(const int n[3], const double dx[3], int precision, int batch, const void* psi,
                            const wfs_box* boxes, int nboxes, double* out, void* stream);

/*
 * nsteps steps with the box observables of the state before steps 0,
 * observe_every, 2 observe_every, ... (< nsteps) written to
 * out[o][packet][box][7] (device; o < ceil(nsteps / observe_every)): the
 * results of alternating wfs_observables and wfs_step(observe_every) calls,
 * which is how core_sim.run_iteration measures before each step
 * (sources/core_sim.py:19-65, sources/measurement.py:644-704), here with
 * observe_every steps between measurements.  With the slab scheme (WFS_K_YZ_SLAB) the exit launch of one
 * stretch, the observation and the entry launch of the next run as one launch;
 * the state is bit-identical to the alternating calls and the observables
 * agree to FP64 rounding (summation order).  V as for wfs_step.
 */
int wfs_step_observe(wfs_plan* plan, void* psi, const void* V, int nsteps, int observe_every,
                     const wfs_box* boxes, int nboxes, double* out, void* stream);

/*
 * Called by wfs_step_observe_between after each step (0-based index `step`), once that step's
 * kernels are queued on the call's stream and before the next step's are; a nonzero return
 * stops the call with WFS_ERR_INVALID.
 */
typedef int (*wfs_between_steps)(void* ctx, int step);

/*
 * wfs_step_observe with a callback between steps (no reference counterpart: the reference's
 * dynamic mode updates the potential after every step, SimState.update_potential between
 * evolve_state calls, sources/core_sim.py:19-24, sources/sim_state.py:466-509).  The callback
 * may rewrite V and the plan's table (wfs_set_potential, wfs_set_potential_planes, the
 * generators) on the same stream; the next step reads what it queued.  It must not touch psi,
 * which holds an intermediate layout until the call returns, nor wait on the stream.  With the
 * slab scheme the steps run at the fused rate (no entry and exit launches per step).  With
 * observe_every = 1 the state and observables are those of one wfs_step_observe call per step
 * with the same potentials (bit for bit, as a stretch equals its alternating calls); with
 * larger intervals the kinetic half steps between observations merge as in any stretch (float
 * rounding).  Without the slab scheme each step is one call, bit for bit those calls.  between
 * may be NULL (then as wfs_step_observe).
 */
int wfs_step_observe_between(wfs_plan* plan, void* psi, const void* V, int nsteps, int observe_every,
                             const wfs_box* boxes, int nboxes, double* out, wfs_between_steps between, void* ctx,
                             void* stream);

/*
 * Probability current through a plane (sources/measurement.py:190-302,
 * sources/cuda_kernels/probability_current_density.cu:4-218).  count[2] are
 * the (odd) pixel counts, size[2] the plane's width and height in Bohr radii,
 * normal a unit vector.  out (device, 1 double) = Simpson integral of j.n over
 * the plane (FP64 accumulation); density (device, optional, state precision
 * real) = j.n per pixel as [count[0]][count[1]].  Pixels outside the grid
 * contribute zero.
 */
int wfs_plane_current(wfs_plan* plan, const void* psi, const double center[3], const double normal[3],
                      const double size[2], const int count[2], void* density, double* out, void* stream);

/*
 * Gaussian wave packet (sources/cuda_kernels/gaussian_wave_packet.cu:6-43,
 * common.cu:565-605), as launched by GaussianWavePacket.init_wave_packet
 * (sources/wave_packet.py:97-151).  Writes (overwrites) psi.
 */
int wfs_init_gaussian(const int n[3], const double dx[3], int precision, const double sigma[3],
                      const double r0[3], const double k0[3], void* psi, void* stream);

/*
 * Complex absorbing potential: V += i * pot (absorbing_potential.cu:3-307,
 * launched by AbsorbingBoundaryCondition.add_potential, sources/potential.py:108-158).
 */
int wfs_add_cap(const int n[3], const double dx[3], const double bottom[3], const double top[3], double offset,
                const double pot_pos[3], const double pot_neg[3], const double exp_pos[3],
                const double exp_neg[3], void* V, void* stream);

/*
 * wfs_add_cap on x planes [x_begin, x_end) of the n-grid V only (no reference counterpart; for
 * wfs_set_potential_planes): those planes get exactly the values of the whole-grid call, the
 * others are not touched.  0 <= x_begin <= x_end <= n[0].
 */
int wfs_add_cap_planes(const int n[3], const double dx[3], const double bottom[3], const double top[3],
                       double offset, const double pot_pos[3], const double pot_neg[3], const double exp_pos[3],
                       const double exp_neg[3], int x_begin, int x_end, void* V, void* stream);

/*
 * Potential wall with 0-4 slits: V += pot (potential_wall.cu:4-83, launched by
 * PotentialWall.add_potential, sources/potential.py:201-242).  normal must be
 * unit length.
 */
int wfs_add_wall(const int n[3], const double dx[3], double potential, const double center[3],
                 const double normal[3], double plateau, double slope, double slope_exponent, int slit_count,
                 double slit_spacing, double slit_width, double slit_rotation, void* V, void* stream);

/*
 * wfs_add_wall on x planes [x_begin, x_end) of the n-grid V only (no reference counterpart; for
 * wfs_set_potential_planes), as wfs_add_cap_planes.
 */
int wfs_add_wall_planes(const int n[3], const double dx[3], double potential, const double center[3],
                        const double normal[3], double plateau, double slope, double slope_exponent,
                        int slit_count, double slit_spacing, double slit_width, double slit_rotation, int x_begin,
                        int x_end, void* V, void* stream);

/*
 * Harmonic potential V += 1/2 m w^2 |r - c|^2 (formula of the uncalled helper
 * sources/potential.py:414-423; exposed here for the harmonic workloads).
 */
int wfs_add_harmonic(const int n[3], const double dx[3], double omega, double mass, const double center[3],
                     void* V, void* stream);

/* cos^2 optical lattice V += depth * sum_d cos^2(pi r_d / period) (no reference counterpart). */
int wfs_add_lattice(const int n[3], const double dx[3], double depth, double period, void* V, void* stream);

/*
 * Kinetic half-step operator table out = exp(-i |k|^2 dt / (4 m)) on the full
 * grid, k = 2*pi*fftfreq.  The intent of operators.init_kinetic_operator
 * (sources/operators.py:9-34, whose body references undefined names) with the
 * phase of kinetic_operator.cu:20-22.  out: state precision.
 */
int wfs_kinetic_operator(const int n[3], const double dx[3], double dt, double mass, int precision, void* out,
                         void* stream);

/*
 * Potential operator table out = cexp_i(-dt V) elementwise (count points).
 * The intent of operators.init_potential_operator (sources/operators.py:41-55)
 * with the phase of potential_operator.cu:10-15.  V complex64; out: state precision.
 */
int wfs_potential_operator(long count, const void* V, double dt, int precision, void* out, void* stream);

/*
 * Per-kernel timing (no reference counterpart: the reference times whole
 * iterations with time.time(), sources/core_sim.py:43-49).  While enabled,
 * every kernel launched for the plan records a pair of HIP events on its
 * stream: the kernel's own start and end (hipExtLaunchKernel events; on a
 * stream under capture, events recorded around the launch).  Kernels launched
 * together for one class (a moments pass and its sum) share one pair.
 * wfs_plan_profile() waits for the recorded events and returns,
 * per kernel class, the launch count and the summed duration in ms since the
 * last reset (reset != 0 clears the sums after reading).
 * Kernel classes: WFS_K_* below.
 */
#define WFS_K_Z_FWD 0
#define WFS_K_Z_INV 1
#define WFS_K_Z_IVF 2
#define WFS_K_Y_FWD 3
#define WFS_K_Y_INV 4
#define WFS_K_X_FKI 5
#define WFS_K_X_FWD 6
#define WFS_K_X_INV 7
#define WFS_K_HORNER 8
#define WFS_K_OBSERVABLES 9
#define WFS_K_YZ_PLANE 10
#define WFS_K_YZ_SLAB 11
#define WFS_K_X_SLAB 12
#define WFS_K_SLAB_EDGE 13
#define WFS_K_SLAB_OBSERVE 14
#define WFS_K_PV_PLANES 15
#define WFS_K_COUNT 16
int wfs_plan_set_profiling(wfs_plan* plan, int enable);

/*
 * FFT tile autotuner (no reference counterpart).  Times each tile variant
 * (0, 1, 2 = half, default, double the columns per block) of every
 * split-step pass on plan-owned scratch buffers and keeps the fastest per
 * kernel class.  Variants give bit-identical results.  Synchronises stream.
 * WFS_K_HORNER: variant 0 runs one launch per power-series stage, 1 to 5
 * run the stages in fused pairs on 16 x 64, 8 x 64 and 8 x 128 (y, z) tiles
 * of 256 threads, 8 x 128 tiles of 512 threads and 8 x 256 tiles of 1024
 * threads (default 5 where nz >= 256, 4 where nz >= 128, else 1); the
 * results are bit-identical.
 * WFS_K_YZ_PLANE: variant 1 (default where supported: complex64 planes of
 * ny = nz <= 128, complex128 planes of ny = nz <= 64) runs the y inverse, z
 * inverse / potential / z forward and y forward passes of each step as one
 * kernel on whole x planes held in LDS; variant 0 runs the three passes.  The
 * results are identical; the autotuner keeps the faster.
 * WFS_K_YZ_SLAB: variants 1 and 2 (the default where supported: complex64
 * grids with ny = nz = 256 or 192 and nx <= 256, 2 at 256 and 1 at 192; at
 * ny = nz = 128 variant 2 below 256 x planes) run a step as two launches, the
 * y transform split in a length-ny/B stage (B = 4; 2 at 128) fused with the z
 * work (kernel class WFS_K_YZ_SLAB; the entry and exit launches of a call
 * count as WFS_K_SLAB_EDGE, the launches of wfs_step_observe that also take
 * the box moments as WFS_K_SLAB_OBSERVE) and a length-B stage fused with the x pass
 * (WFS_K_X_SLAB); variant 2 runs the y-z launch as one persistent block per
 * CU that loads the next row group during the current one's arithmetic;
 * variant 0 runs the four axis passes.  The results agree to float rounding
 * (the length-ny y transform is factored differently); the slab variants
 * among themselves are bit-identical.  WFS_K_X_SLAB: variant 1 (default at
 * 192) puts 16 z points in an x block, variant 2 (default at 256) 8, pairing
 * the two blocks of a 128-byte line on one XCD.  The autotuner keeps the
 * fastest.  WFS_K_PV_PLANES (no launches of its own): variant 1 (default)
 * lets the slab scheme's y-z launch read one copy of each run of x planes
 * whose potential phase factors are bitwise equal (the planes away from walls
 * and absorbing x layers), so that those reads hit the caches instead of HBM;
 * variant 0 reads every plane's own copy.  The results are bit-identical.
 */
int wfs_plan_autotune(wfs_plan* plan, void* stream);
int wfs_plan_set_variant(wfs_plan* plan, int kernel_class, int variant);
int wfs_plan_get_variant(wfs_plan* plan, int kernel_class);
int wfs_plan_profile(wfs_plan* plan, int kernel_class, int reset, long* launches, double* total_ms);

/*
 * Strided batched GEMM on the bf16 matrix cores (no reference counterpart;
 * building block of the Fourier neural operator).
 *   C[b](m,n) = alpha * sum_k A[b](m,k) B[b](k,n) + beta * C[b](m,n)
 * dims = {batch, M, N, K}; strides (elements) = {sAb, sAm, sAk, sBb, sBk, sBn,
 * sCb, sCm, sCn}.  A, B, C are FP32 in memory; A and B are rounded to bf16
 * (nearest even) on load, products accumulate in FP32.  flags: WFS_GEMM_SPLIT_A
 * / WFS_GEMM_SPLIT_B stage that operand as bf16 hi + lo parts (about 16
 * significant bits; Ahi Bhi + Alo Bhi + Ahi Blo).  splits > 1 splits K over
 * blocks and needs work_elems >= batch*splits*M*N floats; the splits are
 * added in a fixed order.  Epilogue: bias (may be NULL) adds bias[m] after
 * alpha/beta; act (may be NULL) receives GELU of the result with C's strides;
 * WFS_GEMM_NO_C (with act) leaves C unwritten.
 */
#define WFS_GEMM_SPLIT_A 1
#define WFS_GEMM_SPLIT_B 2
#define WFS_GEMM_NO_C 4
int wfs_gemm_bf16(const float* A, const float* B, float* C, const long dims[4], const long strides[9], float alpha,
                  float beta, int flags, int splits, float* work, long work_elems, const float* bias, float* act,
                  void* stream);

/*
 * Fourier-neural-operator and parareal elementwise kernels (no reference
 * counterpart).  Activations are channel-major FP32 planes x[c][p].
 *  wfs_fno_bias_act       u += bias[c]; v = GELU(u) (act != 0) or v = u (v may be NULL)
 *  wfs_fno_gelu_backward  du = dv * GELU'(u)
 *  wfs_fno_axpby          y = alpha x + beta y (y not read when beta == 0)
 *  wfs_fno_pack           a = [pscale Re psi, pscale Im psi, vscale Re V] planes (V may be NULL: 2 planes)
 *  wfs_fno_unpack         complex64 psi = oscale (plane 0 + i plane 1)
 *  wfs_fno_mse            out2 = {sum (pred-target)^2, sum target^2} (FP64);
 *                         dpred = 2 (pred-target) gscale (optional)
 *  wfs_fno_adam           Adam step on FP32 master weights
 *  wfs_parareal_update    U = Gnew + Fold - Gold in FP64 (complex64 inputs, complex128 U);
 *                         out2 = {|U - Uprev|^2, |U|^2}
 *  wfs_narrow_c128        complex128 -> complex64
 * Workspaces (partials) need 8192 doubles.
 */
int wfs_fno_bias_act(float* u, const float* bias, float* v, long C, long P, int act, void* stream);
/*
 * wfs_fno_modes (no reference counterpart): complex per-mode channel mixing
 * O[o][q] = sum_i R[q][i][o] X[i][q] over m3 modes, R [m3][C][C] as separate real and
 * imaginary arrays, X and O [C][2][m3] (real parts, then imaginary parts); R rounded to
 * bf16, X as bf16 hi + lo, FP32 accumulation (the operand rounding of wfs_gemm_bf16).
 */
int wfs_fno_modes(const float* Rr, const float* Ri, const float* X, float* O, int m3, int C, void* stream);
/*
 * wfs_fno_layer (no reference counterpart): one network layer fused over the grid lines
 * (x, y) of an n^3 grid, all channels:  u = W v + Z Gz + b;  vn = GELU(u) (act) or u;
 * T = vn Fz.  W [C][cin], v [cin][n^3], Z [C][n^2][tm] (the layer's inverse spectral path
 * before its z stage; NULL: no spectral term), Gz [tm][n], Fz [n][tm]; outputs u, vn
 * [C][n^3] and T [C][n^2][tm] (the next layer's z stage), each optional (T needs vn).
 * W, Gz, Fz enter the matrix cores as bf16, v, Z and vn as bf16 hi + lo parts, FP32
 * accumulation.  n in {32, 64, 96, 128}; C, cin <= 32; tm a multiple of 4 up to 32;
 * u, vn, Z 16-byte aligned.
 */
int wfs_fno_layer(const float* W, const float* v, const float* Z, const float* bias, const float* Gz,
                  const float* Fz, float* u, float* vn, float* T, int cin, int C, int n, int tm, int act,
                  void* stream);
int wfs_fno_gelu_backward(const float* u, const float* dv, float* du, long n, void* stream);
/*
 * wfs_spectral_resample (no reference counterpart): the spectrum of an ndst grid from that of
 * an nsrc grid, complex64, batch grids stacked: dst[k] = scale * src[k + shift] for every
 * signed frequency k of dst (numpy fftfreq order) whose k + shift src holds, 0 for the others;
 * shift (integer frequencies per axis, |shift| <= nsrc) may be NULL for none.  Between ortho
 * wfs_fft3d transforms, scale = (|ndst| / |nsrc|)^(1/2) (point counts) keeps the values of a
 * band-limited function, and a shift s takes the band around frequency s to the band around 0
 * (the function times exp(-2 pi i s.j / n)); the inverse map uses -s.  The coarse-grid coarse
 * propagator of parareal runs its network on such a resampled state.  src and dst must not
 * overlap.
 */
int wfs_spectral_resample(const void* src, const int nsrc[3], void* dst, const int ndst[3], long batch, double scale,
                          const int shift[3], void* stream);
int wfs_fno_axpby(float* y, const float* x, float alpha, float beta, long n, void* stream);
int wfs_fno_pack(const void* psi, int precision, const void* V, float pscale, float vscale, float* a, long P,
                 void* stream);
int wfs_fno_unpack(const float* out, float oscale, void* psi, long P, void* stream);
int wfs_fno_mse(const float* pred, const float* target, float* dpred, double gscale, long n, double* partials,
                long partial_elems, double* out2, void* stream);
int wfs_fno_adam(float* w, const float* g, float* m, float* v, long n, float lr, float beta1, float beta2, float eps,
                 int step, void* stream);
int wfs_parareal_update(const void* gnew, const void* fold, const void* gold, void* U, long n, double* partials,
                        long partial_elems, double* out2, void* stream);
int wfs_narrow_c128(const void* U, void* out, long n, void* stream);

/*
 * hipGraph capture of library calls (no reference counterpart: the reference
 * launches every kernel from Python, sources/core_sim.py:42-55).
 * wfs_graph_begin starts capturing `stream` (not the null stream); every
 * wfs_* call ordered on it (and on streams joined to it by events) is
 * recorded instead of executed; wfs_graph_end instantiates the graph;
 * wfs_graph_launch replays it.  Plans used inside must not allocate during
 * capture: power-series scratch is allocated by a first uncaptured wfs_step,
 * the potential table by a first uncaptured wfs_set_potential, a plane-current
 * workspace by a first uncaptured wfs_plane_current; a call that would allocate
 * on a capturing stream returns WFS_ERR_INVALID instead.  A graph replays the
 * table copies of its capture (wfs_plan_set_step_index).
 */
typedef struct wfs_graph wfs_graph;
int wfs_graph_begin(void* stream, wfs_graph** out);
int wfs_graph_end(wfs_graph* graph, void* stream);
int wfs_graph_launch(wfs_graph* graph, void* stream);
long wfs_graph_node_count(wfs_graph* graph);
int wfs_graph_destroy(wfs_graph* graph);

#ifdef __cplusplus
}
#endif

#endif /* WFSIM_H */
