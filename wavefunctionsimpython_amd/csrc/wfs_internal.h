// Internal declarations shared by the wfsim translation units.
#pragma once

#include <hip/hip_ext.h>

#include "wfs_common.h"
#include "../../include/wfsim.h"

namespace wfs {

// Per-kernel timing (wfs_plan_set_profiling): inside a timed launch scope the host thread's
// pointer names an event pair.  The first launch made through WFS_LAUNCH records its start
// timestamp into `start`, every launch its end timestamp into `stop` (hipExtLaunchKernel takes
// them from the dispatch itself), so the pair spans the scope's kernels alone, without the
// dispatch time that events recorded beside a launch add to it.  Null outside a timed scope.
struct LaunchTiming {
  hipEvent_t start = nullptr, stop = nullptr;
  int launches = 0;
};
LaunchTiming*& launch_timing();

}  // namespace wfs

// hipLaunchKernelGGL, or hipExtLaunchKernelGGL with the events of the current timed scope
#define WFS_LAUNCH(kernel, grid, block, lds, stream, ...)                                              \
  do {                                                                                               \
    ::wfs::LaunchTiming* wfs_lt_ = ::wfs::launch_timing();                                          \
    if (wfs_lt_) {                                                                                   \
      hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), std::uint32_t(lds), stream,             \
                            wfs_lt_->launches++ ? nullptr : wfs_lt_->start, wfs_lt_->stop, 0u,        \
                            __VA_ARGS__);                                                            \
    } else {                                                                                         \
      hipLaunchKernelGGL(kernel, grid, block, lds, stream, __VA_ARGS__);                            \
    }                                                                                                \
  } while (0)

namespace wfs {

enum { LAYOUT_ROWS = 0, LAYOUT_COLS = 1 };
enum { MODE_FWD = 0, MODE_INV = 1, MODE_FKI = 2, MODE_IVF = 3 };

// Parameters of one axis pass (see wfs_fft.hip).
struct PassParams {
  const void* tw = nullptr;     // forward twiddles exp(-2 pi i j / N), j < N, in the state precision
  long rows = 0;                // ROWS layout: number of contiguous rows of length N
  long inner = 0;               // COLS layout: stride of the axis = columns per outer slab
  const void* kx = nullptr;     // FKI: per-axis kinetic factors (normalisation folded in), state precision
  const void* ky = nullptr;
  const void* kz = nullptr;
  long nz = 1;                  // FKI: column index -> (ky, kz) = (c / nz, c % nz)
  const void* pv = nullptr;     // IVF: potential phase table cexp_i(-dt V), state precision, state layout
  // IVF with a separable real potential V = fx(x) + fy(y) + fz(z): per-axis phase factors
  // exp(-i dt f_d), [packet][n_d], state precision (used instead of pv when spx is set)
  const void* spx = nullptr;
  const void* spy = nullptr;
  const void* spz = nullptr;
  long snx = 0, sny = 0;        // grid lengths along x and y (row -> packet, x, y)
  double scale = 1.0;           // FWD/INV: output scale (1 = none)
};

bool fft_length_supported(int n);
template <typename R>
int launch_fft_pass_t(int n, int layout, int mode, void* data, const PassParams& p, long outer, int variant,
                      hipStream_t s);
int launch_fft_pass(int precision, int n, int layout, int mode, void* data, const PassParams& p, long outer,
                    int variant, hipStream_t s);
// Fused y-z pass of the split step on whole x planes (ny == nz, plane held in LDS):
// y inverse, z inverse, potential phase, z forward, y forward, in place.
// Order of a potential phase table read by yz_lines_kernel in its z stage: the table of
// row group g (LY rows of NZ points) holds, at (r * n1 + i) * threads + t, the factor of row
// (t mod rh) + r rh, z = (t div rh) + n2z i: the points thread t multiplies in round r, so
// that each wavefront reads one contiguous run.  Group g covers x plane g div b (over all
// packets) and rows y = b * row + (g mod b).
struct LinesOrder {
  long ny = 0, nz = 0;
  int ly = 0, b = 1, rh = 0, n2z = 0, n1 = 0, threads = 0;
};
// natural (C-order) index of element o of a lines-ordered table
__host__ __device__ inline long lines_source(long o, const LinesOrder& m) {
  const long per = long(m.ly) * m.nz;
  const long g = o / per, rem = o % per;
  const long ri = rem / m.threads;
  const int t = int(rem % m.threads);
  const int r = int(ri / m.n1), i = int(ri % m.n1);
  const long row = long(m.b) * ((t % m.rh) + long(r) * m.rh) + g % m.b;
  return (g / m.b) * m.ny * m.nz + row * m.nz + (t / m.rh) + long(m.n2z) * i;
}

// boxes measured in one launch_box_moments call (and one LINES_OBSERVE launch)
constexpr int kBoxesPerPass = 4;
// Simpson box moments taken inside a LINES_OBSERVE launch: one record of 7 sums per row group
// and box, partials[group][box][7] (the summing kernel applies dx dy dz / 27)
struct LinesMoments {
  int nb = 0;
  int bot[kBoxesPerPass][3] = {}, num[kBoxesPerPass][3] = {};
  int n[3] = {};
  double dx[3] = {};
  double* partials = nullptr;
};

struct YzPlaneParams {
  // y twiddles of the inverse / forward y transforms and the z twiddles, state precision:
  // the length-N axis tables, or (slab scheme) each thread's twiddles [u][k1]
  const void* tw_yi = nullptr;
  const void* tw_yf = nullptr;
  const void* tw_z = nullptr;
  const void* tw_zf = nullptr;  // LINES_OBSERVE: z twiddles of the forward z transform
  LinesMoments mom;             // LINES_OBSERVE, LINES_ENTRY_OBSERVE: boxes and partial records
  const void* pv = nullptr;     // potential phase table, state layout (null when spx is set)
  const int* pv_planes = nullptr;  // slab scheme: the table plane holding each x plane's factors
                                   // (the first of a run of identical planes), or null
  const void* spx = nullptr;    // separable potential: per-axis phase factors [packet][n_d]
  const void* spy = nullptr;
  const void* spz = nullptr;
  long nx = 1;                  // x planes per packet
  long planes = 0;              // packets * nx
};
// halves: lines per thread (1 or 2), the plane passing through LDS in that many rounds
bool yz_plane_supported(int precision, int ny, int nz, int halves);
int launch_yz_plane(int precision, int n, int halves, void* data, const YzPlaneParams& p, hipStream_t s);

// Two-launch split step for complex64 grids with ny = nz = 256 or 192 (wfs_slab.hip): the y
// transform split in a length-ny/4 stage (on rows y = 4 a + b, with the z work) and a
// length-4 stage (with the x pass); see there.
enum {
  LINES_FULL = 0,    // y inverse, z inverse, potential phase, z forward, y forward
  LINES_ENTRY = 1,   // z forward, y forward (position space in)
  LINES_EXIT = 2,    // y inverse, z inverse (position space out)
  LINES_OBSERVE = 3,  // y inverse, z inverse, box moments of the position-space rows, z forward, y forward
  LINES_ENTRY_OBSERVE = 4  // box moments of the position-space rows, z forward, y forward
};
struct XSlabParams {
  const void* tw_x = nullptr;   // forward twiddles exp(-2 pi i j / nx)
  const void* tw_y = nullptr;   // forward twiddles exp(-2 pi i j / ny) (the y-stage twiddles)
  const void* kx = nullptr;     // per-axis kinetic factors, normalisation folded in
  const void* ky = nullptr;
  const void* kz = nullptr;
  long packets = 0;             // grids stacked along x
};
bool slab_supported(int precision, const int n[3]);
// Per x plane (packets * nx planes of plane_elems complex64 factors, contiguous per plane): src[p]
// = the first plane of the run of bitwise-identical consecutive planes holding p, within its
// packet.  flags: workspace of packets * nx ints.
// flags[p] = plane p of the table equals plane p - 1 of its packet, for planes first .. first +
// count - 1 (the others keep theirs); then src[p] = the first plane of p's run, for every plane
int launch_plane_dedup(const void* table, long plane_elems, int nx, int packets, long first, long count, int* flags,
                       int* src, hipStream_t s);
// thread geometry of the slab scheme's y-z launch: y lines split over uy threads, z lines
// over uz threads, n1 points per thread; y-line twiddle stride b in the ny table
struct SlabRows {
  int b, uy, uz, n1;
};
SlabRows slab_rows(int n);
// the potential table order of the slab scheme's y-z launch (ny = nz = n)
LinesOrder slab_lines_order(int n);
// YzPlaneParams.planes = packets * nx; variant 2 = persistent (one block per CU walking the
// row groups, loading the next group during the current one's arithmetic), else one block
// per group
int launch_yz_slab(int n, int mode, int variant, void* data, const YzPlaneParams& p, hipStream_t s);
// variant 2: 8 z points per block, else 16
int launch_x_slab(int nx, int n, int variant, void* data, const XSlabParams& p, hipStream_t s);

// pv = cexp_i(-dt V) elementwise in the state precision
int launch_potential_phase(int precision, long count, const float2* V, double dt, void* pv, hipStream_t s);
// the same table written in lines order: pv[o] = cexp_i(-dt V[lines_source(o)])
int launch_potential_phase_lines(int precision, long count, const float2* V, double dt, void* pv,
                                 const LinesOrder& m, hipStream_t s);
// move a table between natural and lines order (to_lines: dst[o] = src[lines_source(o)],
// else dst[lines_source(o)] = src[o])
int launch_table_reorder(int precision, long count, const void* src, void* dst, const LinesOrder& m, bool to_lines,
                         hipStream_t s);

// power series: one Horner stage  u_out = psi + (-i dt / n) H u_in
int launch_horner_pair(int precision, int batch, int nx, int ny, int nz, const void* u_in, void* u_out,
                       const void* psi, const float2* V, double c1, double c2, double inv_2m, double inv_dx2,
                       int variant, hipStream_t s);
int launch_horner_stage(int precision, int batch, int nx, int ny, int nz, const void* u_in, void* u_out,
                        const void* psi, const float2* V, double dt_over_n, double inv_2m, double inv_dx2,
                        int tile_x, hipStream_t s);

// observables over one box of every packet: writes 7 doubles per packet to out (device)
int launch_box_moments(int precision, int batch, int nx, int ny, int nz, const double dx[3], const void* psi,
                       const wfs_box* boxes, int nb, double* partials, long max_partials, double* out,
                       long out_stride, hipStream_t s);
long box_moments_partials_needed(const wfs_box& box);
// out[p * out_stride + 7 b + q] = scale * sum over the nparts records partials[p][r][b][q]
int launch_sum_partials(const double* partials, long nparts, int nb, int batch, double scale, double* out,
                        long out_stride, hipStream_t s);

// probability current through a plane: out7[0] = flux (FP64), density optional [cu][cv]
long plane_current_partials_needed(int cu, int cv);
int launch_plane_current(int precision, int nx, int ny, int nz, const double dx[3], const void* psi,
                         const double center[3], const double normal[3], const double size[2], const int count[2],
                         double mass, void* density, double* partials, long max_partials, double* out7,
                         hipStream_t s);

}  // namespace wfs
