// Initial state and potential generators (gfx950).  Elementwise, run at set-up
// or when walls move; one thread per voxel over the C-order grid.
#include "wfs_internal.h"

#include <cmath>

namespace wfs {

namespace {

constexpr int kGenThreads = 256;

// grid size and the window of x planes one launch covers: element i of the launch's
// (offset) array is voxel (x0 + i / (ny nz), (i / nz) mod ny, i mod nz)
struct Dims {
  int nx, ny, nz;
  int x0, planes;
};

Dims full_dims(const int n[3]) { return Dims{n[0], n[1], n[2], 0, n[0]}; }

__device__ __forceinline__ void voxel_of(long i, const Dims& d, int& x, int& y, int& z) {
  z = int(i % d.nz);
  const long r = i / d.nz;
  y = int(r % d.ny);
  x = d.x0 + int(r / d.ny);
}

__device__ __forceinline__ long window_points(const Dims& d) { return long(d.planes) * d.ny * d.nz; }

unsigned blocks_for(long n) { return unsigned((n + kGenThreads - 1) / kGenThreads); }

// ---------------------------------------------------------------------------
// Gaussian wave packet, evaluated in FP64 and rounded to the state precision.
// sources/cuda_kernels/common.cu:565-605 with the coordinates of
// gaussian_wave_packet.cu:36-40: a = 2 sigma,
// psi = prod_d (2/(pi a_d^2))^(1/4) exp(-(r_d - r0_d)^2 / a_d^2) * exp(i k0.r).
// ---------------------------------------------------------------------------
struct GaussParams {
  double dx[3], sigma[3], r0[3], k0[3];
};

template <typename R>
__global__ __launch_bounds__(kGenThreads) void gaussian_kernel(cx<R>* __restrict__ psi, Dims d, GaussParams p) {
  const long n = window_points(d);
  const long i = long(blockIdx.x) * kGenThreads + threadIdx.x;
  if (i >= n) return;
  int v[3];
  voxel_of(i, d, v[0], v[1], v[2]);
  const int N[3] = {d.nx, d.ny, d.nz};
  double amp = 1.0, phase = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double r = p.dx[k] * double(v[k] - N[k] / 2);
    const double a = 2.0 * p.sigma[k];
    const double e = r - p.r0[k];
    amp *= pow(2.0 / M_PI / a / a, 0.25) * exp(-e * e / a / a);
    phase += p.k0[k] * r;
  }
  double s, c;
  sincos(phase, &s, &c);
  psi[i] = mk<R>(R(amp * c), R(amp * s));
}

// ---------------------------------------------------------------------------
// Complex absorbing potential, float32 as in the reference
// (sources/cuda_kernels/absorbing_potential.cu:38-306).  Every branch of the
// reference needs each axis to be strictly below the box, strictly inside, or
// strictly above it; a voxel on a box plane gets nothing.  Three "outside"
// axes use the corner blend, two use an edge blend (including the
// reference's own mix() parameters of :162-276), one uses a face ramp.
// ---------------------------------------------------------------------------
struct CapParams {
  float dr[3], bottom[3], top[3], offset, ppos[3], pneg[3], epos[3], eneg[3];
};

constexpr float kPiF = 3.14159265358979323846f;

__device__ __forceinline__ float mixf(float x, float y, float z, float u, float v) {
  // common.cu:394-397
  return z * v + (1.0f - v) * (y * u + x * (1.0f - u));
}
__device__ __forceinline__ float angf(float num, float den) { return atanf(num / den) / kPiF * 2.0f; }
__device__ __forceinline__ float len3f(float a, float b, float c) { return sqrtf(a * a + b * b + c * c); }

__global__ __launch_bounds__(kGenThreads) void cap_kernel(float2* __restrict__ V, Dims dm, CapParams p) {
#pragma clang fp contract(off)
  const long n = window_points(dm);
  const long i = long(blockIdx.x) * kGenThreads + threadIdx.x;
  if (i >= n) return;
  int vx[3];
  voxel_of(i, dm, vx[0], vx[1], vx[2]);
  const int NN[3] = {dm.nx, dm.ny, dm.nz};
  float d[3], bh[3], corner[3], e3[3], p3[3];
  int cls[3];  // -1 below, 0 strictly inside, +1 above, 2 on a plane
  int nside = 0;
  bool tie = false;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float Nf = float(NN[k]);
    const float r = p.dr[k] * float(vx[k]) - 0.5f * Nf * p.dr[k];   // :52
    bh[k] = 0.5f * Nf * p.dr[k];                                    // :53
    const float dB = r - p.bottom[k];
    const float dT = r - p.top[k];
    if (dB < 0.0f) {
      cls[k] = -1; d[k] = dB; corner[k] = p.bottom[k]; e3[k] = p.eneg[k]; p3[k] = p.pneg[k]; ++nside;
    } else if (dT > 0.0f) {
      cls[k] = 1; d[k] = dT; corner[k] = p.top[k]; e3[k] = p.epos[k]; p3[k] = p.ppos[k]; ++nside;
    } else if (dB > 0.0f && dT < 0.0f) {
      cls[k] = 0; d[k] = 0.0f; corner[k] = 0.0f; e3[k] = 0.0f; p3[k] = 0.0f;
    } else {
      cls[k] = 2; tie = true; d[k] = 0.0f; corner[k] = 0.0f; e3[k] = 0.0f; p3[k] = 0.0f;
    }
  }
  float pot = 0.0f;
  const float off = p.offset;
  if (!tie && nside > 0) {
    float m3[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) m3[k] = bh[k] - (fabsf(corner[k]) + (cls[k] != 0 ? off : 0.0f));
    if (nside == 3) {
      // corners, :58-155
      const float u = angf(fabsf(d[1]), fabsf(d[0]));
      const float v = angf(fabsf(d[2]), sqrtf(d[0] * d[0] + d[1] * d[1]));
      const float maxD = mixf(m3[0], m3[1], m3[2], u, v);
      const float ex = mixf(e3[0], e3[1], e3[2], u, v);
      const float mp = mixf(p3[0], p3[1], p3[2], u, v);
      pot = powf(fmaxf(len3f(d[0], d[1], d[2]) - off, 0.0f) / maxD, ex) * mp;
    } else if (nside == 2) {
      // edges, :156-278
      if (cls[2] == 0) {
        // x-y edges (:157-166, :177-186, :239-248, :259-268): u from the x-y angle, v = 0
        const float u = angf(fabsf(d[1]), fabsf(d[0]));
        const float maxD = mixf(m3[0], m3[1], m3[2], u, 0.0f);
        const float ex = mixf(e3[0], e3[1], 0.0f, u, 0.0f);
        const float mp = mixf(p3[0], p3[1], 0.0f, u, 0.0f);
        pot = powf(fmaxf(len3f(d[0], d[1], 0.0f) - off, 0.0f) / maxD, ex) * mp;
      } else if (cls[1] == 0) {
        // x-z pillar edges (:198-237): v from the z-x angle, u = 0
        const float v = angf(fabsf(d[2]), fabsf(d[0]));
        const float maxD = mixf(m3[0], m3[1], m3[2], 0.0f, v);
        const float ex = mixf(e3[0], 0.0f, e3[2], 0.0f, v);
        const float mp = mixf(p3[0], 0.0f, p3[2], 0.0f, v);
        pot = powf(fmaxf(len3f(d[0], 0.0f, d[2]) - off, 0.0f) / maxD, ex) * mp;
      } else {
        // y-z edges (:167-176, :187-196, :249-258, :269-278): v from the z-y angle.
        // The reference blends maxD with u = 1 only on the (+y,+z) edge and
        // exponent / outer potential with u = 0 only on the (+y,-z) edge.
        const float v = angf(fabsf(d[2]), fabsf(d[1]));
        const float um = (cls[1] > 0 && cls[2] > 0) ? 1.0f : 0.0f;
        const float ue = (cls[1] > 0 && cls[2] < 0) ? 0.0f : 1.0f;
        const float maxD = mixf(m3[0], m3[1], m3[2], um, v);
        const float ex = mixf(0.0f, e3[1], e3[2], ue, v);
        const float mp = mixf(0.0f, p3[1], p3[2], ue, v);
        pot = powf(fmaxf(len3f(0.0f, d[1], d[2]) - off, 0.0f) / maxD, ex) * mp;
      }
    } else {
      // faces, :279-303
      const int k = (cls[0] != 0) ? 0 : (cls[1] != 0) ? 1 : 2;
      const float dist = (cls[k] < 0) ? -d[k] : d[k];
      const float maxD = bh[k] - (fabsf(corner[k]) + off);
      pot = powf(fmaxf(dist - off, 0.0f) / maxD, e3[k]) * p3[k];
    }
  }
  V[i].y += pot;   // :305-306
}

// ---------------------------------------------------------------------------
// Potential wall with slits, float32 (sources/cuda_kernels/potential_wall.cu:30-82).
// The in-plane 'right' vector (:39-45) is the same for every voxel and is
// computed on the host in float32.
// ---------------------------------------------------------------------------
struct WallParams {
  float dr[3], center[3], normal[3], right[3];
  float potential, plateau, slope, slope_exponent, slit_spacing, slit_width;
  int slit_count;
};

__global__ __launch_bounds__(kGenThreads) void wall_kernel(float2* __restrict__ V, Dims dm, WallParams p) {
#pragma clang fp contract(off)
  const long n = window_points(dm);
  const long i = long(blockIdx.x) * kGenThreads + threadIdx.x;
  if (i >= n) return;
  int vx[3];
  voxel_of(i, dm, vx[0], vx[1], vx[2]);
  const int NN[3] = {dm.nx, dm.ny, dm.nz};
  float e[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float r = p.dr[k] * (float(vx[k]) - 0.5f * float(NN[k]));  // :37
    e[k] = r - p.center[k];
  }
  const float d = fabsf(p.normal[0] * e[0] + p.normal[1] * e[1] + p.normal[2] * e[2]);  // :38
  const float dR = fabsf(p.right[0] * e[0] + p.right[1] * e[1] + p.right[2] * e[2]);    // :46
  const float w = p.slit_width, sp = p.slit_spacing;
  bool hole = false;  // :47-75
  switch (p.slit_count) {
    case 1: hole = dR < w * 0.5f; break;
    case 2: hole = dR > (sp - w) * 0.5f && dR < (sp + w) * 0.5f; break;
    case 3: hole = (dR < w * 0.5f) || (dR > sp - w * 0.5f && dR < sp + w * 0.5f); break;
    case 4:
      hole = (dR > (sp - w) * 0.5f && dR < (sp + w) * 0.5f) ||
             (dR > sp * 1.5f - w * 0.5f && dR < sp * 1.5f + w * 0.5f);
      break;
    default: break;
  }
  if (hole) return;
  const float pl = p.plateau * 0.5f;
  if (d <= pl) {
    V[i].x += p.potential;                                               // :76-79
  } else if (d <= pl + p.slope) {
    V[i].x += p.potential * powf(1.0f - (d - pl) / p.slope, p.slope_exponent);  // :80-82
  }
}

// harmonic: V += 1/2 m w^2 |r - c|^2 (FP64, rounded)
__global__ __launch_bounds__(kGenThreads) void harmonic_kernel(float2* __restrict__ V, Dims dm, double dx, double dy,
                                                               double dz, double k, double cx_, double cy_,
                                                               double cz_) {
  const long n = window_points(dm);
  const long i = long(blockIdx.x) * kGenThreads + threadIdx.x;
  if (i >= n) return;
  int x, y, z;
  voxel_of(i, dm, x, y, z);
  const double rx = dx * double(x - dm.nx / 2) - cx_;
  const double ry = dy * double(y - dm.ny / 2) - cy_;
  const double rz = dz * double(z - dm.nz / 2) - cz_;
  V[i].x = float(double(V[i].x) + k * (rx * rx + ry * ry + rz * rz));
}

// cos^2 lattice: V += depth * sum_d cos^2(pi r_d / period) (FP64, rounded)
__global__ __launch_bounds__(kGenThreads) void lattice_kernel(float2* __restrict__ V, Dims dm, double dx, double dy,
                                                              double dz, double depth, double period) {
  const long n = window_points(dm);
  const long i = long(blockIdx.x) * kGenThreads + threadIdx.x;
  if (i >= n) return;
  int x, y, z;
  voxel_of(i, dm, x, y, z);
  const double r[3] = {dx * double(x - dm.nx / 2), dy * double(y - dm.ny / 2), dz * double(z - dm.nz / 2)};
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double c = cos(M_PI * r[k] / period);
    acc += depth * (c * c);
  }
  V[i].x = float(double(V[i].x) + acc);
}

// exp(i theta), theta = -|k|^2 dt / (4 m), with the phase reduced modulo 2 pi in FP64
template <typename R>
__global__ __launch_bounds__(kGenThreads) void kinetic_table_kernel(cx<R>* __restrict__ out, Dims dm, double dx,
                                                                    double dy, double dz, double c) {
  const long n = window_points(dm);
  const long i = long(blockIdx.x) * kGenThreads + threadIdx.x;
  if (i >= n) return;
  int v[3];
  voxel_of(i, dm, v[0], v[1], v[2]);
  const int N[3] = {dm.nx, dm.ny, dm.nz};
  const double d[3] = {dx, dy, dz};
  double k2 = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int f = (v[a] <= (N[a] - 1) / 2) ? v[a] : v[a] - N[a];   // numpy.fft.fftfreq ordering
    const double k = 2.0 * M_PI * double(f) / (double(N[a]) * d[a]);
    k2 += k * k;
  }
  double th = fmod(-k2 * c, 2.0 * M_PI);
  double s, co;
  sincos(th, &s, &co);
  out[i] = mk<R>(R(co), R(s));
}

// cexp_i(-dt V) = (cos(-dt Vr), sin(-dt Vr)) exp(dt Vi), FP64 then rounded
template <typename R>
__device__ __forceinline__ cx<R> potential_factor(float2 v, double dt) {
  double s, c;
  sincos(-dt * double(v.x), &s, &c);
  const double a = exp(dt * double(v.y));
  return mk<R>(R(c * a), R(s * a));
}

template <typename R>
__global__ __launch_bounds__(kGenThreads) void potential_table_kernel(cx<R>* __restrict__ out,
                                                                      const float2* __restrict__ V, long n,
                                                                      double dt) {
  const long i = long(blockIdx.x) * kGenThreads + threadIdx.x;
  if (i >= n) return;
  out[i] = potential_factor<R>(V[i], dt);
}

// the same factors in the order the slab scheme's y-z launch reads them
template <typename R>
__global__ __launch_bounds__(kGenThreads) void potential_table_lines_kernel(cx<R>* __restrict__ out,
                                                                            const float2* __restrict__ V, long n,
                                                                            double dt, LinesOrder m) {
  const long o = long(blockIdx.x) * kGenThreads + threadIdx.x;
  if (o >= n) return;
  out[o] = potential_factor<R>(V[lines_source(o, m)], dt);
}

template <typename R>
__global__ __launch_bounds__(kGenThreads) void table_reorder_kernel(cx<R>* __restrict__ dst,
                                                                    const cx<R>* __restrict__ src, long n,
                                                                    LinesOrder m, int to_lines) {
  const long o = long(blockIdx.x) * kGenThreads + threadIdx.x;
  if (o >= n) return;
  const long j = lines_source(o, m);
  if (to_lines) dst[o] = src[j];
  else dst[j] = src[o];
}

}  // namespace

static bool dims_ok(const int n[3]) { return n && n[0] > 0 && n[1] > 0 && n[2] > 0; }

}  // namespace wfs

using namespace wfs;

extern "C" int wfs_init_gaussian(const int n[3], const double dx[3], int precision, const double sigma[3],
                                 const double r0[3], const double k0[3], void* psi, void* stream) {
  if (!dims_ok(n) || !dx || !sigma || !r0 || !k0 || !psi) { set_last_error("wfs_init_gaussian: bad argument"); return WFS_ERR_INVALID; }
  const Dims d = full_dims(n);
  GaussParams p;
  for (int k = 0; k < 3; ++k) { p.dx[k] = dx[k]; p.sigma[k] = sigma[k]; p.r0[k] = r0[k]; p.k0[k] = k0[k]; }
  const long total = long(n[0]) * n[1] * n[2];
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (precision == WFS_C128)
    hipLaunchKernelGGL(gaussian_kernel<double>, dim3(blocks_for(total)), dim3(kGenThreads), 0, s,
                       static_cast<double2*>(psi), d, p);
  else
    hipLaunchKernelGGL(gaussian_kernel<float>, dim3(blocks_for(total)), dim3(kGenThreads), 0, s,
                       static_cast<float2*>(psi), d, p);
  WFS_HIP_CHECK(hipGetLastError());
  return 0;
}

// The x planes [x_begin, x_end) of an n[0] x n[1] x n[2] grid: the generators below write only
// those planes' elements of V and compute every voxel there exactly as a whole-grid launch does.
static bool window_ok(const int n[3], int x_begin, int x_end) {
  return x_begin >= 0 && x_begin <= x_end && x_end <= n[0];
}

static int add_cap_window(const int n[3], const double dx[3], const double bottom[3], const double top[3],
                          double offset, const double pot_pos[3], const double pot_neg[3], const double exp_pos[3],
                          const double exp_neg[3], int x_begin, int x_end, void* V, void* stream, const char* who) {
  if (!dims_ok(n) || !dx || !bottom || !top || !pot_pos || !pot_neg || !exp_pos || !exp_neg || !V ||
      !window_ok(n, x_begin, x_end)) {
    set_last_error("%s: bad argument", who);
    return WFS_ERR_INVALID;
  }
  if (x_begin == x_end) return 0;
  const Dims d{n[0], n[1], n[2], x_begin, x_end - x_begin};
  CapParams p;
  for (int k = 0; k < 3; ++k) {
    p.dr[k] = float(dx[k]); p.bottom[k] = float(bottom[k]); p.top[k] = float(top[k]);
    p.ppos[k] = float(pot_pos[k]); p.pneg[k] = float(pot_neg[k]);
    p.epos[k] = float(exp_pos[k]); p.eneg[k] = float(exp_neg[k]);
  }
  p.offset = float(offset);
  const long total = long(d.planes) * n[1] * n[2];
  hipLaunchKernelGGL(cap_kernel, dim3(blocks_for(total)), dim3(kGenThreads), 0, static_cast<hipStream_t>(stream),
                     static_cast<float2*>(V) + long(x_begin) * n[1] * n[2], d, p);
  WFS_HIP_CHECK(hipGetLastError());
  return 0;
}

extern "C" int wfs_add_cap(const int n[3], const double dx[3], const double bottom[3], const double top[3],
                           double offset, const double pot_pos[3], const double pot_neg[3], const double exp_pos[3],
                           const double exp_neg[3], void* V, void* stream) {
  return add_cap_window(n, dx, bottom, top, offset, pot_pos, pot_neg, exp_pos, exp_neg, 0, dims_ok(n) ? n[0] : 0, V,
                        stream, "wfs_add_cap");
}

extern "C" int wfs_add_cap_planes(const int n[3], const double dx[3], const double bottom[3], const double top[3],
                                  double offset, const double pot_pos[3], const double pot_neg[3],
                                  const double exp_pos[3], const double exp_neg[3], int x_begin, int x_end, void* V,
                                  void* stream) {
  return add_cap_window(n, dx, bottom, top, offset, pot_pos, pot_neg, exp_pos, exp_neg, x_begin, x_end, V, stream,
                        "wfs_add_cap_planes");
}

static int add_wall_window(const int n[3], const double dx[3], double potential, const double center[3],
                           const double normal[3], double plateau, double slope, double slope_exponent,
                           int slit_count, double slit_spacing, double slit_width, double slit_rotation, int x_begin,
                           int x_end, void* V, void* stream, const char* who) {
#pragma clang fp contract(off)
  if (!dims_ok(n) || !dx || !center || !normal || !V || slit_count < 0 || slit_count > 4 ||
      !window_ok(n, x_begin, x_end)) {
    set_last_error("%s: bad argument", who);
    return WFS_ERR_INVALID;
  }
  if (x_begin == x_end) return 0;
  const Dims d{n[0], n[1], n[2], x_begin, x_end - x_begin};
  WallParams p;
  for (int k = 0; k < 3; ++k) { p.dr[k] = float(dx[k]); p.center[k] = float(center[k]); p.normal[k] = float(normal[k]); }
  // right = rotate(normalize(normal x up), normal, rotation) in float32 (potential_wall.cu:39-45,
  // common.cu:348-369)
  float up[3] = {0.0f, 1.0f, 0.0f};
  const float* nn = p.normal;
  if (std::fabs(nn[0] * up[0] + nn[1] * up[1] + nn[2] * up[2]) > 0.99f) { up[0] = 0.0f; up[2] = 1.0f; }
  float cr[3] = {nn[1] * up[2] - nn[2] * up[1], nn[2] * up[0] - nn[0] * up[2], nn[0] * up[1] - nn[1] * up[0]};
  const float len = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
  float rt[3] = {cr[0] / len, cr[1] / len, cr[2] / len};
  const float h = float(slit_rotation) / 2.0f;
  const float q0 = std::cos(h), sh = std::sin(h);
  const float q1 = sh * nn[0], q2 = sh * nn[1], q3 = sh * nn[2];
  const float Q[3][3] = {
      {q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3, 2.0f * (q1 * q2 - q0 * q3), 2.0f * (q1 * q3 + q0 * q2)},
      {2.0f * (q2 * q1 + q0 * q3), q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3, 2.0f * (q2 * q3 - q0 * q1)},
      {2.0f * (q3 * q1 - q0 * q2), 2.0f * (q3 * q2 + q0 * q1), q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3}};
  for (int k = 0; k < 3; ++k) p.right[k] = Q[k][0] * rt[0] + Q[k][1] * rt[1] + Q[k][2] * rt[2];
  p.potential = float(potential);
  p.plateau = float(plateau);
  p.slope = float(slope);
  p.slope_exponent = float(slope_exponent);
  p.slit_spacing = float(slit_spacing);
  p.slit_width = float(slit_width);
  p.slit_count = slit_count;
  const long total = long(d.planes) * n[1] * n[2];
  hipLaunchKernelGGL(wall_kernel, dim3(blocks_for(total)), dim3(kGenThreads), 0, static_cast<hipStream_t>(stream),
                     static_cast<float2*>(V) + long(x_begin) * n[1] * n[2], d, p);
  WFS_HIP_CHECK(hipGetLastError());
  return 0;
}

extern "C" int wfs_add_wall(const int n[3], const double dx[3], double potential, const double center[3],
                            const double normal[3], double plateau, double slope, double slope_exponent,
                            int slit_count, double slit_spacing, double slit_width, double slit_rotation, void* V,
                            void* stream) {
  return add_wall_window(n, dx, potential, center, normal, plateau, slope, slope_exponent, slit_count, slit_spacing,
                         slit_width, slit_rotation, 0, dims_ok(n) ? n[0] : 0, V, stream, "wfs_add_wall");
}

extern "C" int wfs_add_wall_planes(const int n[3], const double dx[3], double potential, const double center[3],
                                   const double normal[3], double plateau, double slope, double slope_exponent,
                                   int slit_count, double slit_spacing, double slit_width, double slit_rotation,
                                   int x_begin, int x_end, void* V, void* stream) {
  return add_wall_window(n, dx, potential, center, normal, plateau, slope, slope_exponent, slit_count, slit_spacing,
                         slit_width, slit_rotation, x_begin, x_end, V, stream, "wfs_add_wall_planes");
}

extern "C" int wfs_add_harmonic(const int n[3], const double dx[3], double omega, double mass,
                                const double center[3], void* V, void* stream) {
  if (!dims_ok(n) || !dx || !center || !V) { set_last_error("wfs_add_harmonic: bad argument"); return WFS_ERR_INVALID; }
  const Dims d = full_dims(n);
  const long total = long(n[0]) * n[1] * n[2];
  hipLaunchKernelGGL(harmonic_kernel, dim3(blocks_for(total)), dim3(kGenThreads), 0, static_cast<hipStream_t>(stream),
                     static_cast<float2*>(V), d, dx[0], dx[1], dx[2], 0.5 * mass * omega * omega, center[0],
                     center[1], center[2]);
  WFS_HIP_CHECK(hipGetLastError());
  return 0;
}

extern "C" int wfs_add_lattice(const int n[3], const double dx[3], double depth, double period, void* V,
                               void* stream) {
  if (!dims_ok(n) || !dx || !V || period == 0.0) { set_last_error("wfs_add_lattice: bad argument"); return WFS_ERR_INVALID; }
  const Dims d = full_dims(n);
  const long total = long(n[0]) * n[1] * n[2];
  hipLaunchKernelGGL(lattice_kernel, dim3(blocks_for(total)), dim3(kGenThreads), 0, static_cast<hipStream_t>(stream),
                     static_cast<float2*>(V), d, dx[0], dx[1], dx[2], depth, period);
  WFS_HIP_CHECK(hipGetLastError());
  return 0;
}

extern "C" int wfs_kinetic_operator(const int n[3], const double dx[3], double dt, double mass, int precision,
                                    void* out, void* stream) {
  if (!dims_ok(n) || !dx || !out || !(mass > 0.0)) { set_last_error("wfs_kinetic_operator: bad argument"); return WFS_ERR_INVALID; }
  const Dims d = full_dims(n);
  const long total = long(n[0]) * n[1] * n[2];
  const double c = dt / (4.0 * mass);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (precision == WFS_C128)
    hipLaunchKernelGGL(kinetic_table_kernel<double>, dim3(blocks_for(total)), dim3(kGenThreads), 0, s,
                       static_cast<double2*>(out), d, dx[0], dx[1], dx[2], c);
  else
    hipLaunchKernelGGL(kinetic_table_kernel<float>, dim3(blocks_for(total)), dim3(kGenThreads), 0, s,
                       static_cast<float2*>(out), d, dx[0], dx[1], dx[2], c);
  WFS_HIP_CHECK(hipGetLastError());
  return 0;
}

extern "C" int wfs_potential_operator(long count, const void* V, double dt, int precision, void* out,
                                      void* stream) {
  if (count < 0 || (count > 0 && (!V || !out))) { set_last_error("wfs_potential_operator: bad argument"); return WFS_ERR_INVALID; }
  if (count == 0) return 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (precision == WFS_C128)
    hipLaunchKernelGGL(potential_table_kernel<double>, dim3(blocks_for(count)), dim3(kGenThreads), 0, s,
                       static_cast<double2*>(out), static_cast<const float2*>(V), count, dt);
  else
    hipLaunchKernelGGL(potential_table_kernel<float>, dim3(blocks_for(count)), dim3(kGenThreads), 0, s,
                       static_cast<float2*>(out), static_cast<const float2*>(V), count, dt);
  WFS_HIP_CHECK(hipGetLastError());
  return 0;
}

namespace wfs {
int launch_potential_phase(int precision, long count, const float2* V, double dt, void* pv, hipStream_t s) {
  if (count <= 0) return 0;
  if (precision == WFS_C128)
    hipLaunchKernelGGL(potential_table_kernel<double>, dim3(blocks_for(count)), dim3(kGenThreads), 0, s,
                       static_cast<double2*>(pv), V, count, dt);
  else
    hipLaunchKernelGGL(potential_table_kernel<float>, dim3(blocks_for(count)), dim3(kGenThreads), 0, s,
                       static_cast<float2*>(pv), V, count, dt);
  return int(hipGetLastError());
}

static bool lines_order_ok(long count, const LinesOrder& m) {
  return m.ly > 0 && m.nz > 0 && m.b > 0 && m.rh > 0 && m.n1 > 0 && m.threads > 0 &&
         count % (long(m.ly) * m.nz) == 0 && long(m.ly) * m.b == m.ny;
}

int launch_potential_phase_lines(int precision, long count, const float2* V, double dt, void* pv,
                                 const LinesOrder& m, hipStream_t s) {
  if (count <= 0) return 0;
  if (!lines_order_ok(count, m)) { set_last_error("potential table: bad lines order"); return WFS_ERR_INVALID; }
  if (precision == WFS_C128)
    hipLaunchKernelGGL(potential_table_lines_kernel<double>, dim3(blocks_for(count)), dim3(kGenThreads), 0, s,
                       static_cast<double2*>(pv), V, count, dt, m);
  else
    hipLaunchKernelGGL(potential_table_lines_kernel<float>, dim3(blocks_for(count)), dim3(kGenThreads), 0, s,
                       static_cast<float2*>(pv), V, count, dt, m);
  return int(hipGetLastError());
}

int launch_table_reorder(int precision, long count, const void* src, void* dst, const LinesOrder& m, bool to_lines,
                         hipStream_t s) {
  if (count <= 0) return 0;
  if (!lines_order_ok(count, m)) { set_last_error("potential table: bad lines order"); return WFS_ERR_INVALID; }
  if (precision == WFS_C128)
    hipLaunchKernelGGL(table_reorder_kernel<double>, dim3(blocks_for(count)), dim3(kGenThreads), 0, s,
                       static_cast<double2*>(dst), static_cast<const double2*>(src), count, m, int(to_lines));
  else
    hipLaunchKernelGGL(table_reorder_kernel<float>, dim3(blocks_for(count)), dim3(kGenThreads), 0, s,
                       static_cast<float2*>(dst), static_cast<const float2*>(src), count, m, int(to_lines));
  return int(hipGetLastError());
}
}  // namespace wfs
