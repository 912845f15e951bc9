// Simpson-weighted moments of |psi|^2 over a box (gfx950).
//
// For each point of the box (local index i, sample counts n) the weight is
//   w = s(ix) s(iy) s(iz) / 27,  s = 1 at both ends, 4 at odd and 2 at even
// interior positions (sources/cuda_kernels/common.cu:466-492), and the position
// is r_d = dx_d (bottom_d + i_d - N_d / 2) with integer N_d/2
// (sources/cuda_kernels/expected_location.cu:33-37).  One pass accumulates in
// FP64
//   P = sum w|psi|^2 dV,  E[r_d] = sum w r_d |psi|^2 dV,  E[r_d^2] = sum w r_d^2 |psi|^2 dV
// which the reference computes in two kernels with float atomics
// (volume_probability.cu:29-41, expected_location.cu:44-66).
//
// Up to four boxes are measured in one pass over their bounding box (each point
// read once, added to every box that holds it).  Each block owns 32 (x, y) rows,
// a wavefront one row at a time (lanes along z, coalesced); the lanes reduce with
// cross-lane shuffles at the end, waves combine through LDS, and the block
// writes one partial record per box.  A second kernel (one block per packet) sums the partials in
// a fixed order, so results are bit-reproducible run to run.  A stack of
// independent grids of the same shape is handled by one launch
// (blockIdx.y = packet).
#include "wfs_internal.h"

#include <algorithm>
#include <cmath>
#include <mutex>

namespace wfs {

constexpr int kObsThreads = 256;
constexpr int kObsQ = 7;
constexpr int kRowsPerBlock = 32;   // rows (x, y) per block; wave w takes rows w, w + 4, ...

__device__ __forceinline__ double simpson_s(int i, int n) {
  if (i == 0 || i == n - 1) return 1.0;
  return (i & 1) ? 4.0 : 2.0;
}

__device__ __forceinline__ double wave_sum(double v) {
  v += xor_lane<32>(v);
  v += xor_lane<16>(v);
  v += xor_lane<8>(v);
  v += xor_lane<4>(v);
  v += xor_lane<2>(v);
  v += xor_lane<1>(v);
  return v;
}

// Up to kBoxesPerPass boxes measured in one pass over their bounding box: every point is
// read once and adds to each box that holds it, with that box's Simpson weights.
struct BoxSet {
  int n;
  int lo[3], cnt[3];                      // bounding box of the set
  int bot[kBoxesPerPass][3], num[kBoxesPerPass][3];
};

template <typename R>
__global__ __launch_bounds__(kObsThreads) void box_moments_kernel(const cx<R>* __restrict__ psi, int nx, int ny,
                                                                  int nz, double dx, double dy, double dz,
                                                                  BoxSet set, double* __restrict__ partials) {
  constexpr int kWaves = kObsThreads / 64;
  constexpr int NB = kBoxesPerPass;
  const long nrows = long(set.cnt[0]) * set.cnt[1];
  // packet blockIdx.y of a stack of independent grids
  psi += long(blockIdx.y) * nx * ny * nz;
  partials += long(blockIdx.y) * gridDim.x * set.n * kObsQ;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const long row0 = long(blockIdx.x) * kRowsPerBlock;
  // lane partial sums; every term is linear in the lane's z sums, so the row weights
  // (uniform over the wavefront) multiply them before the one reduction at the end
  double acc[NB][kObsQ];
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int q = 0; q < kObsQ; ++q) acc[b][q] = 0.0;

  for (int rr = wave; rr < kRowsPerBlock; rr += kWaves) {
    const long row = row0 + rr;
    if (row >= nrows) break;
    const int gx = set.lo[0] + int(row / set.cnt[1]);
    const int gy = set.lo[1] + int(row % set.cnt[1]);
    double wxy[NB];
    bool any = false;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int ix = gx - set.bot[b][0], iy = gy - set.bot[b][1];
      const bool in = b < set.n && ix >= 0 && ix < set.num[b][0] && iy >= 0 && iy < set.num[b][1];
      wxy[b] = in ? simpson_s(ix, set.num[b][0]) * simpson_s(iy, set.num[b][1]) : 0.0;
      any = any || in;
    }
    if (!any) continue;
    const double rx = dx * double(gx - nx / 2);
    const double ry = dy * double(gy - ny / 2);
    const cx<R>* rowp = psi + (long(gx) * ny + gy) * long(nz);
    double s0[NB], sz1[NB], sz2[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) s0[b] = sz1[b] = sz2[b] = 0.0;
    // 256 points of the row at a time: four loads issued together (indices clamped into the
    // row, the excess lanes masked afterwards), so that each wavefront has four in flight
    const int zend = set.lo[2] + set.cnt[2];
    for (int z0 = set.lo[2]; z0 < zend; z0 += 256) {
      cx<R> v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = rowp[min(z0 + lane + 64 * u, zend - 1)];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int gz = z0 + lane + 64 * u;
        const double rho = gz < zend ? double(v[u].x) * double(v[u].x) + double(v[u].y) * double(v[u].y) : 0.0;
        const double rz = dz * double(gz - nz / 2);
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          const int iz = gz - set.bot[b][2];
          const bool in = wxy[b] != 0.0 && iz >= 0 && iz < set.num[b][2];
          const double wr = in ? simpson_s(iz, set.num[b][2]) * rho : 0.0;
          s0[b] += wr;
          sz1[b] += wr * rz;
          sz2[b] += wr * rz * rz;
        }
      }
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const double w = wxy[b];
      acc[b][0] += w * s0[b];
      acc[b][1] += w * rx * s0[b];
      acc[b][2] += w * ry * s0[b];
      acc[b][3] += w * sz1[b];
      acc[b][4] += w * rx * rx * s0[b];
      acc[b][5] += w * ry * ry * s0[b];
      acc[b][6] += w * sz2[b];
    }
  }

  __shared__ double red[kWaves][NB * kObsQ];
  static_assert(kObsQ == 7, "the wavefront sums take seven values");
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    if (b >= set.n) break;
    const double v = wave_sum7(acc[b], lane);
    const int q = (lane >> 3) & 7;
    if ((lane & 7) == 0 && q < kObsQ) red[wave][b * kObsQ + q] = v;
  }
  __syncthreads();
  if (int(threadIdx.x) < set.n * kObsQ) {
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) v += red[w][threadIdx.x];
    partials[long(blockIdx.x) * set.n * kObsQ + threadIdx.x] = v;
  }
}

// one block per (box, packet): the box's records summed in a fixed order
__global__ __launch_bounds__(kObsThreads) void sum_partials_kernel(const double* __restrict__ partials, long nparts,
                                                                   int nb, double scale, double* __restrict__ out,
                                                                   long out_stride) {
  const int box = blockIdx.x;
  partials += long(blockIdx.y) * nparts * nb * kObsQ + long(box) * kObsQ;
  out += long(blockIdx.y) * out_stride + long(box) * kObsQ;
  double acc[kObsQ];
#pragma unroll
  for (int q = 0; q < kObsQ; ++q) acc[q] = 0.0;
  for (long r = threadIdx.x; r < nparts; r += kObsThreads) {
#pragma unroll
    for (int q = 0; q < kObsQ; ++q) acc[q] += partials[r * nb * kObsQ + q];
  }
  __shared__ double red[kObsThreads / 64][kObsQ];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < kObsQ; ++q) {
    const double v = wave_sum(acc[q]);
    if (lane == 0) red[wave][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < kObsQ) {
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < kObsThreads / 64; ++w) v += red[w][threadIdx.x];
    out[threadIdx.x] = v * scale;
  }
}

int launch_sum_partials(const double* partials, long nparts, int nb, int batch, double scale, double* out,
                        long out_stride, hipStream_t s) {
  WFS_LAUNCH(sum_partials_kernel, dim3(unsigned(nb), unsigned(batch)), dim3(kObsThreads), 0, s, partials,
                     nparts, nb, scale, out, out_stride);
  return int(hipGetLastError());
}

long box_moments_partials_needed(const wfs_box& box) {
  const long nrows = long(box.count[0]) * box.count[1];
  return ((nrows + kRowsPerBlock - 1) / kRowsPerBlock) * kObsQ * kBoxesPerPass;
}

// out[p * out_stride + 7 b + q], q < 7, for the nb <= kBoxesPerPass boxes and every packet p
int launch_box_moments(int precision, int batch, int nx, int ny, int nz, const double dx[3], const void* psi,
                       const wfs_box* boxes, int nb, double* partials, long max_partials, double* out,
                       long out_stride, hipStream_t s) {
  if (nb < 1 || nb > kBoxesPerPass) {
    set_last_error("box moments: %d boxes in one pass (1..%d)", nb, kBoxesPerPass);
    return WFS_ERR_INVALID;
  }
  BoxSet set{};
  set.n = nb;
  for (int a = 0; a < 3; ++a) {
    int lo = boxes[0].bottom[a], hi = boxes[0].bottom[a] + boxes[0].count[a];
    for (int b = 1; b < nb; ++b) {
      lo = std::min(lo, boxes[b].bottom[a]);
      hi = std::max(hi, boxes[b].bottom[a] + boxes[b].count[a]);
    }
    set.lo[a] = lo;
    set.cnt[a] = hi - lo;
    for (int b = 0; b < nb; ++b) {
      set.bot[b][a] = boxes[b].bottom[a];
      set.num[b][a] = boxes[b].count[a];
    }
  }
  const long nrows = long(set.cnt[0]) * set.cnt[1];
  const long nblocks = (nrows + kRowsPerBlock - 1) / kRowsPerBlock;
  if (nblocks * kObsQ * nb * batch > max_partials) {
    set_last_error("observable workspace too small");
    return WFS_ERR_INVALID;
  }
  // the Simpson weight products carry 1/27 and the volume element dx dy dz
  const double sc = dx[0] * dx[1] * dx[2] / 27.0;
  if (precision == WFS_C128) {
    WFS_LAUNCH(box_moments_kernel<double>, dim3(unsigned(nblocks), unsigned(batch)), dim3(kObsThreads), 0, s,
                       static_cast<const double2*>(psi), nx, ny, nz, dx[0], dx[1], dx[2], set, partials);
  } else {
    WFS_LAUNCH(box_moments_kernel<float>, dim3(unsigned(nblocks), unsigned(batch)), dim3(kObsThreads), 0, s,
                       static_cast<const float2*>(psi), nx, ny, nz, dx[0], dx[1], dx[2], set, partials);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return int(e);
  WFS_LAUNCH(sum_partials_kernel, dim3(unsigned(nb), unsigned(batch)), dim3(kObsThreads), 0, s, partials,
                     nblocks, nb, sc, out, out_stride);
  return int(hipGetLastError());
}

}  // namespace wfs

// ---------------------------------------------------------------------------
// Probability current through a tilted plane
// (sources/cuda_kernels/probability_current_density.cu:33-218).
//
// Pixel (i, j) of a count_u x count_v grid on the plane sits at
//   r = c + right * dW * (i - count_u/2) + up * dH * (j - count_v/2)
// (float32, :41-58), falls into voxel floor(r/dx + N/2) (:61-62), and gets
//   j.n = Im(conj(psi) * n.grad psi) / m
// with the 6th-order central difference of :174-176 whose k-step neighbour
// falls back to the centre voxel when it leaves the grid (:68-144).  The flux
// is the 2-D Simpson integral of j.n dW dH (:198-214), accumulated in FP64.
// The density map is written as [count_u][count_v]; pixels whose voxel lies
// outside the grid get zero density (the reference reads out of bounds there).
// ---------------------------------------------------------------------------
namespace wfs {

struct PlaneParams {
  float center[3], right[3], up[3];
  float dW, dH;            // float copies used for the pixel positions (:56-57)
  double dWd, dHd;         // state-precision widths used for the weights (:51-52, :202)
  double normal[3];
  int cu, cv;
  double inv_mass;
};

template <typename R>
__global__ __launch_bounds__(kObsThreads) void plane_current_kernel(const cx<R>* __restrict__ psi, int nx, int ny,
                                                                    int nz, double dx, double dy, double dz,
                                                                    PlaneParams p, R* __restrict__ density,
                                                                    double* __restrict__ partials) {
#pragma clang fp contract(off)
  const long npix = long(p.cu) * p.cv;
  const long pix = long(blockIdx.x) * kObsThreads + threadIdx.x;
  double contrib = 0.0;
  if (pix < npix) {
    const int i = int(pix / p.cv);
    const int j = int(pix % p.cv);
    const float si = float(i - p.cu / 2), sj = float(j - p.cv / 2);
    float fr[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) fr[a] = (p.center[a] + p.right[a] * p.dW * si) + p.up[a] * p.dH * sj;
    const int N[3] = {nx, ny, nz};
    const double d[3] = {dx, dy, dz};
    int v[3];
    bool inside = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const R fv = R(fr[a]) / R(d[a]) + R(0.5f) * R(N[a]);
      inside = inside && fv >= R(0) && fv < R(N[a]);
      v[a] = inside ? int(fv) : 0;
    }
    R jn = R(0);
    if (inside) {
      const long sx = long(ny) * nz, sy = nz;
      const long c = long(v[0]) * sx + long(v[1]) * sy + v[2];
      const long strides[3] = {sx, sy, 1};
      const cx<R> p0 = psi[c];
      cx<R> grad[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        cx<R> f[7];   // offsets -3..3
#pragma unroll
        for (int k = 1; k <= 3; ++k) {
          f[3 - k] = (v[a] >= k) ? psi[c - k * strides[a]] : p0;
          f[3 + k] = (v[a] < N[a] - k) ? psi[c + k * strides[a]] : p0;
        }
        cx<R> g;
        g.x = (f[6].x - R(9) * f[5].x + R(45) * f[4].x - R(45) * f[2].x + R(9) * f[1].x - f[0].x) / R(60) / R(d[a]);
        g.y = (f[6].y - R(9) * f[5].y + R(45) * f[4].y - R(45) * f[2].y + R(9) * f[1].y - f[0].y) / R(60) / R(d[a]);
        grad[a] = g;
      }
      cx<R> dpsi;
      dpsi.x = R(p.normal[0]) * grad[0].x + R(p.normal[1]) * grad[1].x + R(p.normal[2]) * grad[2].x;
      dpsi.y = R(p.normal[0]) * grad[0].y + R(p.normal[1]) * grad[1].y + R(p.normal[2]) * grad[2].y;
      // (-1/2m) i (conj(psi) dpsi - psi conj(dpsi)) = Im(conj(psi) dpsi) / m
      jn = R(p.inv_mass) * (p0.x * dpsi.y - p0.y * dpsi.x);
    }
    if (density) density[pix] = jn;
    const double su = (i == 0 || i == p.cu - 1) ? 1.0 : ((i & 1) ? 4.0 : 2.0);
    const double sv = (j == 0 || j == p.cv - 1) ? 1.0 : ((j & 1) ? 4.0 : 2.0);
    contrib = double(jn) * su * sv;
  }
  __shared__ double red[kObsThreads / 64];
  const double w = wave_sum(contrib);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = w;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < kObsThreads / 64; ++k) t += red[k];
    partials[long(blockIdx.x) * kObsQ] = t;
#pragma unroll
    for (int q = 1; q < kObsQ; ++q) partials[long(blockIdx.x) * kObsQ + q] = 0.0;
  }
}

long plane_current_partials_needed(int cu, int cv) {
  return ((long(cu) * cv + kObsThreads - 1) / kObsThreads) * kObsQ;
}

int launch_plane_current(int precision, int nx, int ny, int nz, const double dx[3], const void* psi,
                         const double center[3], const double normal[3], const double size[2], const int count[2],
                         double mass, void* density, double* partials, long max_partials, double* out7,
                         hipStream_t s) {
#pragma clang fp contract(off)
  const long nblocks = (long(count[0]) * count[1] + kObsThreads - 1) / kObsThreads;
  if (nblocks * kObsQ > max_partials) { set_last_error("plane current workspace too small"); return WFS_ERR_INVALID; }
  PlaneParams p;
  // right/up in float32 as the reference (:41-49): up = (0,1,0), or (0,0,1) when |n.up| > 0.99
  float n[3] = {float(normal[0]), float(normal[1]), float(normal[2])};
  float pu[3] = {0.0f, 1.0f, 0.0f};
  if (std::fabs(n[0] * pu[0] + n[1] * pu[1] + n[2] * pu[2]) > 0.99) { pu[1] = 0.0f; pu[2] = 1.0f; }
  float r[3] = {n[1] * pu[2] - n[2] * pu[1], n[2] * pu[0] - n[0] * pu[2], n[0] * pu[1] - n[1] * pu[0]};
  float rl = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  for (int a = 0; a < 3; ++a) r[a] = r[a] / rl;
  float u[3] = {r[1] * n[2] - r[2] * n[1], r[2] * n[0] - r[0] * n[2], r[0] * n[1] - r[1] * n[0]};
  float ul = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  for (int a = 0; a < 3; ++a) {
    u[a] = u[a] / ul;
    p.center[a] = float(center[a]);
    p.right[a] = r[a];
    p.up[a] = u[a];
    p.normal[a] = double(n[a]);
  }
  if (precision == WFS_C128) {
    p.dWd = size[0] / double(count[0] - 1);
    p.dHd = size[1] / double(count[1] - 1);
  } else {
    p.dWd = double(float(size[0]) / float(count[0] - 1));
    p.dHd = double(float(size[1]) / float(count[1] - 1));
  }
  p.dW = float(p.dWd);
  p.dH = float(p.dHd);
  p.cu = count[0];
  p.cv = count[1];
  p.inv_mass = 1.0 / mass;
  if (precision == WFS_C128)
    WFS_LAUNCH(plane_current_kernel<double>, dim3(unsigned(nblocks)), dim3(kObsThreads), 0, s,
                       static_cast<const double2*>(psi), nx, ny, nz, dx[0], dx[1], dx[2], p,
                       static_cast<double*>(density), partials);
  else
    WFS_LAUNCH(plane_current_kernel<float>, dim3(unsigned(nblocks)), dim3(kObsThreads), 0, s,
                       static_cast<const float2*>(psi), nx, ny, nz, dx[0], dx[1], dx[2], p,
                       static_cast<float*>(density), partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return int(e);
  // Simpson 1/9 and the pixel area dW dH
  WFS_LAUNCH(sum_partials_kernel, dim3(1), dim3(kObsThreads), 0, s, partials, nblocks, 1,
                     p.dWd * p.dHd / 9.0, out7, 7L);
  return int(hipGetLastError());
}

}  // namespace wfs

// ---------------------------------------------------------------------------
// Box moments without a plan and without any allocation (wfs_observables_noalloc)
// ---------------------------------------------------------------------------
// The partial records live in this file-scope device array: each GPU the library runs on holds
// its own copy until the process ends.  Host code never hands the array's name to a kernel (in
// host code it is only a stand-in, not a device address); it passes the address that
// hipGetSymbolAddress returns on the current device, looked up once per device.
// Every record the summing kernel reads is written by the moments kernel of the same call, so
// nothing in the array has to start at zero.  Calls that overlap on one GPU share the array and
// race.
namespace wfs {
namespace {

constexpr long kScratchDoubles = 262144;   // 2 MiB of partial records
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

int scratch_address(double** out) {
  int dev = 0;
  WFS_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_last_error("wfs_observables_noalloc: device %d beyond the %d the library tracks", dev, kMaxDevices);
    return WFS_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_last_error("wfs_observables_noalloc: hipGetSymbolAddress of the static workspace failed: %s",
                     hipGetErrorString(e));
      return e != hipSuccess ? int(e) : WFS_ERR_INVALID;
    }
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace
}  // namespace wfs

using namespace wfs;

extern "C" int wfs_observables_noalloc(const int n[3], const double dx[3], int precision, int batch, const void* psi,
                                       const wfs_box* boxes, int nboxes, double* out, void* stream) {
  if (!n || !dx || !psi || nboxes < 0 || batch < 1 || (nboxes > 0 && (!boxes || !out)) ||
      (precision != WFS_C64 && precision != WFS_C128)) {
    set_last_error("wfs_observables_noalloc: bad argument");
    return WFS_ERR_INVALID;
  }
  for (int a = 0; a < 3; ++a)
    if (n[a] < 1 || !(dx[a] > 0.0)) { set_last_error("wfs_observables_noalloc: bad grid"); return WFS_ERR_INVALID; }
  for (int b = 0; b < nboxes; ++b)
    for (int a = 0; a < 3; ++a)
      if (boxes[b].bottom[a] < 0 || boxes[b].count[a] < 1 || long(boxes[b].bottom[a]) + boxes[b].count[a] > n[a]) {
        set_last_error("wfs_observables_noalloc: box %d axis %d [%d, %d) outside the grid of %d voxels", b, a,
                       boxes[b].bottom[a], boxes[b].bottom[a] + boxes[b].count[a], n[a]);
        return WFS_ERR_INVALID;
      }
  // every pass of up to kBoxesPerPass boxes must fit its records into the static workspace
  for (int b0 = 0; b0 < nboxes; b0 += kBoxesPerPass) {
    const int nb = std::min(kBoxesPerPass, nboxes - b0);
    long lo[2], hi[2];
    for (int a = 0; a < 2; ++a) {
      lo[a] = boxes[b0].bottom[a];
      hi[a] = long(boxes[b0].bottom[a]) + boxes[b0].count[a];
      for (int b = b0 + 1; b < b0 + nb; ++b) {
        lo[a] = std::min(lo[a], long(boxes[b].bottom[a]));
        hi[a] = std::max(hi[a], long(boxes[b].bottom[a]) + boxes[b].count[a]);
      }
    }
    const long rows = (hi[0] - lo[0]) * (hi[1] - lo[1]);
    const long need = (rows + kRowsPerBlock - 1) / kRowsPerBlock * kObsQ * nb * long(batch);
    if (need > kScratchDoubles) {
      set_last_error("wfs_observables_noalloc: boxes %d..%d of %d packets need %ld partial records; the static "
                     "workspace holds %ld (2 MiB): use wfs_observables with a plan", b0, b0 + nb - 1, batch, need,
                     kScratchDoubles);
      return WFS_ERR_UNSUPPORTED;
    }
  }
  if (nboxes == 0) return 0;
  double* work = nullptr;
  const int rc = scratch_address(&work);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  for (int b0 = 0; b0 < nboxes; b0 += kBoxesPerPass) {
    const int nb = std::min(kBoxesPerPass, nboxes - b0);
    const int r = launch_box_moments(precision, batch, n[0], n[1], n[2], dx, psi, boxes + b0, nb, work,
                                     kScratchDoubles, out + 7 * b0, 7L * nboxes, s);
    if (r) return r;
  }
  return 0;
}
