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
// Each block owns a run of (x, y) rows of the box; a wavefront reduces with
// cross-lane shuffles, waves combine through LDS, and the block writes one
// partial record.  A second kernel (one block per packet) sums the partials in
// a fixed order, so results are bit-reproducible run to run.  A stack of
// independent grids of the same shape is handled by one launch
// (blockIdx.y = packet).
#include "wfs_internal.h"

#include <cmath>

namespace wfs {

constexpr int kObsThreads = 256;
constexpr int kObsQ = 7;
constexpr int kRowsPerBlock = 4;

__device__ __forceinline__ double simpson_s(int i, int n) {
  if (i == 0 || i == n - 1) return 1.0;
  return (i & 1) ? 4.0 : 2.0;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <typename R>
__global__ __launch_bounds__(kObsThreads) void box_moments_kernel(const cx<R>* __restrict__ psi, int nx, int ny,
                                                                  int nz, double dx, double dy, double dz,
                                                                  wfs_box box, double* __restrict__ partials) {
  const long nrows = long(box.count[0]) * box.count[1];
  // packet blockIdx.y of a stack of independent grids
  psi += long(blockIdx.y) * nx * ny * nz;
  partials += long(blockIdx.y) * gridDim.x * kObsQ;
  const long row0 = long(blockIdx.x) * kRowsPerBlock;
  double acc[kObsQ];
#pragma unroll
  for (int q = 0; q < kObsQ; ++q) acc[q] = 0.0;

  for (int rr = 0; rr < kRowsPerBlock; ++rr) {
    const long row = row0 + rr;
    if (row >= nrows) break;
    const int ix = int(row / box.count[1]);
    const int iy = int(row % box.count[1]);
    const double wxy = simpson_s(ix, box.count[0]) * simpson_s(iy, box.count[1]);
    const double rx = dx * double(box.bottom[0] + ix - nx / 2);
    const double ry = dy * double(box.bottom[1] + iy - ny / 2);
    const cx<R>* rowp = psi + (long(box.bottom[0] + ix) * ny + (box.bottom[1] + iy)) * long(nz) + box.bottom[2];
    double s0 = 0.0, sz1 = 0.0, sz2 = 0.0;
    for (int iz = threadIdx.x; iz < box.count[2]; iz += kObsThreads) {
      const cx<R> v = rowp[iz];
      const double rho = double(v.x) * double(v.x) + double(v.y) * double(v.y);
      const double wr = simpson_s(iz, box.count[2]) * rho;
      const double rz = dz * double(box.bottom[2] + iz - nz / 2);
      s0 += wr;
      sz1 += wr * rz;
      sz2 += wr * rz * rz;
    }
    const double w = wxy;
    acc[0] += w * s0;
    acc[1] += w * rx * s0;
    acc[2] += w * ry * s0;
    acc[3] += w * sz1;
    acc[4] += w * rx * rx * s0;
    acc[5] += w * ry * ry * s0;
    acc[6] += w * sz2;
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
    partials[long(blockIdx.x) * kObsQ + threadIdx.x] = v;
  }
}

__global__ __launch_bounds__(kObsThreads) void sum_partials_kernel(const double* __restrict__ partials, long nparts,
                                                                   double scale, double* __restrict__ out,
                                                                   long out_stride) {
  // one block per packet
  partials += long(blockIdx.x) * nparts * kObsQ;
  out += long(blockIdx.x) * out_stride;
  double acc[kObsQ];
#pragma unroll
  for (int q = 0; q < kObsQ; ++q) acc[q] = 0.0;
  for (long b = threadIdx.x; b < nparts; b += kObsThreads) {
#pragma unroll
    for (int q = 0; q < kObsQ; ++q) acc[q] += partials[b * kObsQ + q];
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

long box_moments_partials_needed(const wfs_box& box) {
  const long nrows = long(box.count[0]) * box.count[1];
  return ((nrows + kRowsPerBlock - 1) / kRowsPerBlock) * kObsQ;
}

// out[b * out_stride + q], q < 7, for every packet b < batch
int launch_box_moments(int precision, int batch, int nx, int ny, int nz, const double dx[3], const void* psi,
                       const wfs_box& box, double* partials, long max_partials, double* out, long out_stride,
                       hipStream_t s) {
  const long nrows = long(box.count[0]) * box.count[1];
  const long nblocks = (nrows + kRowsPerBlock - 1) / kRowsPerBlock;
  if (nblocks * kObsQ * batch > max_partials) {
    set_last_error("observable workspace too small");
    return WFS_ERR_INVALID;
  }
  // the Simpson weight products carry 1/27 and the volume element dx dy dz
  const double scale = dx[0] * dx[1] * dx[2] / 27.0;
  if (precision == WFS_C128) {
    hipLaunchKernelGGL(box_moments_kernel<double>, dim3(unsigned(nblocks), unsigned(batch)), dim3(kObsThreads), 0, s,
                       static_cast<const double2*>(psi), nx, ny, nz, dx[0], dx[1], dx[2], box, partials);
  } else {
    hipLaunchKernelGGL(box_moments_kernel<float>, dim3(unsigned(nblocks), unsigned(batch)), dim3(kObsThreads), 0, s,
                       static_cast<const float2*>(psi), nx, ny, nz, dx[0], dx[1], dx[2], box, partials);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return int(e);
  hipLaunchKernelGGL(sum_partials_kernel, dim3(unsigned(batch)), dim3(kObsThreads), 0, s, partials, nblocks, scale, out,
                     out_stride);
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
    hipLaunchKernelGGL(plane_current_kernel<double>, dim3(unsigned(nblocks)), dim3(kObsThreads), 0, s,
                       static_cast<const double2*>(psi), nx, ny, nz, dx[0], dx[1], dx[2], p,
                       static_cast<double*>(density), partials);
  else
    hipLaunchKernelGGL(plane_current_kernel<float>, dim3(unsigned(nblocks)), dim3(kObsThreads), 0, s,
                       static_cast<const float2*>(psi), nx, ny, nz, dx[0], dx[1], dx[2], p,
                       static_cast<float*>(density), partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return int(e);
  // Simpson 1/9 and the pixel area dW dH
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(kObsThreads), 0, s, partials, nblocks,
                     p.dWd * p.dHd / 9.0, out7, 7L);
  return int(hipGetLastError());
}

}  // namespace wfs
