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
