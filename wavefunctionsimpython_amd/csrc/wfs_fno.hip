// Elementwise kernels of the Fourier neural operator and of the parareal
// driver (no reference counterpart).  Channel-major activations: x[c][p],
// c < C channels, p < P grid points, FP32.
#include "wfs_internal.h"

#include <cmath>

namespace wfs {
namespace {

constexpr int kT = 256;

unsigned blocks_of(long n) {
  long b = (n + kT - 1) / kT;
  if (b > 65535L * 16) b = 65535L * 16;
  return unsigned(b);
}

// u[c][p] += bias[c]; v = gelu(u) (or v = u)
__global__ __launch_bounds__(kT) void bias_act_kernel(float* __restrict__ u, const float* __restrict__ bias,
                                                      float* __restrict__ v, long C, long P, int act) {
  const long n = C * P;
  for (long i = long(blockIdx.x) * kT + threadIdx.x; i < n; i += long(gridDim.x) * kT) {
    const float x = u[i] + bias[i / P];
    u[i] = x;
    if (v) v[i] = act ? gelu_f(x) : x;
  }
}

// du = dv * gelu'(u)
__global__ __launch_bounds__(kT) void gelu_back_kernel(const float* __restrict__ u, const float* __restrict__ dv,
                                                       float* __restrict__ du, long n) {
  for (long i = long(blockIdx.x) * kT + threadIdx.x; i < n; i += long(gridDim.x) * kT) du[i] = dv[i] * gelu_grad(u[i]);
}

// features of a state: [pscale Re psi, pscale Im psi, vscale Re V] as float planes
template <typename R>
__global__ __launch_bounds__(kT) void pack_kernel(const cx<R>* __restrict__ psi, const float2* __restrict__ V,
                                                  float pscale, float vscale, float* __restrict__ a, long P,
                                                  int with_v) {
  for (long i = long(blockIdx.x) * kT + threadIdx.x; i < P; i += long(gridDim.x) * kT) {
    const cx<R> z = psi[i];
    a[i] = pscale * float(z.x);
    a[P + i] = pscale * float(z.y);
    if (with_v) a[2 * P + i] = vscale * V[i].x;
  }
}

// complex64 state = oscale * (plane 0 + i plane 1)
__global__ __launch_bounds__(kT) void unpack_kernel(const float* __restrict__ out, float oscale,
                                                    float2* __restrict__ psi, long P) {
  for (long i = long(blockIdx.x) * kT + threadIdx.x; i < P; i += long(gridDim.x) * kT)
    psi[i] = make_float2(oscale * out[i], oscale * out[P + i]);
}

// partial sums of (pred - target)^2 and target^2 in FP64; dpred = 2 (pred - target) * gscale
__global__ __launch_bounds__(kT) void mse_kernel(const float* __restrict__ pred, const float* __restrict__ target,
                                                 float* __restrict__ dpred, double gscale, long n,
                                                 double* __restrict__ partials) {
  double e = 0.0, t = 0.0;
  for (long i = long(blockIdx.x) * kT + threadIdx.x; i < n; i += long(gridDim.x) * kT) {
    const double d = double(pred[i]) - double(target[i]);
    e += d * d;
    t += double(target[i]) * double(target[i]);
    if (dpred) dpred[i] = float(2.0 * d * gscale);
  }
  for (int off = 32; off > 0; off >>= 1) {
    e += __shfl_xor(e, off, 64);
    t += __shfl_xor(t, off, 64);
  }
  __shared__ double se[kT / 64], st[kT / 64];
  if ((threadIdx.x & 63) == 0) { se[threadIdx.x >> 6] = e; st[threadIdx.x >> 6] = t; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int w = 0; w < kT / 64; ++w) { a += se[w]; b += st[w]; }
    partials[2 * blockIdx.x] = a;
    partials[2 * blockIdx.x + 1] = b;
  }
}

__global__ __launch_bounds__(kT) void sum2_kernel(const double* __restrict__ partials, long nb, double* __restrict__ out) {
  double a = 0.0, b = 0.0;
  for (long i = threadIdx.x; i < nb; i += kT) { a += partials[2 * i]; b += partials[2 * i + 1]; }
  for (int off = 32; off > 0; off >>= 1) { a += __shfl_xor(a, off, 64); b += __shfl_xor(b, off, 64); }
  __shared__ double sa[kT / 64], sb[kT / 64];
  if ((threadIdx.x & 63) == 0) { sa[threadIdx.x >> 6] = a; sb[threadIdx.x >> 6] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 0.0, y = 0.0;
    for (int w = 0; w < kT / 64; ++w) { x += sa[w]; y += sb[w]; }
    out[0] = x;
    out[1] = y;
  }
}

// Adam on FP32 master weights (bias-corrected)
__global__ __launch_bounds__(kT) void adam_kernel(float* __restrict__ w, const float* __restrict__ g,
                                                  float* __restrict__ m, float* __restrict__ v, long n, float lr,
                                                  float b1, float b2, float eps, float c1, float c2) {
  for (long i = long(blockIdx.x) * kT + threadIdx.x; i < n; i += long(gridDim.x) * kT) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// parareal update in FP64: U = Gnew + Fold - Gold (complex64 inputs, complex128 result);
// also the FP64 squared change |U - Uprev|^2 and |U|^2 per block
__global__ __launch_bounds__(kT) void parareal_kernel(const float2* __restrict__ gnew, const float2* __restrict__ fold,
                                                      const float2* __restrict__ gold, double2* __restrict__ U,
                                                      long n, double* __restrict__ partials) {
  double dch = 0.0, nrm = 0.0;
  for (long i = long(blockIdx.x) * kT + threadIdx.x; i < n; i += long(gridDim.x) * kT) {
    const double x = (double(gnew[i].x) + double(fold[i].x)) - double(gold[i].x);
    const double y = (double(gnew[i].y) + double(fold[i].y)) - double(gold[i].y);
    const double2 old = U[i];
    dch += (x - old.x) * (x - old.x) + (y - old.y) * (y - old.y);
    nrm += x * x + y * y;
    U[i] = make_double2(x, y);
  }
  for (int off = 32; off > 0; off >>= 1) { dch += __shfl_xor(dch, off, 64); nrm += __shfl_xor(nrm, off, 64); }
  __shared__ double sa[kT / 64], sb[kT / 64];
  if ((threadIdx.x & 63) == 0) { sa[threadIdx.x >> 6] = dch; sb[threadIdx.x >> 6] = nrm; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int w = 0; w < kT / 64; ++w) { a += sa[w]; b += sb[w]; }
    partials[2 * blockIdx.x] = a;
    partials[2 * blockIdx.x + 1] = b;
  }
}

// y = alpha x + beta y (y is not read when beta == 0)
__global__ __launch_bounds__(kT) void axpby_kernel(float* __restrict__ y, const float* __restrict__ x, float alpha,
                                                   float beta, long n) {
  for (long i = long(blockIdx.x) * kT + threadIdx.x; i < n; i += long(gridDim.x) * kT)
    y[i] = (beta != 0.0f) ? alpha * x[i] + beta * y[i] : alpha * x[i];
}

// complex128 -> complex64 (rounded) for feeding the propagators
__global__ __launch_bounds__(kT) void narrow_kernel(const double2* __restrict__ U, float2* __restrict__ out, long n) {
  for (long i = long(blockIdx.x) * kT + threadIdx.x; i < n; i += long(gridDim.x) * kT)
    out[i] = make_float2(float(U[i].x), float(U[i].y));
}

// Spectrum of an n_dst^3 grid from that of an n_src^3 grid (complex64, batch stacked grids):
// dst[k] = scale * src[k] for every signed frequency k (numpy fftfreq order) both grids hold,
// 0 for the frequencies only dst holds.  Downsampling keeps the lowest n_dst per axis,
// upsampling pads with zeros; with ortho transforms scale = (n_dst / n_src)^(3/2) keeps the
// sampled function's values.
__device__ __forceinline__ int signed_freq(int j, int n) { return j <= (n - 1) / 2 ? j : j - n; }

__global__ __launch_bounds__(kT) void resample_kernel(const float2* __restrict__ src, float2* __restrict__ dst,
                                                      int s0, int s1, int s2, int d0, int d1, int d2, long batch,
                                                      float scale, int h0, int h1, int h2) {
  const long per = long(d0) * d1 * d2;
  const long n = per * batch;
  for (long i = long(blockIdx.x) * kT + threadIdx.x; i < n; i += long(gridDim.x) * kT) {
    const long b = i / per;
    long r = i - b * per;
    const int j2 = int(r % d2);
    r /= d2;
    const int j1 = int(r % d1);
    const int j0 = int(r / d1);
    // destination frequency f reads source frequency f + h (h = 0: the same frequency)
    const int f0 = signed_freq(j0, d0) + h0, f1 = signed_freq(j1, d1) + h1, f2 = signed_freq(j2, d2) + h2;
    // the source holds the signed frequencies -s/2 .. (s-1)/2 of each axis
    const bool in = f0 >= -(s0 / 2) && f0 <= (s0 - 1) / 2 && f1 >= -(s1 / 2) && f1 <= (s1 - 1) / 2 &&
                    f2 >= -(s2 / 2) && f2 <= (s2 - 1) / 2;
    float2 v = make_float2(0.0f, 0.0f);
    if (in) {
      const int k0 = f0 < 0 ? f0 + s0 : f0, k1 = f1 < 0 ? f1 + s1 : f1, k2 = f2 < 0 ? f2 + s2 : f2;
      const float2 x = src[b * (long(s0) * s1 * s2) + (long(k0) * s1 + k1) * s2 + k2];
      v = make_float2(x.x * scale, x.y * scale);
    }
    dst[i] = v;
  }
}

}  // namespace
}  // namespace wfs

using namespace wfs;

#define WFS_LAUNCH_CHECK() WFS_HIP_CHECK(hipGetLastError())

extern "C" int wfs_fno_bias_act(float* u, const float* bias, float* v, long C, long P, int act, void* stream) {
  if (!u || !bias || C < 1 || P < 1) { set_last_error("wfs_fno_bias_act: bad argument"); return WFS_ERR_INVALID; }
  hipLaunchKernelGGL(bias_act_kernel, dim3(blocks_of(C * P)), dim3(kT), 0, static_cast<hipStream_t>(stream), u, bias, v,
                     C, P, act);
  WFS_LAUNCH_CHECK();
  return 0;
}

extern "C" int wfs_fno_axpby(float* y, const float* x, float alpha, float beta, long n, void* stream) {
  if (!y || !x || n < 1) { set_last_error("wfs_fno_axpby: bad argument"); return WFS_ERR_INVALID; }
  hipLaunchKernelGGL(axpby_kernel, dim3(blocks_of(n)), dim3(kT), 0, static_cast<hipStream_t>(stream), y, x, alpha,
                     beta, n);
  WFS_LAUNCH_CHECK();
  return 0;
}

extern "C" int wfs_fno_gelu_backward(const float* u, const float* dv, float* du, long n, void* stream) {
  if (!u || !dv || !du || n < 1) { set_last_error("wfs_fno_gelu_backward: bad argument"); return WFS_ERR_INVALID; }
  hipLaunchKernelGGL(gelu_back_kernel, dim3(blocks_of(n)), dim3(kT), 0, static_cast<hipStream_t>(stream), u, dv, du, n);
  WFS_LAUNCH_CHECK();
  return 0;
}

extern "C" int wfs_fno_pack(const void* psi, int precision, const void* V, float pscale, float vscale, float* a,
                            long P, void* stream) {
  if (!psi || !a || P < 1) { set_last_error("wfs_fno_pack: bad argument"); return WFS_ERR_INVALID; }
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (precision == WFS_C128)
    hipLaunchKernelGGL(pack_kernel<double>, dim3(blocks_of(P)), dim3(kT), 0, s, static_cast<const double2*>(psi),
                       static_cast<const float2*>(V), pscale, vscale, a, P, V ? 1 : 0);
  else
    hipLaunchKernelGGL(pack_kernel<float>, dim3(blocks_of(P)), dim3(kT), 0, s, static_cast<const float2*>(psi),
                       static_cast<const float2*>(V), pscale, vscale, a, P, V ? 1 : 0);
  WFS_LAUNCH_CHECK();
  return 0;
}

extern "C" int wfs_fno_unpack(const float* out, float oscale, void* psi, long P, void* stream) {
  if (!out || !psi || P < 1) { set_last_error("wfs_fno_unpack: bad argument"); return WFS_ERR_INVALID; }
  hipLaunchKernelGGL(unpack_kernel, dim3(blocks_of(P)), dim3(kT), 0, static_cast<hipStream_t>(stream), out,
                     oscale, static_cast<float2*>(psi), P);
  WFS_LAUNCH_CHECK();
  return 0;
}

// out2 (device, 2 doubles) = {sum (pred-target)^2, sum target^2}; dpred = 2 (pred-target) * gscale if non-null
extern "C" int wfs_fno_mse(const float* pred, const float* target, float* dpred, double gscale, long n,
                           double* partials, long partial_elems, double* out2, void* stream) {
  if (!pred || !target || !partials || !out2 || n < 1) { set_last_error("wfs_fno_mse: bad argument"); return WFS_ERR_INVALID; }
  long nb = (n + kT - 1) / kT;
  if (nb > 4096) nb = 4096;
  if (partial_elems < 2 * nb) { set_last_error("wfs_fno_mse: workspace needs 8192 doubles"); return WFS_ERR_INVALID; }
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(mse_kernel, dim3(unsigned(nb)), dim3(kT), 0, s, pred, target, dpred, gscale, n, partials);
  WFS_LAUNCH_CHECK();
  hipLaunchKernelGGL(sum2_kernel, dim3(1), dim3(kT), 0, s, partials, nb, out2);
  WFS_LAUNCH_CHECK();
  return 0;
}

extern "C" int wfs_fno_adam(float* w, const float* g, float* m, float* v, long n, float lr, float beta1, float beta2,
                            float eps, int step, void* stream) {
  if (!w || !g || !m || !v || n < 1 || step < 1) { set_last_error("wfs_fno_adam: bad argument"); return WFS_ERR_INVALID; }
  const float c1 = 1.0f - std::pow(beta1, float(step));
  const float c2 = 1.0f - std::pow(beta2, float(step));
  hipLaunchKernelGGL(adam_kernel, dim3(blocks_of(n)), dim3(kT), 0, static_cast<hipStream_t>(stream), w, g, m, v, n,
                     lr, beta1, beta2, eps, c1, c2);
  WFS_LAUNCH_CHECK();
  return 0;
}

// out2 (device) = {|U - Uprev|^2, |U|^2} in FP64
extern "C" int wfs_parareal_update(const void* gnew, const void* fold, const void* gold, void* U, long n,
                                   double* partials, long partial_elems, double* out2, void* stream) {
  if (!gnew || !fold || !gold || !U || !partials || !out2 || n < 1) { set_last_error("wfs_parareal_update: bad argument"); return WFS_ERR_INVALID; }
  long nb = (n + kT - 1) / kT;
  if (nb > 4096) nb = 4096;
  if (partial_elems < 2 * nb) { set_last_error("wfs_parareal_update: workspace needs 8192 doubles"); return WFS_ERR_INVALID; }
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(parareal_kernel, dim3(unsigned(nb)), dim3(kT), 0, s, static_cast<const float2*>(gnew),
                     static_cast<const float2*>(fold), static_cast<const float2*>(gold), static_cast<double2*>(U), n,
                     partials);
  WFS_LAUNCH_CHECK();
  hipLaunchKernelGGL(sum2_kernel, dim3(1), dim3(kT), 0, s, partials, nb, out2);
  WFS_LAUNCH_CHECK();
  return 0;
}

extern "C" int wfs_narrow_c128(const void* U, void* out, long n, void* stream) {
  if (!U || !out || n < 1) { set_last_error("wfs_narrow_c128: bad argument"); return WFS_ERR_INVALID; }
  hipLaunchKernelGGL(narrow_kernel, dim3(blocks_of(n)), dim3(kT), 0, static_cast<hipStream_t>(stream),
                     static_cast<const double2*>(U), static_cast<float2*>(out), n);
  WFS_LAUNCH_CHECK();
  return 0;
}

extern "C" int wfs_spectral_resample(const void* src, const int nsrc[3], void* dst, const int ndst[3], long batch,
                                     double scale, const int shift[3], void* stream) {
  if (!src || !dst || !nsrc || !ndst || batch < 1) { set_last_error("wfs_spectral_resample: bad argument"); return WFS_ERR_INVALID; }
  for (int a = 0; a < 3; ++a)
    if (nsrc[a] < 1 || ndst[a] < 1) { set_last_error("wfs_spectral_resample: bad grid"); return WFS_ERR_INVALID; }
  if (src == dst) { set_last_error("wfs_spectral_resample: source and destination must differ"); return WFS_ERR_INVALID; }
  int h[3] = {0, 0, 0};
  if (shift)
    for (int a = 0; a < 3; ++a) {
      if (shift[a] < -nsrc[a] || shift[a] > nsrc[a]) {
        set_last_error("wfs_spectral_resample: shift %d on axis %d beyond the source grid of %d", shift[a], a, nsrc[a]);
        return WFS_ERR_INVALID;
      }
      h[a] = shift[a];
    }
  const long n = long(ndst[0]) * ndst[1] * ndst[2] * batch;
  hipLaunchKernelGGL(resample_kernel, dim3(blocks_of(n)), dim3(kT), 0, static_cast<hipStream_t>(stream),
                     static_cast<const float2*>(src), static_cast<float2*>(dst), nsrc[0], nsrc[1], nsrc[2], ndst[0],
                     ndst[1], ndst[2], batch, float(scale), h[0], h[1], h[2]);
  WFS_LAUNCH_CHECK();
  return 0;
}
