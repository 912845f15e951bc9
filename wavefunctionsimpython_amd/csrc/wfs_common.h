// Shared device helpers for the wfsim HIP library (gfx950).
//
// Complex values are interleaved (re, im) pairs stored as HIP float2/double2
// so that a grid point moves in one 8-byte (complex64) or 16-byte
// (complex128) access.  Grids are C-order: x slowest, z fastest
// (sources/cuda_kernels/common.cu:294-299).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <utility>

namespace wfs {

template <typename R> struct cplx_t;
template <> struct cplx_t<float> { using type = float2; };
template <> struct cplx_t<double> { using type = double2; };
template <typename R> using cx = typename cplx_t<R>::type;

template <typename R>
__host__ __device__ __forceinline__ cx<R> mk(R x, R y) { cx<R> r; r.x = x; r.y = y; return r; }

template <typename T>
__host__ __device__ __forceinline__ T cadd(T a, T b) { a.x += b.x; a.y += b.y; return a; }
template <typename T>
__host__ __device__ __forceinline__ T csub(T a, T b) { a.x -= b.x; a.y -= b.y; return a; }
// Complex products with the fused multiply-adds written out: which of the two
// products a contracting compiler would fuse depends on the surrounding code, so a
// kernel that repeats another's arithmetic could round differently.
__host__ __device__ __forceinline__ float fma_r(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__host__ __device__ __forceinline__ double fma_r(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename T>
__host__ __device__ __forceinline__ T cmul(T a, T b) {
  T r;
  r.x = fma_r(a.x, b.x, -(a.y * b.y));
  r.y = fma_r(a.x, b.y, a.y * b.x);
  return r;
}
// a * conj(b)
template <typename T>
__host__ __device__ __forceinline__ T cmulc(T a, T b) {
  T r;
  r.x = fma_r(a.x, b.x, a.y * b.y);
  r.y = fma_r(a.y, b.x, -(a.x * b.y));
  return r;
}
template <typename T, typename S>
__host__ __device__ __forceinline__ T cscale(T a, S s) { a.x *= s; a.y *= s; return a; }

// element T at a 32-bit byte offset from a base address: with a block-uniform base the
// address is the scalar base plus one 32-bit vector offset (no 64-bit vector arithmetic per
// access); the caller keeps the offset below 2^32
template <typename T>
__device__ __forceinline__ T& byte_at(T* base, unsigned off) {
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + off);
}
template <typename T>
__device__ __forceinline__ const T& byte_at(const T* base, unsigned off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
}

// a * (hi + lo) and a * conj(hi + lo) for a factor held as an unevaluated sum of two
// floats (hi = float(w), lo = float(w - hi)).  A float-rounded factor applied every
// step to the same mode errs the same way every step, so its error grows linearly
// with the step count; with the lo part the factor is exact to about 2^-48 and only
// the (data-dependent) rounding of the products remains.
// The small products a*lo are formed first and enter the fused multiply-adds of a*hi,
// so they are added before the main products are rounded (added afterwards they would
// mostly vanish below half an ulp of the rounded product).
__host__ __device__ __forceinline__ float2 cmul2(float2 a, float2 hi, float2 lo) {
  const float cx = __builtin_fmaf(a.x, lo.x, -(a.y * lo.y));
  const float cy = __builtin_fmaf(a.x, lo.y, a.y * lo.x);
  float2 r;
  r.x = __builtin_fmaf(a.x, hi.x, __builtin_fmaf(-a.y, hi.y, cx));
  r.y = __builtin_fmaf(a.x, hi.y, __builtin_fmaf(a.y, hi.x, cy));
  return r;
}
__host__ __device__ __forceinline__ float2 cmulc2(float2 a, float2 hi, float2 lo) {
  const float cx = __builtin_fmaf(a.x, lo.x, a.y * lo.y);
  const float cy = __builtin_fmaf(a.y, lo.x, -(a.x * lo.y));
  float2 r;
  r.x = __builtin_fmaf(a.x, hi.x, __builtin_fmaf(a.y, hi.y, cx));
  r.y = __builtin_fmaf(a.y, hi.x, __builtin_fmaf(-a.x, hi.y, cy));
  return r;
}

// ---------------------------------------------------------------------------
// Compile-time cos/sin of 2*pi*e/M for the in-register DFT twiddles.
// Exact values are returned for multiples of pi/4; otherwise the argument is
// reduced to [-pi/4, pi/4] by octant and a Taylor series (error < 1e-17 on
// that interval after 14 terms) is summed in double.
// ---------------------------------------------------------------------------
constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kSqrtHalf = 0.707106781186547524400844362104849039;

constexpr double taylor_sin(double x) {
  double term = x, sum = x;
  for (int k = 1; k < 14; ++k) {
    term *= -x * x / double((2 * k) * (2 * k + 1));
    sum += term;
  }
  return sum;
}
constexpr double taylor_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 14; ++k) {
    term *= -x * x / double((2 * k - 1) * (2 * k));
    sum += term;
  }
  return sum;
}
// returns {cos, sin} of 2*pi*e/M for 0 <= e < M
struct cs_pair { double c, s; };
constexpr cs_pair cos_sin_2pi(long e, long M) {
  // 8*e/M = octant + frac, angle = (octant + frac) * pi/4
  long num = 8 * e;           // angle in units of pi/4 = num / M
  long oct = num / M;         // 0..7
  long rem = num - oct * M;   // remainder in [0, M)
  // reduce to the nearest multiple of pi/4: angle = q*pi/4 + d, |d| <= pi/8
  long q = oct;
  double d = double(rem) / double(M) * (kPi / 4.0);
  if (2 * rem > M) { q = oct + 1; d = -double(M - rem) / double(M) * (kPi / 4.0); }
  q &= 7;
  const double cq[8] = {1.0, kSqrtHalf, 0.0, -kSqrtHalf, -1.0, -kSqrtHalf, 0.0, kSqrtHalf};
  const double sq[8] = {0.0, kSqrtHalf, 1.0, kSqrtHalf, 0.0, -kSqrtHalf, -1.0, -kSqrtHalf};
  if (rem == 0 || 2 * rem == 2 * M) return {cq[q], sq[q]};
  double cd = taylor_cos(d), sd = taylor_sin(d);
  return {cq[q] * cd - sq[q] * sd, sq[q] * cd + cq[q] * sd};
}

// multiply by exp(SIGN * 2*pi*i * E / M), E known at compile time
template <int M, int E, int SIGN, typename T>
__device__ __forceinline__ T twiddle_ct(T a) {
  constexpr int e = ((E % M) + M) % M;
  using R = decltype(a.x);
  if constexpr (e == 0) {
    return a;
  } else if constexpr (2 * e == M) {
    a.x = -a.x; a.y = -a.y; return a;
  } else if constexpr (4 * e == M) {        // exp(SIGN*i*pi/2) = SIGN*i
    T r;
    if constexpr (SIGN > 0) { r.x = -a.y; r.y = a.x; } else { r.x = a.y; r.y = -a.x; }
    return r;
  } else if constexpr (4 * e == 3 * M) {    // exp(SIGN*3i*pi/2) = -SIGN*i
    T r;
    if constexpr (SIGN > 0) { r.x = a.y; r.y = -a.x; } else { r.x = -a.y; r.y = a.x; }
    return r;
  } else if constexpr (sizeof(R) == 4 && (8 * e) % M == 0) {
    // odd multiples of pi/4: w = (cx + i cy) sqrt(1/2) with cx, cy = +-1, so
    // a w = ((cx a.x - cy a.y) + i (cx a.y + cy a.x)) sqrt(1/2): one addition per part, then
    // the product with sqrt(1/2) as hi + lo (6 instructions instead of the 8 of cmul2)
    constexpr int o = 8 * e / M;                               // 1, 3, 5 or 7
    constexpr R cx = (o == 1 || o == 7) ? R(1) : R(-1);
    constexpr R cy = ((o == 1 || o == 3) == (SIGN > 0)) ? R(1) : R(-1);
    constexpr R h = R(kSqrtHalf);
    constexpr R l = R(kSqrtHalf - double(R(kSqrtHalf)));
    const R px = cx * a.x - cy * a.y;
    const R py = cx * a.y + cy * a.x;
    T r;
    r.x = __builtin_fmaf(px, h, px * l);
    r.y = __builtin_fmaf(py, h, py * l);
    return r;
  } else {
    constexpr cs_pair cs = cos_sin_2pi(e, M);
    T w;
    w.x = R(cs.c);
    w.y = R(SIGN > 0 ? cs.s : -cs.s);
    if constexpr (sizeof(R) == 4) {
      // float state: the twiddle as hi + lo (see cmul2)
      T lo;
      lo.x = R(cs.c - double(R(cs.c)));
      lo.y = R((SIGN > 0 ? cs.s : -cs.s) - double(R(SIGN > 0 ? cs.s : -cs.s)));
      return cmul2(a, w, lo);
    } else {
      return cmul(a, w);
    }
  }
}

// The value of lane (lane ^ M) of the wavefront, M in {1, 2, 4, 8, 16, 32}, by cross-lane VALU
// moves instead of LDS-unit permutes (exact moves, so sums formed from them are unchanged):
// DPP quad permutations for 1 and 2; for 4, row rotations by 4 and 12 written to alternate
// 4-lane banks; a row rotation by 8 for 8; gfx950's row and half-wave swaps of the value with
// itself (v_permlane16_swap, v_permlane32_swap) for 16 and 32, each leaving the partner's value
// in one of its two results, picked by the lane.
template <int M>
__device__ __forceinline__ int xor_lane_b32(int v) {
  if constexpr (M == 1) {
    return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);      // quad_perm [1, 0, 3, 2]
  } else if constexpr (M == 2) {
    return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);      // quad_perm [2, 3, 0, 1]
  } else if constexpr (M == 4) {
    const int odd = __builtin_amdgcn_update_dpp(0, v, 0x124, 0xF, 0xA, false);   // row_ror:4, banks 1, 3
    return __builtin_amdgcn_update_dpp(odd, v, 0x12C, 0xF, 0x5, false);          // row_ror:12, banks 0, 2
  } else if constexpr (M == 8) {
    return __builtin_amdgcn_update_dpp(0, v, 0x128, 0xF, 0xF, false);     // row_ror:8
  } else if constexpr (M == 16) {
    const auto r = __builtin_amdgcn_permlane16_swap(unsigned(v), unsigned(v), false, false);
    return int((__lane_id() & 16) ? r[0] : r[1]);
  } else {
    static_assert(M == 32, "xor distance 1, 2, 4, 8, 16 or 32");
    const auto r = __builtin_amdgcn_permlane32_swap(unsigned(v), unsigned(v), false, false);
    return int((__lane_id() & 32) ? r[0] : r[1]);
  }
}
template <int M>
__device__ __forceinline__ double xor_lane(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = xor_lane_b32<M>(int(b & 0xffffffffLL)), hi = xor_lane_b32<M>(int(b >> 32));
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}

// Sums of seven values over the 64 lanes of a wavefront by recursive halving: at the xor-32,
// -16 and -8 steps each lane keeps half of its remaining values and sends the other half to
// its partner, then xor 4, 2, 1 finish one value; 10 exchanges instead of 7 x 6.  Lane l
// returns the sum of value 4 b5 + 2 b4 + b3 (b the bits of l; 7 means none).
__device__ __forceinline__ double wave_sum7(const double (&v)[7], int lane) {
  const bool h5 = lane & 32, h4 = lane & 16, h3 = lane & 8;
  double y[4], z[2];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const double lo = v[j], hi = j + 4 < 7 ? v[j + 4] : 0.0;
    y[j] = (h5 ? hi : lo) + xor_lane<32>(h5 ? lo : hi);
  }
#pragma unroll
  for (int j = 0; j < 2; ++j) z[j] = (h4 ? y[j + 2] : y[j]) + xor_lane<16>(h4 ? y[j] : y[j + 2]);
  double s = (h3 ? z[1] : z[0]) + xor_lane<8>(h3 ? z[0] : z[1]);
  s += xor_lane<4>(s);
  s += xor_lane<2>(s);
  s += xor_lane<1>(s);
  return s;
}

// compile-time loop
template <typename F, int... Is>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, Is...>) {
  (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// ---------------------------------------------------------------------------
// In-register DFT of length M, natural order in and out:
//   a[k] <- sum_n a[n] exp(SIGN * 2*pi*i * n*k / M)
// Built by decimation in time, M = R0 * Q (R0 in {4, 3, 2}), fully unrolled.
// ---------------------------------------------------------------------------
template <int M, int SIGN>
struct DFT;

template <int SIGN>
struct DFT<1, SIGN> {
  template <typename T> __device__ __forceinline__ static void run(T*) {}
};

template <int SIGN>
struct DFT<2, SIGN> {
  template <typename T> __device__ __forceinline__ static void run(T* a) {
    T t0 = cadd(a[0], a[1]);
    T t1 = csub(a[0], a[1]);
    a[0] = t0; a[1] = t1;
  }
};

template <int SIGN>
struct DFT<3, SIGN> {
  template <typename T> __device__ __forceinline__ static void run(T* a) {
    using R = decltype(a[0].x);
    const R c = R(-0.5);
    const R s = R(0.866025403784438646763723170752936183);  // sqrt(3)/2
    T t = cadd(a[1], a[2]);
    T d = csub(a[1], a[2]);
    T m;  // a0 + c*(a1+a2)
    m.x = a[0].x + c * t.x;
    m.y = a[0].y + c * t.y;
    // SIGN*i*s*(a1 - a2); for float, s as hi + lo with the low product inside the fma
    R sdx, sdy;
    if constexpr (sizeof(R) == 4) {
      constexpr double sd = 0.866025403784438646763723170752936183;
      const R slo = R(sd - double(R(sd)));
      sdx = __builtin_fmaf(s, d.x, slo * d.x);
      sdy = __builtin_fmaf(s, d.y, slo * d.y);
    } else {
      sdx = s * d.x;
      sdy = s * d.y;
    }
    T q;
    if constexpr (SIGN > 0) { q.x = -sdy; q.y = sdx; } else { q.x = sdy; q.y = -sdx; }
    a[0] = cadd(a[0], t);
    a[1] = cadd(m, q);
    a[2] = csub(m, q);
  }
};

template <int SIGN>
struct DFT<4, SIGN> {
  template <typename T> __device__ __forceinline__ static void run(T* a) {
    T s02 = cadd(a[0], a[2]);
    T d02 = csub(a[0], a[2]);
    T s13 = cadd(a[1], a[3]);
    T d13 = csub(a[1], a[3]);
    // SIGN*i*d13
    T j;
    if constexpr (SIGN > 0) { j.x = -d13.y; j.y = d13.x; } else { j.x = d13.y; j.y = -d13.x; }
    a[0] = cadd(s02, s13);
    a[2] = csub(s02, s13);
    a[1] = cadd(d02, j);
    a[3] = csub(d02, j);
  }
};

template <int M>
struct radix_of {
  static constexpr int value = (M % 4 == 0) ? 4 : (M % 3 == 0) ? 3 : 2;
};

template <int M, int SIGN>
struct DFT {
  static constexpr int R0 = radix_of<M>::value;
  static constexpr int Q = M / R0;
  static_assert(R0 * Q == M, "unsupported DFT length");
  template <typename T> __device__ __forceinline__ static void run(T* a) {
    T y[R0][Q];
    static_for<R0>([&](auto r) {
      static_for<Q>([&](auto q) { y[r][q] = a[R0 * q + r]; });
      DFT<Q, SIGN>::run(y[r]);
    });
    static_for<Q>([&](auto k1) {
      T t[R0];
      static_for<R0>([&](auto r) {
        t[r] = twiddle_ct<M, decltype(r)::value * decltype(k1)::value, SIGN>(y[r][k1]);
      });
      DFT<R0, SIGN>::run(t);
      static_for<R0>([&](auto k2) { a[k1 + Q * k2] = t[k2]; });
    });
  }
};

// ---------------------------------------------------------------------------
// One length-N transform of a column shared by N2 threads, N = N1 * N2, N1 % N2 == 0:
//   X[k1 + N1 k2] = sum_n2 W_N2^{n2 k2} W_N^{n2 k1} sum_n1 x[N2 n1 + n2] W_N1^{n1 k1}
// Thread t (0 <= t < N2) holds a[i] = x[t + N2 i] on entry and a[i] = X[t + N2 i] on
// exit; w[k1] = W_N^{t k1} (forward twiddles; the inverse uses their conjugates).
// lidx(k1, n2) names the LDS slot of element (k1, n2) of this column.  The routine
// synchronises the block once between its LDS writes and reads; the caller
// synchronises before the same slots are written again.
// ---------------------------------------------------------------------------
// First step of column_dft: the length-N1 DFT in registers and the table twiddles.
template <int N, int N2, int SIGN, typename C>
__device__ __forceinline__ void column_dft_first(C (&a)[N / N2], const C (&w)[N / N2]) {
  constexpr int N1 = N / N2;
  static_assert(N1 * N2 == N && N1 % N2 == 0, "bad split");
  DFT<N1, SIGN>::run(a);
#pragma unroll
  for (int k1 = 1; k1 < N1; ++k1) a[k1] = (SIGN < 0) ? cmul(a[k1], w[k1]) : cmulc(a[k1], w[k1]);
}

// Second step of column_dft: LDS exchange and the length-N2 DFTs.
template <int N, int N2, int SIGN, typename C, typename LIDX>
__device__ __forceinline__ void column_dft_second(C (&a)[N / N2], C* lds, int t, LIDX lidx) {
  constexpr int N1 = N / N2;
  constexpr int M = N1 / N2;            // length-N2 DFTs per thread in the second step
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) lds[lidx(k1, t)] = a[k1];
  __syncthreads();
#pragma unroll
  for (int m = 0; m < M; ++m) {
    C b[N2];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) b[n2] = lds[lidx(t + N2 * m, n2)];
    DFT<N2, SIGN>::run(b);
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) a[m + M * k2] = b[k2];
  }
}

template <int N, int N2, int SIGN, typename C, typename LIDX>
__device__ __forceinline__ void column_dft(C (&a)[N / N2], const C (&w)[N / N2], C* lds, int t, LIDX lidx) {
  column_dft_first<N, N2, SIGN>(a, w);
  column_dft_second<N, N2, SIGN>(a, lds, t, lidx);
}

// GELU, tanh form: 0.5 u (1 + tanh(sqrt(2/pi) (u + 0.044715 u^3))), and its derivative
// tanh(t) = 1 - 2 / (exp(2t) + 1), with exp(2t) = 2^(2 t log2 e) from the hardware
// exponential (v_exp_f32) and the hardware reciprocal (v_rcp_f32): a few ulp of 1 in
// absolute terms (what 1 + tanh in the GELU needs), saturating to +-1 for large |t|.
__device__ __forceinline__ float tanh_hw(float t) {
  const float e = __builtin_amdgcn_exp2f(t * 2.8853900817779268f);   // 2 log2(e)
  return 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
}
__device__ __forceinline__ float gelu_f(float u) {
  const float k = 0.7978845608028654f;
  const float t = tanh_hw(k * (u + 0.044715f * u * u * u));
  return 0.5f * u * (1.0f + t);
}
__device__ __forceinline__ float gelu_grad(float u) {
  const float k = 0.7978845608028654f;
  const float a = k * (u + 0.044715f * u * u * u);
  const float t = tanh_hw(a);
  const float da = k * (1.0f + 3.0f * 0.044715f * u * u);
  return 0.5f * (1.0f + t) + 0.5f * u * (1.0f - t * t) * da;
}

// error reporting shared by all translation units
void set_last_error(const char* fmt, ...);

}  // namespace wfs

#define WFS_HIP_CHECK(expr)                                                    \
  do {                                                                         \
    hipError_t _e = (expr);                                                    \
    if (_e != hipSuccess) {                                                    \
      ::wfs::set_last_error("%s failed: %s (%s:%d)", #expr,                    \
                            hipGetErrorString(_e), __FILE__, __LINE__);        \
      return int(_e);                                                          \
    }                                                                          \
  } while (0)
