// Axis passes of the 3-D FFT with fused phase multiplies (gfx950).
//
// A length-N transform along one axis is split N = N1 * N2 (N1 % N2 == 0):
//   X[k1 + N1 k2] = sum_n2 W_N2^{n2 k2} W_N^{n2 k1} sum_n1 x[N2 n1 + n2] W_N1^{n1 k1}
// Thread t (0 <= t < N2) of a column holds a[i] = x[t + N2 i], i < N1, does the
// inner length-N1 DFT in registers, multiplies by the table twiddle W_N^{t k1},
// exchanges through LDS, and finishes with N1/N2 length-N2 DFTs.  With the
// outputs written back to a[m + (N1/N2) k2] the result satisfies
// a[i] = X[t + N2 i]: the register layout of the output equals that of the
// input, so a forward transform, a pointwise phase and an inverse transform
// chain in registers with no extra memory traffic.
//
// Passes (one kernel launch each, in place, unnormalised transforms):
//   ROWS layout  (contiguous z axis):  FWD, INV, IVF = inverse -> exp(-i dt V) -> forward
//   COLS layout  (strided x / y axes): FWD, INV, FKI = forward -> K(kx,ky,kz) -> inverse
// where K holds the kinetic phase and the 1/(Nx Ny Nz) normalisation of the
// round trip.  The reference performs the same arithmetic as four cuFFT calls
// and three elementwise kernels per step (sources/sim_state.py:641-697).
#include "wfs_internal.h"

namespace wfs {

template <typename R, int N, int N2, int LAYOUT, int MODE, int G>
__global__ __launch_bounds__(N2 * G) void fft_pass_kernel(cx<R>* __restrict__ data, const PassParams p) {
  constexpr int N1 = N / N2;
  static_assert(N1 * N2 == N && N1 % N2 == 0, "bad split");
  using C = cx<R>;
  // LDS exchange buffer: element (k1, n2) of group g
  constexpr int ROW_STRIDE = N1 * (N2 + 1);   // padded: conflict-free step-2 reads
  constexpr int LDS_ELEMS = (LAYOUT == LAYOUT_ROWS) ? G * ROW_STRIDE : G * N;
  __shared__ C lds[LDS_ELEMS];

  const int tid = threadIdx.x;
  int t, g;
  if constexpr (LAYOUT == LAYOUT_ROWS) { t = tid % N2; g = tid / N2; }
  else { g = tid % G; t = tid / G; }

  // Addresses: a block-uniform base pointer plus a 32-bit per-lane element
  // offset, so loads and stores can use the scalar-base addressing form (one
  // VGPR per address instead of two).  Offsets stay below 2^31 elements for
  // every supported grid (at most 512^3 points).
  C* __restrict__ blk;
  unsigned loff;          // lane offset of element i = 0
  unsigned step;          // offset between consecutive elements of the column
  bool active;
  long col = 0;           // ROWS: row index; COLS: column index inside one outer slab
  if constexpr (LAYOUT == LAYOUT_ROWS) {
    blk = data + long(blockIdx.x) * G * N;
    col = long(blockIdx.x) * G + g;
    active = col < p.rows;
    loff = unsigned(g * N + t);
    step = N2;
  } else {
    blk = data + long(blockIdx.y) * N * p.inner + long(blockIdx.x) * G;
    col = long(blockIdx.x) * G + g;
    active = col < p.inner;
    loff = unsigned(g) + unsigned(t) * unsigned(p.inner);
    step = unsigned(N2) * unsigned(p.inner);
  }

  auto lidx = [&](int k1, int n2) -> int {
    if constexpr (LAYOUT == LAYOUT_ROWS) return g * ROW_STRIDE + k1 * (N2 + 1) + n2;
    else return (k1 * N2 + n2) * G + g;
  };

  // Table twiddles and kinetic factors are plain floats for complex64 states; the plan
  // cycles dithered copies of those tables step by step so that their rounding does not
  // repeat, and the in-register DFT twiddles (compile-time constants) are applied as
  // hi + lo pairs (wfs_common.h).
  const C* __restrict__ tw = reinterpret_cast<const C*>(p.tw);
  C w[N1];
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) w[k1] = tw[(t * k1) % N];

  C a[N1];
  if (active) {
#pragma unroll
    for (int i = 0; i < N1; ++i) a[i] = blk[loff + unsigned(i) * step];
  } else {
#pragma unroll
    for (int i = 0; i < N1; ++i) a[i] = mk<R>(R(0), R(0));
  }

  // one length-N transform of the column held in a[] (SIGN -1 forward, +1 inverse)
  auto transform = [&](auto sign_c) {
    column_dft<N, N2, decltype(sign_c)::value>(a, w, lds, t, lidx);
  };

  if constexpr (MODE == MODE_FWD || MODE == MODE_INV) {
    transform(std::integral_constant<int, (MODE == MODE_FWD) ? -1 : +1>{});
    if (p.scale != 1.0) {
      const R sc = R(p.scale);
#pragma unroll
      for (int i = 0; i < N1; ++i) a[i] = cscale(a[i], sc);
    }
  } else if constexpr (MODE == MODE_FKI) {
    // x pass in full k-space: column = (ky, kz)
    transform(std::integral_constant<int, -1>{});
    const C* __restrict__ kx = reinterpret_cast<const C*>(p.kx);
    const C* __restrict__ ky = reinterpret_cast<const C*>(p.ky);
    const C* __restrict__ kz = reinterpret_cast<const C*>(p.kz);
    const long cc = active ? col : 0;
    const C kyz = cmul(ky[cc / p.nz], kz[cc % p.nz]);
#pragma unroll
    for (int i = 0; i < N1; ++i) a[i] = cmul(a[i], cmul(kx[t + N2 * i], kyz));
    __syncthreads();
    transform(std::integral_constant<int, +1>{});
  } else if constexpr (MODE == MODE_IVF) {
    // z pass at the turning point: inverse, potential phase table, forward.
    // pv holds cexp_i(-dt V) = (cos(-dt Vr), sin(-dt Vr)) exp(dt Vi) per point
    // (sources/cuda_kernels/common.cu:98-111, potential_operator.cu:10-15).
    transform(std::integral_constant<int, +1>{});
    if (p.spx) {
      // separable potential: exp(-i dt (fx + fy + fz)) = px(x) py(y) pz(z), no N^3 table read
      const C* __restrict__ px = reinterpret_cast<const C*>(p.spx);
      const C* __restrict__ py = reinterpret_cast<const C*>(p.spy);
      const C* __restrict__ pz = reinterpret_cast<const C*>(p.spz);
      const long row = active ? col : 0;
      const long plane = p.snx * p.sny;
      const long pk = row / plane, r = row % plane;
      const C pxy = cmul(px[pk * p.snx + r / p.sny], py[pk * p.sny + r % p.sny]);
      const C* __restrict__ pzb = pz + pk * N;
      if (active) {
#pragma unroll
        for (int i = 0; i < N1; ++i) a[i] = cmul(a[i], cmul(pzb[t + N2 * i], pxy));
      }
    } else {
      const C* __restrict__ pvb = reinterpret_cast<const C*>(p.pv) + long(blockIdx.x) * G * N;
      if (active) {
#pragma unroll
        for (int i = 0; i < N1; ++i) a[i] = cmul(a[i], pvb[loff + unsigned(i) * step]);
      }
    }
    __syncthreads();
    transform(std::integral_constant<int, -1>{});
  }

  if (active) {
#pragma unroll
    for (int i = 0; i < N1; ++i) blk[loff + unsigned(i) * step] = a[i];
  }
}

// ---------------------------------------------------------------------------
// host-side dispatch
// ---------------------------------------------------------------------------

template <int N> struct split_of;
template <> struct split_of<8>   { static constexpr int n2 = 2; };
template <> struct split_of<16>  { static constexpr int n2 = 4; };
template <> struct split_of<32>  { static constexpr int n2 = 4; };
template <> struct split_of<64>  { static constexpr int n2 = 8; };
template <> struct split_of<96>  { static constexpr int n2 = 4; };
template <> struct split_of<128> { static constexpr int n2 = 8; };
template <> struct split_of<192> { static constexpr int n2 = 8; };
template <> struct split_of<256> { static constexpr int n2 = 16; };
template <> struct split_of<512> { static constexpr int n2 = 16; };

// ---------------------------------------------------------------------------
// Fused y-z pass on whole x planes (ny = nz = N).  One block holds one plane,
// N*N points, in registers (N1 per thread) and moves it through LDS:
//   y inverse -> transpose -> z inverse -> potential phase -> z forward
//   -> transpose -> y forward
// Each transform is column_dft with the split, table twiddles and operation
// order of the axis passes y_inv, z_inv_v_fwd and y_fwd, so the result is the
// same; the plane crosses HBM once (16 B per point for complex64) instead of
// three times (48 B per point).  The LDS holds the largest of the column
// exchange (N*N), the padded row exchange (N * N1 * (N2+1)) and the padded
// transposition (N * (N + N2)); for N = 128 complex64 that is 144 KiB, one
// block of 1024 threads per CU.
// ---------------------------------------------------------------------------
constexpr int max3(int a, int b, int c) { return a > b ? (a > c ? a : c) : (b > c ? b : c); }

template <typename R, int N>
constexpr long yz_plane_lds_bytes() {
  constexpr int N2 = split_of<N>::n2, N1 = N / N2;
  return long(max3(N * N, N * N1 * (N2 + 1), N * (N + N2)) + N) * long(sizeof(cx<R>));
}

// An opaque copy of v: the compiler recomputes what derives from it instead of
// keeping values from an earlier stage alive (register pressure).
__device__ __forceinline__ int opaque_v(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

template <typename R, int N>
__global__ __launch_bounds__(N * split_of<N>::n2) void yz_plane_kernel(cx<R>* __restrict__ data,
                                                                        const YzPlaneParams p) {
  constexpr int N2 = split_of<N>::n2;
  constexpr int N1 = N / N2;
  constexpr int TS = N + N2;                 // padded row stride of the transposition
  constexpr int RS = N1 * (N2 + 1);          // padded row stride of the row exchange
  // y-column thread index uniform over a wavefront: y twiddles by scalar loads
  constexpr bool Y_UNIFORM = (N % 64) == 0;
  using C = cx<R>;
  using CC = const __attribute__((address_space(4))) C;
  __shared__ C lds[max3(N * N, N * RS, N * TS)];
  __shared__ C twz[N];                       // z twiddles

  const int tid = threadIdx.x;
  const long plane = blockIdx.x;
  C* __restrict__ blk = data + plane * long(N) * N;
  if (tid < N) twz[tid] = reinterpret_cast<const C*>(p.tw_z)[tid];

  C a[N1];
  C w[N1];
  // y transform on columns: thread ty of column gz holds y = ty + N2 i
  auto y_stage = [&](const void* twp, auto sign_c) {
    constexpr int SIGN = decltype(sign_c)::value;
    const int gz = opaque_v(tid) % N;
    if constexpr (Y_UNIFORM) {
      const int ty = __builtin_amdgcn_readfirstlane(tid / N);
      CC* tw = (CC*)(twp);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) w[k1] = mk<R>(tw[(ty * k1) % N].x, tw[(ty * k1) % N].y);
      column_dft<N, N2, SIGN>(a, w, lds, ty, [&](int k1, int n2) { return (k1 * N2 + n2) * N + gz; });
    } else {
      const int ty = opaque_v(tid) / N;
      const C* __restrict__ tw = reinterpret_cast<const C*>(twp);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) w[k1] = tw[(ty * k1) % N];
      column_dft<N, N2, SIGN>(a, w, lds, ty, [&](int k1, int n2) { return (k1 * N2 + n2) * N + gz; });
    }
  };
  // transposition through LDS between the column and row layouts
  auto cols_to_rows = [&]() {
    const int t = opaque_v(tid);
    const int gz = t % N, ty = t / N, tz = t % N2, gy = t / N2;
#pragma unroll
    for (int i = 0; i < N1; ++i) lds[(ty + N2 * i) * TS + gz] = a[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < N1; ++i) a[i] = lds[gy * TS + tz + N2 * i];
  };
  auto rows_to_cols = [&]() {
    const int t = opaque_v(tid);
    const int gz = t % N, ty = t / N, tz = t % N2, gy = t / N2;
#pragma unroll
    for (int i = 0; i < N1; ++i) lds[gy * TS + tz + N2 * i] = a[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < N1; ++i) a[i] = lds[(ty + N2 * i) * TS + gz];
  };

  {
    const int gz = tid % N, ty = tid / N;
#pragma unroll
    for (int i = 0; i < N1; ++i) a[i] = blk[unsigned(gz + (ty + N2 * i) * N)];
  }

  // y inverse
  y_stage(p.tw_yi, std::integral_constant<int, +1>{});
  __syncthreads();
  cols_to_rows();
  __syncthreads();

  // z inverse, potential phase, z forward on rows: thread tz of row gy holds z = tz + N2 i
  {
    const int tz = tid % N2, gy = tid / N2;
    auto lidx_r = [&](int k1, int n2) -> int { return gy * RS + k1 * (N2 + 1) + n2; };
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) w[k1] = twz[(tz * k1) % N];
    column_dft<N, N2, +1>(a, w, lds, tz, lidx_r);
    if (p.spx) {
      const long b = plane / p.nx;
      const C* __restrict__ px = reinterpret_cast<const C*>(p.spx);
      const C* __restrict__ py = reinterpret_cast<const C*>(p.spy) + b * N;
      const C* __restrict__ pzb = reinterpret_cast<const C*>(p.spz) + b * N;
      const C pxy = cmul(px[plane], py[gy]);
#pragma unroll
      for (int i = 0; i < N1; ++i) a[i] = cmul(a[i], cmul(pzb[tz + N2 * i], pxy));
    } else {
      const C* __restrict__ pvb = reinterpret_cast<const C*>(p.pv) + plane * long(N) * N;
#pragma unroll
      for (int i = 0; i < N1; ++i) a[i] = cmul(a[i], pvb[unsigned(gy * N + tz + N2 * i)]);
    }
    __syncthreads();
    column_dft<N, N2, -1>(a, w, lds, tz, lidx_r);
  }
  __syncthreads();
  rows_to_cols();
  __syncthreads();

  // y forward
  y_stage(p.tw_yf, std::integral_constant<int, -1>{});
  {
    const int t = opaque_v(tid);
    const int gz = t % N, ty = t / N;
#pragma unroll
    for (int i = 0; i < N1; ++i) blk[unsigned(gz + (ty + N2 * i) * N)] = a[i];
  }
}

template <typename R, int N>
static int launch_yz_plane_n(void* data, const YzPlaneParams& p, hipStream_t s) {
  static_assert(yz_plane_lds_bytes<R, N>() <= 160 * 1024, "plane does not fit in LDS");
  if (p.planes <= 0 || p.planes > 0x7fffffffL) {
    set_last_error("yz plane pass: bad plane count %ld", p.planes);
    return WFS_ERR_INVALID;
  }
  hipLaunchKernelGGL((yz_plane_kernel<R, N>), dim3(unsigned(p.planes)), dim3(N * split_of<N>::n2), 0, s,
                     reinterpret_cast<cx<R>*>(data), p);
  return int(hipGetLastError());
}

// Columns (COLS) or rows (ROWS) per block.  The default aims at 256 threads
// with the LDS exchange buffer within 64 KiB (two or more blocks per CU); the
// autotuned variants halve or double it (LDS up to 80 KiB, 1024 threads).
template <typename R, int N, int N2, int LAYOUT>
constexpr long lds_bytes(int g) {
  const long e = (LAYOUT == LAYOUT_ROWS) ? long(g) * (N / N2) * (N2 + 1) : long(g) * N;
  return e * long(sizeof(cx<R>));
}

template <typename R, int N, int N2, int LAYOUT>
constexpr int pick_groups() {
  int g = 256 / N2;
  if (g > 64) g = 64;
  while (g > 4 && lds_bytes<R, N, N2, LAYOUT>(g) > 65536) g /= 2;
  return g;
}

template <typename R, int N, int N2, int LAYOUT>
constexpr bool groups_ok(int g) {
  return g >= 2 && g <= 128 && N2 * g <= 1024 && lds_bytes<R, N, N2, LAYOUT>(g) <= 81920;
}

template <typename R, int N, int LAYOUT, int MODE, int G>
static int launch_g(void* data, const PassParams& p, long outer, hipStream_t s) {
  constexpr int N2 = split_of<N>::n2;
  const long ncols = (LAYOUT == LAYOUT_ROWS) ? p.rows : p.inner;
  dim3 grid(unsigned((ncols + G - 1) / G), unsigned(LAYOUT == LAYOUT_ROWS ? 1 : outer), 1);
  hipLaunchKernelGGL((fft_pass_kernel<R, N, N2, LAYOUT, MODE, G>), grid, dim3(N2 * G), 0, s,
                     reinterpret_cast<cx<R>*>(data), p);
  return int(hipGetLastError());
}

// variant 0 / 1 / 2 = half / default / double group count (complex64, N >= 64)
template <typename R, int N, int LAYOUT, int MODE>
static int launch_n(void* data, const PassParams& p, long outer, int variant, hipStream_t s) {
  constexpr int N2 = split_of<N>::n2;
  constexpr int G0 = pick_groups<R, N, N2, LAYOUT>();
  if constexpr (sizeof(R) == 4 && N >= 64) {
    constexpr int GS = groups_ok<R, N, N2, LAYOUT>(G0 / 2) ? G0 / 2 : G0;
    constexpr int GL = groups_ok<R, N, N2, LAYOUT>(G0 * 2) ? G0 * 2 : G0;
    if (variant == 0) return launch_g<R, N, LAYOUT, MODE, GS>(data, p, outer, s);
    if (variant == 2) return launch_g<R, N, LAYOUT, MODE, GL>(data, p, outer, s);
  }
  return launch_g<R, N, LAYOUT, MODE, G0>(data, p, outer, s);
}

template <typename R, int LAYOUT, int MODE>
static int launch_mode(int n, void* data, const PassParams& p, long outer, int variant, hipStream_t s) {
  switch (n) {
    case 8:   return launch_n<R, 8, LAYOUT, MODE>(data, p, outer, variant, s);
    case 16:  return launch_n<R, 16, LAYOUT, MODE>(data, p, outer, variant, s);
    case 32:  return launch_n<R, 32, LAYOUT, MODE>(data, p, outer, variant, s);
    case 64:  return launch_n<R, 64, LAYOUT, MODE>(data, p, outer, variant, s);
    case 96:  return launch_n<R, 96, LAYOUT, MODE>(data, p, outer, variant, s);
    case 128: return launch_n<R, 128, LAYOUT, MODE>(data, p, outer, variant, s);
    case 192: return launch_n<R, 192, LAYOUT, MODE>(data, p, outer, variant, s);
    case 256: return launch_n<R, 256, LAYOUT, MODE>(data, p, outer, variant, s);
    case 512: return launch_n<R, 512, LAYOUT, MODE>(data, p, outer, variant, s);
    default:
      set_last_error("unsupported FFT length %d (supported: 8,16,32,64,96,128,192,256,512)", n);
      return WFS_ERR_UNSUPPORTED;
  }
}

template <typename R>
int launch_fft_pass_t(int n, int layout, int mode, void* data, const PassParams& p, long outer, int variant,
                      hipStream_t s) {
  if (layout == LAYOUT_ROWS) {
    if (mode == MODE_FWD) return launch_mode<R, LAYOUT_ROWS, MODE_FWD>(n, data, p, outer, variant, s);
    if (mode == MODE_INV) return launch_mode<R, LAYOUT_ROWS, MODE_INV>(n, data, p, outer, variant, s);
    if (mode == MODE_IVF) return launch_mode<R, LAYOUT_ROWS, MODE_IVF>(n, data, p, outer, variant, s);
  } else {
    if (mode == MODE_FWD) return launch_mode<R, LAYOUT_COLS, MODE_FWD>(n, data, p, outer, variant, s);
    if (mode == MODE_INV) return launch_mode<R, LAYOUT_COLS, MODE_INV>(n, data, p, outer, variant, s);
    if (mode == MODE_FKI) return launch_mode<R, LAYOUT_COLS, MODE_FKI>(n, data, p, outer, variant, s);
  }
  set_last_error("invalid FFT pass layout/mode %d/%d", layout, mode);
  return WFS_ERR_INVALID;
}

// This file is compiled twice: WFS_FFT_PREC=0 (complex64) and 1 (complex128).
#if WFS_FFT_PREC == 0
template int launch_fft_pass_t<float>(int, int, int, void*, const PassParams&, long, int, hipStream_t);

int launch_yz_plane_c128(int n, void* data, const YzPlaneParams& p, hipStream_t s);

// planes that fit in LDS: complex64 up to 128 x 128, complex128 up to 64 x 64
bool yz_plane_supported(int precision, int ny, int nz) {
  if (ny != nz) return false;
  if (precision == WFS_C128) return ny == 16 || ny == 32 || ny == 64;
  return ny == 16 || ny == 32 || ny == 64 || ny == 96 || ny == 128;
}

int launch_yz_plane(int precision, int n, void* data, const YzPlaneParams& p, hipStream_t s) {
  if (!yz_plane_supported(precision, n, n)) {
    set_last_error("yz plane pass: unsupported plane %d x %d", n, n);
    return WFS_ERR_UNSUPPORTED;
  }
  if (precision == WFS_C128) return launch_yz_plane_c128(n, data, p, s);
  switch (n) {
    case 16:  return launch_yz_plane_n<float, 16>(data, p, s);
    case 32:  return launch_yz_plane_n<float, 32>(data, p, s);
    case 64:  return launch_yz_plane_n<float, 64>(data, p, s);
    case 96:  return launch_yz_plane_n<float, 96>(data, p, s);
    default:  return launch_yz_plane_n<float, 128>(data, p, s);
  }
}

bool fft_length_supported(int n) {
  switch (n) {
    case 8: case 16: case 32: case 64: case 96: case 128: case 192: case 256: case 512: return true;
    default: return false;
  }
}

int launch_fft_pass(int precision, int n, int layout, int mode, void* data, const PassParams& p, long outer,
                    int variant, hipStream_t s) {
  if (precision == WFS_C128) return launch_fft_pass_t<double>(n, layout, mode, data, p, outer, variant, s);
  return launch_fft_pass_t<float>(n, layout, mode, data, p, outer, variant, s);
}
#else
template int launch_fft_pass_t<double>(int, int, int, void*, const PassParams&, long, int, hipStream_t);

int launch_yz_plane_c128(int n, void* data, const YzPlaneParams& p, hipStream_t s) {
  switch (n) {
    case 16:  return launch_yz_plane_n<double, 16>(data, p, s);
    case 32:  return launch_yz_plane_n<double, 32>(data, p, s);
    default:  return launch_yz_plane_n<double, 64>(data, p, s);
  }
}
#endif

}  // namespace wfs
