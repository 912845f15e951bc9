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
// N*N points, in registers and moves it through LDS:
//   y inverse -> transpose -> z inverse -> potential phase -> z forward
//   -> transpose -> y forward
// Each transform is column_dft with the split, table twiddles and operation
// order of the axis passes y_inv, z_inv_v_fwd and y_fwd, so the result is the
// same; the plane crosses HBM once (16 B per point for complex64) instead of
// three times (48 B per point).
// Thread t owns lines c0 + h N/H (h < H; columns z for the y transforms, rows y
// for the z transforms) as thread u = t / (N/H) of each, N1 points per line:
// the same mapping for both axes, so with N/H a multiple of 64 the table
// twiddles of both transforms are wavefront-uniform (scalar loads).  The
// exchanges and transpositions go through LDS in H rounds of N/H lines: the
// LDS holds N/H * (N + 1) points (the transposition, padded; the exchanges
// take N/H * N).  For N = 128 complex64 and H = 2: 66 KiB and 512 threads, two
// blocks per CU, so one block's loads and stores overlap the other's
// arithmetic.
// ---------------------------------------------------------------------------
template <typename R, int N, int H>
constexpr long yz_plane_lds_bytes() {
  return long(N / H) * (N + 1) * long(sizeof(cx<R>));
}

// An opaque copy of v: the compiler recomputes what derives from it instead of
// keeping values from an earlier stage alive (register pressure).
__device__ __forceinline__ int opaque_v(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

template <typename R, int N, int H>
__global__ __launch_bounds__(N * split_of<N>::n2 / H)
__attribute__((amdgpu_waves_per_eu(H == 1 ? 1 : 4))) void yz_plane_kernel(cx<R>* __restrict__ data,
                                                                           const YzPlaneParams p) {
  constexpr int N2 = split_of<N>::n2;
  constexpr int N1 = N / N2;
  constexpr int NH = N / H;                  // lines per round
  constexpr int N1H = N1 / H;                // points of a line per round of the transposition
  constexpr int TS = N + 1;                  // padded row stride of the transposition
  static_assert(N1 % H == 0 && N2 >= H, "bad plane split");
  constexpr bool UNIFORM = (NH % 64) == 0;   // u is the same across a wavefront
  using C = cx<R>;
  using CC = const __attribute__((address_space(4))) C;
  __shared__ C lds[NH * TS];

  const int tid = threadIdx.x;
  const long plane = blockIdx.x;
  C* __restrict__ blk = data + plane * long(N) * N;

  C a[H][N1];
  C w[N1];
  auto thread_u = [&]() -> int {
    if constexpr (UNIFORM) return __builtin_amdgcn_readfirstlane(tid / NH);
    else return opaque_v(tid) / NH;
  };
  auto load_w = [&](const void* twp, int u) {
    if constexpr (UNIFORM) {
      CC* tw = (CC*)(twp);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) w[k1] = mk<R>(tw[(u * k1) % N].x, tw[(u * k1) % N].y);
    } else {
      const C* __restrict__ tw = reinterpret_cast<const C*>(twp);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) w[k1] = tw[(u * k1) % N];
    }
  };
  // the exchange of column_dft for the lines of one round
  auto exchange = [&](C (&line)[N1], int u, auto sign_c) {
    const int c0 = opaque_v(tid) % NH;
    column_dft_second<N, N2, decltype(sign_c)::value>(line, lds, u,
                                                      [&](int k1, int n2) { return (k1 * N2 + n2) * NH + c0; });
  };
  // one transform of every line, round by round (one round's temporaries live at a time)
  auto transform = [&](const void* twp, auto sign_c) {
    const int u = thread_u();
#pragma unroll
    for (int h = 0; h < H; ++h) {
      if (h) __syncthreads();
      load_w(twp, u);
      column_dft_first<N, N2, decltype(sign_c)::value>(a[h], w);
      exchange(a[h], u, sign_c);
    }
  };
  // transpositions between the column (y) and row (z) lines, N/H rows a round
  auto cols_to_rows = [&]() {
    const int t = opaque_v(tid);
    const int c0 = t % NH, u = t / NH;
    C b[H][N1];
#pragma unroll
    for (int r = 0; r < H; ++r) {
      if (r) __syncthreads();
#pragma unroll
      for (int h = 0; h < H; ++h)
#pragma unroll
        for (int i = 0; i < N1H; ++i) lds[(u + N2 * i) * TS + c0 + h * NH] = a[h][r * N1H + i];
      __syncthreads();
#pragma unroll
      for (int i = 0; i < N1; ++i) b[r][i] = lds[c0 * TS + u + N2 * i];
    }
#pragma unroll
    for (int r = 0; r < H; ++r)
#pragma unroll
      for (int i = 0; i < N1; ++i) a[r][i] = b[r][i];
  };
  auto rows_to_cols = [&]() {
    const int t = opaque_v(tid);
    const int c0 = t % NH, u = t / NH;
    C b[H][N1];
#pragma unroll
    for (int r = 0; r < H; ++r) {
      if (r) __syncthreads();
#pragma unroll
      for (int i = 0; i < N1; ++i) lds[c0 * TS + u + N2 * i] = a[r][i];
      __syncthreads();
#pragma unroll
      for (int h = 0; h < H; ++h)
#pragma unroll
        for (int i = 0; i < N1H; ++i) b[h][r * N1H + i] = lds[(u + N2 * i) * TS + c0 + h * NH];
    }
#pragma unroll
    for (int h = 0; h < H; ++h)
#pragma unroll
      for (int i = 0; i < N1; ++i) a[h][i] = b[h][i];
  };

  {
    const int c0 = tid % NH, u = tid / NH;
#pragma unroll
    for (int h = 0; h < H; ++h)
#pragma unroll
      for (int i = 0; i < N1; ++i) a[h][i] = blk[unsigned(c0 + h * NH + (u + N2 * i) * N)];
  }

  // y inverse on columns
  transform(p.tw_yi, std::integral_constant<int, +1>{});
  __syncthreads();
  cols_to_rows();
  __syncthreads();

  // z inverse, potential phase and z forward on rows y = c0 + r N/H (z = u + N2 i).  Each
  // row round is finished (exchange, potential, first forward step) before the next one's
  // exchange, so that only one round's potential factors are held.
  {
    const int u = thread_u();
#pragma unroll
    for (int r = 0; r < H; ++r) {
      if (r) __syncthreads();
      load_w(p.tw_z, u);
      column_dft_first<N, N2, +1>(a[r], w);
      exchange(a[r], u, std::integral_constant<int, +1>{});
      const int c0 = opaque_v(tid) % NH;
      if (p.spx) {
        const long b = plane / p.nx;
        const C* __restrict__ px = reinterpret_cast<const C*>(p.spx);
        const C* __restrict__ py = reinterpret_cast<const C*>(p.spy) + b * N;
        const C* __restrict__ pzb = reinterpret_cast<const C*>(p.spz) + b * N;
        const C pxy = cmul(px[plane], py[c0 + r * NH]);
        const int ur = opaque_v(u);     // separate loads per round (not held across rounds)
#pragma unroll
        for (int i = 0; i < N1; ++i) {
          a[r][i] = cmul(a[r][i], cmul(pzb[ur + N2 * i], pxy));
          if (H > 1 && i % 4 == 3) asm volatile("" ::: "memory");   // at most 4 factors in flight
        }
      } else {
        const C* __restrict__ pvb = reinterpret_cast<const C*>(p.pv) + plane * long(N) * N;
#pragma unroll
        for (int i = 0; i < N1; ++i) {
          a[r][i] = cmul(a[r][i], pvb[unsigned((c0 + r * NH) * N + u + N2 * i)]);
          if (H > 1 && i % 4 == 3) asm volatile("" ::: "memory");
        }
      }
      load_w(p.tw_z, u);
      column_dft_first<N, N2, -1>(a[r], w);
    }
#pragma unroll
    for (int r = 0; r < H; ++r) {
      __syncthreads();
      exchange(a[r], u, std::integral_constant<int, -1>{});
    }
  }
  __syncthreads();
  rows_to_cols();
  __syncthreads();

  // y forward on columns
  transform(p.tw_yf, std::integral_constant<int, -1>{});
  {
    const int t = opaque_v(tid);
    const int c0 = t % NH, u = t / NH;
#pragma unroll
    for (int h = 0; h < H; ++h)
#pragma unroll
      for (int i = 0; i < N1; ++i) blk[unsigned(c0 + h * NH + (u + N2 * i) * N)] = a[h][i];
  }
}

template <typename R, int N, int H>
static int launch_yz_plane_n(void* data, const YzPlaneParams& p, hipStream_t s) {
  static_assert(yz_plane_lds_bytes<R, N, H>() <= 160 * 1024, "plane does not fit in LDS");
  if (p.planes <= 0 || p.planes > 0x7fffffffL) {
    set_last_error("yz plane pass: bad plane count %ld", p.planes);
    return WFS_ERR_INVALID;
  }
  hipLaunchKernelGGL((yz_plane_kernel<R, N, H>), dim3(unsigned(p.planes)), dim3(N * split_of<N>::n2 / H), 0, s,
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

int launch_yz_plane_c128(int n, int halves, void* data, const YzPlaneParams& p, hipStream_t s);

// planes that fit in LDS: complex64 up to 128 x 128, complex128 up to 64 x 64; two rounds
// (halves = 2) for complex64 64 and 128 and complex128 64
bool yz_plane_supported(int precision, int ny, int nz, int halves) {
  if (ny != nz || halves < 1 || halves > 2) return false;
  if (precision == WFS_C128) return (ny == 16 || ny == 32 || ny == 64) && (halves == 1 || ny == 64);
  return (ny == 16 || ny == 32 || ny == 64 || ny == 96 || ny == 128) && (halves == 1 || ny == 64 || ny == 128);
}

int launch_yz_plane(int precision, int n, int halves, void* data, const YzPlaneParams& p, hipStream_t s) {
  if (!yz_plane_supported(precision, n, n, halves)) {
    set_last_error("yz plane pass: unsupported plane %d x %d (%d rounds)", n, n, halves);
    return WFS_ERR_UNSUPPORTED;
  }
  if (precision == WFS_C128) return launch_yz_plane_c128(n, halves, data, p, s);
  switch (n) {
    case 16:  return launch_yz_plane_n<float, 16, 1>(data, p, s);
    case 32:  return launch_yz_plane_n<float, 32, 1>(data, p, s);
    case 64:  return halves == 2 ? launch_yz_plane_n<float, 64, 2>(data, p, s) : launch_yz_plane_n<float, 64, 1>(data, p, s);
    case 96:  return launch_yz_plane_n<float, 96, 1>(data, p, s);
    default:  return halves == 2 ? launch_yz_plane_n<float, 128, 2>(data, p, s) : launch_yz_plane_n<float, 128, 1>(data, p, s);
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

int launch_yz_plane_c128(int n, int halves, void* data, const YzPlaneParams& p, hipStream_t s) {
  switch (n) {
    case 16:  return launch_yz_plane_n<double, 16, 1>(data, p, s);
    case 32:  return launch_yz_plane_n<double, 32, 1>(data, p, s);
    default:  return halves == 2 ? launch_yz_plane_n<double, 64, 2>(data, p, s) : launch_yz_plane_n<double, 64, 1>(data, p, s);
  }
}
#endif

}  // namespace wfs
