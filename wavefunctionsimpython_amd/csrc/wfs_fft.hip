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
#include "wfs_yz_lines.h"

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

// Fused y-z pass on whole x planes (ny = nz = N): yz_lines_kernel with one residue class
// (B = 1) and LY = N, i.e. the plane's y inverse, z inverse, potential phase, z forward and
// y forward in one launch, bit for bit the three axis passes.  The plane crosses HBM once
// (16 B per point for complex64) instead of three times (48 B per point).  For N = 128
// complex64 and H = 2: 66 KiB of LDS and 512 threads, two blocks per CU, so one block's
// loads and stores overlap the other's arithmetic.
template <typename R, int N, int H>
static int launch_yz_plane_n(void* data, const YzPlaneParams& p, hipStream_t s) {
  static_assert(yz_lines_lds_bytes<R, N, N, H>() <= 160 * 1024, "plane does not fit in LDS");
  if (p.planes <= 0 || p.planes > 0x7fffffffL) {
    set_last_error("yz plane pass: bad plane count %ld", p.planes);
    return WFS_ERR_INVALID;
  }
  WFS_LAUNCH((yz_lines_kernel<R, N, N, 1, H, LINES_FULL>), dim3(unsigned(p.planes)),
                     dim3(yz_lines_threads<N, N, H>()), 0, s, reinterpret_cast<cx<R>*>(data), p);
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
  WFS_LAUNCH((fft_pass_kernel<R, N, N2, LAYOUT, MODE, G>), grid, dim3(N2 * G), 0, s,
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
