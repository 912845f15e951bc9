// Two-launch split step for complex64 grids with ny = nz in {128, 192, 256} and nx <= 256
// (gfx950).
//
// A 256^2 y-z plane (512 KiB) or a 192^2 one (288 KiB) does not fit in one CU, so the y
// transform is split in two stages, ny = LY * B (B = 4, LY = 64 or 48; at 128, B = 2 and
// LY = 64), y = B a + b, ky = ka + LY kb:
//   X[ka + LY kb] = sum_b W_B^{b kb} W_NY^{b ka} sum_a x[B a + b] W_LY^{a ka}
// and each stage rides on a pass that is needed anyway:
//
//   yz_lines_kernel<NZ, LY, B>  (one block per x plane and residue b: the LY rows
//     y = B a + b with all z, 128 KiB at 256, 72 KiB at 192, 66 KiB at 128):  length-LY
//     inverse y stage over a, z inverse, potential phase, z forward, length-LY forward y stage
//   x_slab_kernel  (one block per packet, ka and 8 or 16 consecutive z: the B rows
//     y = B ka + b with all x):  twiddles W_NY^{b ka} and length-B forward y stage over b,
//     x forward, kinetic phase K(kx, ky, kz), x inverse, length-B inverse y stage and twiddles
//
// Between the passes the y axis holds the half-transformed Y[b][ka] at row B ka + b.  The
// y-z launch reads the potential phase table in lines order (LinesOrder), one contiguous run
// per wavefront load.  A step moves 16 B per point in each pass (+8 for a potential table)
// against 64 (+8) for the four axis passes.  The arithmetic differs from the axis passes only
// in how the y transform is factored, so the results agree with theirs to float rounding, not
// bit for bit.
#include "wfs_yz_lines.h"

#include <algorithm>

namespace wfs {

namespace {

// rows y = B a + b of a row group: the length-B stage rides on the x launch (B = 4 at 256 and
// 192, LY = 64 and 48 rows per group; B = 2 at 128, LY = 64)
template <int N> struct SlabShape {
  static constexpr int B = N == 128 ? 2 : 4;
  static constexpr int LY = N / B;
};

// Lane pairs: the partner lanes tid and tid ^ ZB share the (x, z) points of a block; lane
// q = (tid / ZB) % 2 holds rows b = q and q + 2 (and, after the forward y stage, ky
// columns kb = q and q + 2).  The length-4 y stage is a 2-point stage in each lane, one
// complex value swapped with the partner, and a second 2-point stage:
//   forward, Z[b] = W_NY^{b ka} Y[b]:  E_q[k'] = Z[q] + (-1)^k' Z[q + 2],
//            X[k' + 2 k''] = E_0[k'] + (-1)^k'' W_4^{k'} E_1[k']
//   inverse: F_q[beta] = X[q] + (-1)^beta X[q + 2],
//            out[b] = F_0[b mod 2] + W_4^{-b} F_1[b mod 2],  Y'[b] = W_NY^{-b ka} out[b]
// Lane q keeps the k' = q (beta = q) values and receives the partner's.  For B = 2 (128^2
// planes) each lane holds the one row b = q and the length-2 stage is one swap:
//   X[kb] = Z[0] + (-1)^kb Z[1],  out[b] = X[0] + (-1)^b X[1].
// Lane-dependent choices are signs, not selects: with s = (-1)^q, lane q of a B = 4 pair
// holds s times its second column through the x transforms (all linear) and uses s W_NY^{b ka}
// as that row's twiddle.  Negation commutes exactly with every rounded operation, so the
// results are bit for bit those of the selecting form, with one select per value left.
template <typename R>
__device__ __forceinline__ cx<R> shfl_xor_c(cx<R> v, int m) {
  return mk<R>(__shfl_xor(v.x, m, 64), __shfl_xor(v.y, m, 64));
}

// the partner's value, partner = tid ^ ZB: for ZB = 8 the partner sits in the same row of 16
// lanes, 8 along, so a DPP row rotation by 8 (row_ror:8, one VALU move per part) replaces the
// LDS-unit permute; for ZB = 16 it sits in the neighbouring row, and gfx950's row swap
// (v_permlane16_swap: odd rows of the first operand with even rows of the second, both the
// value itself) leaves it in one of the two results, picked by the lane's row; other widths use
// the permute
template <int ZB, typename R>
__device__ __forceinline__ cx<R> partner_c(cx<R> v) {
  if constexpr (ZB == 8 && sizeof(R) == 4) {
    constexpr int kRowRor8 = 0x128;
    const int x = __builtin_amdgcn_update_dpp(0, __float_as_int(v.x), kRowRor8, 0xF, 0xF, false);
    const int y = __builtin_amdgcn_update_dpp(0, __float_as_int(v.y), kRowRor8, 0xF, 0xF, false);
    return mk<R>(__int_as_float(x), __int_as_float(y));
  } else if constexpr (ZB == 16 && sizeof(R) == 4) {
    const bool odd_row = threadIdx.x & 16;
    const auto sx = __builtin_amdgcn_permlane16_swap(__float_as_uint(v.x), __float_as_uint(v.x), false, false);
    const auto sy = __builtin_amdgcn_permlane16_swap(__float_as_uint(v.y), __float_as_uint(v.y), false, false);
    return mk<R>(__uint_as_float(odd_row ? sx[0] : sx[1]), __uint_as_float(odd_row ? sy[0] : sy[1]));
  } else {
    return shfl_xor_c<R>(v, ZB);
  }
}

// ZB = 8 runs four waves per SIMD where a lane holds at most 16 points of an x line (at most 128
// VGPRs: the x transforms' table twiddles are read from LDS one at a time where they are applied,
// so no twiddle array is held; at nx = 256 one spilled register, 52.4 against 55.8 us per launch
// with three waves, bit-identical); the 24 points of nx = 96 or 192 do not fit in 128.
// ZB = 8: a block reads and writes 64-byte halves of 128-byte lines; the two blocks of a line
// get the block indices 16 g + x and 16 g + 8 + x, which the dispatcher sends to the same XCD
// (blocks go round-robin over the 8 XCDs) one after the other, so the second half comes from
// that XCD's L2.
template <typename R, int NX, int NY, int NZ, int B, int ZB>
__global__ __launch_bounds__(2 * ZB * split_of<NX>::n2)
__attribute__((amdgpu_waves_per_eu(ZB == 8 && NX / split_of<NX>::n2 <= 16 ? 4 : 1))) void x_slab_kernel(
    cx<R>* __restrict__ data, const XSlabParams p) {
  constexpr int N2 = split_of<NX>::n2;
  constexpr int N1 = NX / N2;
  constexpr int A = NY / B;
  constexpr int TZ = NZ / ZB;
  static_assert(B == 4 || B == 2, "the lane-pair y stage is written for B = 4 and B = 2");
  constexpr int RPL = B / 2;               // rows (columns after the y stage) per lane
  static_assert(N1 * N2 == NX && N1 % N2 == 0 && NZ % ZB == 0 && NY % B == 0 && 2 * ZB <= 64, "bad slab split");
  using C = cx<R>;
  using CC = const __attribute__((address_space(4))) C;
  constexpr int LINES = 2 * ZB;            // x lines per round: (q, z)
  __shared__ C lds[LINES * NX];

  const int tid = threadIdx.x;
  const int zl = tid % ZB;                 // z within the block's segment
  const int q = (tid / ZB) % 2;            // lane-pair half
  const int t = tid / (2 * ZB);            // thread of the x line: x = t + N2 i
  const int line = q * ZB + zl;
  long bid = blockIdx.x;
  if constexpr (ZB == 8) bid = 2 * (8 * (bid / 16) + bid % 8) + (bid / 8) % 2;
  const int zt = int(bid % TZ);
  const int ka = int((bid / TZ) % A);
  const long pk = bid / (long(TZ) * A);
  C* __restrict__ blk = data + pk * (long(NX) * NY * NZ) + long(B * ka) * NZ + long(zt) * ZB;
  constexpr unsigned XS = unsigned(NY) * NZ;   // x stride
  constexpr unsigned EB = unsigned(sizeof(C));  // byte offsets within a packet: NX NY NZ 8 <= 2^27
  static_assert(long(NX) * NY * NZ * sizeof(C) <= (1L << 32), "32-bit byte offsets");
  // element i of row q + 2 j
  auto off = [&](int i, int j) { return unsigned(zl) + unsigned(q + 2 * j) * NZ + unsigned(t + N2 * i) * XS; };

  // y-stage twiddles W_NY^{b ka} of this lane's rows b = q (+ 2)
  C wq[RPL];
  {
    CC* twy = (CC*)(p.tw_y);
    C wb[B];
#pragma unroll
    for (int b = 0; b < B; ++b) wb[b] = mk<R>(twy[(b * ka) % NY].x, twy[(b * ka) % NY].y);
    wq[0] = q ? wb[1] : wb[0];
    if constexpr (RPL == 2) wq[RPL - 1] = q ? mk<R>(-wb[B - 1].x, -wb[B - 1].y) : wb[B - 2];   // s W_NY^{b ka}
  }

  C a[RPL][N1];
#pragma unroll
  for (int j = 0; j < RPL; ++j)
#pragma unroll
    for (int i = 0; i < N1; ++i) a[j][i] = byte_at(blk, off(i, j) * EB);

  // the x transforms' table twiddles W_NX^{t k1} (t < N2, k1 < N1) and the kinetic factors
  // kx(x), staged in LDS once per block (their loads in flight with the points'); every
  // transform of both columns reads them from there
  __shared__ C tws[N2 * N1];
  __shared__ C kxs[NX];
  {
    const C* __restrict__ tw = reinterpret_cast<const C*>(p.tw_x);
    const C* __restrict__ kx = reinterpret_cast<const C*>(p.kx);
    for (int e = tid; e < N2 * N1; e += 2 * ZB * N2) tws[e] = tw[((e / N1) * (e % N1)) % NX];
    for (int e = tid; e < NX; e += 2 * ZB * N2) kxs[e] = kx[e];
  }

  const R sgn = q ? R(-1) : R(1);          // s = (-1)^q
  if constexpr (B == 2) {
    // forward y stage, B = 2: X[kb] = Z[0] + (-1)^kb Z[1], lane q keeps X[q] = recv + s z
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      const C z = cmul(a[0][i], wq[0]);
      const C recv = partner_c<ZB, R>(z);
      a[0][i] = mk<R>(fma_r(sgn, z.x, recv.x), fma_r(sgn, z.y, recv.y));
    }
  } else {
    // forward y stage, B = 4: a[0] <- X[q], a[1] <- s X[q + 2].  With z1 = s Z[q + 2] the
    // kept value z0 + z1 is E_q[q] and the sent one E_q[1 - q]; lane 0 then holds E_0[0]
    // and receives E_1[0], lane 1 holds E_1[1] and receives E_0[1]:
    //   X[q] = kk + recv,  X[q + 2] = s (kk - recv),  kk = keep or W_4 keep (lane 1)
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      const C z0 = cmul(a[0][i], wq[0]);
      const C z1 = cmul(a[RPL - 1][i], wq[RPL - 1]);
      const C keep = cadd(z0, z1);
      const C recv = partner_c<ZB, R>(csub(z0, z1));
      const C kk = q ? mk<R>(keep.y, -keep.x) : keep;     // W_4^q keep, W_4 = -i
      a[0][i] = cadd(kk, recv);
      a[RPL - 1][i] = csub(kk, recv);
    }
  }

  // x forward, kinetic phase, x inverse on the columns ky = ka + A kb, kb = q + 2 j, one
  // column at a time: the table twiddles and kinetic factors are read where they are used
  // (opaque indices), so that only the other column's points stay live besides the transform
  auto lidx = [&](int k1, int n2) { return (k1 * N2 + n2) * LINES + line; };
  const C* __restrict__ ky = reinterpret_cast<const C*>(p.ky);
  const C* __restrict__ kz = reinterpret_cast<const C*>(p.kz);
  __syncthreads();                         // tws, kxs staged
  // one column transform: the in-register DFT, then each table twiddle read from LDS where it
  // is applied (no twiddle array held across the DFT), then the LDS exchange
  auto column = [&](C (&col)[N1], auto sign_c) {
    constexpr int SG = decltype(sign_c)::value;
    DFT<N1, SG>::run(col);
    const C* tr = tws + opaque_v(t) * N1;
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) {
      const C w = tr[k1];
      col[k1] = SG < 0 ? cmul(col[k1], w) : cmulc(col[k1], w);
    }
    column_dft_second<NX, N2, SG>(col, lds, t, lidx);
  };
#pragma unroll
  for (int j = 0; j < RPL; ++j) {
    if (j) __syncthreads();
    column(a[j], std::integral_constant<int, -1>{});
    {
      const C* kr = kxs + opaque_v(t);
      const C kyz = cmul(ky[ka + A * (q + 2 * j)], kz[zt * ZB + zl]);
#pragma unroll
      for (int i = 0; i < N1; ++i) a[j][i] = cmul(a[j][i], cmul(kr[N2 * i], kyz));
    }
    __syncthreads();
    column(a[j], std::integral_constant<int, +1>{});
  }

  if constexpr (B == 2) {
    // inverse y stage, B = 2: out[b] = X[0] + (-1)^b X[1], Y'[q] = W_NY^{-q ka} out[q]
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      const C v = a[0][i];
      const C recv = partner_c<ZB, R>(v);
      a[0][i] = cmulc(mk<R>(fma_r(sgn, v.x, recv.x), fma_r(sgn, v.y, recv.y)), wq[0]);
    }
  } else {
    // inverse y stage, B = 4, from a[0] = X[q] and a[1] = s X[q + 2]: the kept a[0] + a[1] is
    // F_q[q] and the sent a[0] - a[1] is F_q[1 - q]; with kk = keep or W_4^{-1} keep (lane 1)
    //   out[q] = kk + recv,  out[q + 2] = s (kk - recv),  a[0] <- Y'[q], a[1] <- Y'[q + 2]
    // (s W_NY^{(q+2) ka} is the second row's twiddle, so the sign cancels)
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      const C keep = cadd(a[0][i], a[RPL - 1][i]);
      const C recv = partner_c<ZB, R>(csub(a[0][i], a[RPL - 1][i]));
      const C kk = q ? mk<R>(-keep.y, keep.x) : keep;     // W_4^{-q} keep, W_4^{-1} = +i
      a[0][i] = cmulc(cadd(kk, recv), wq[0]);
      a[RPL - 1][i] = cmulc(csub(kk, recv), wq[RPL - 1]);
    }
  }

  // store addresses recomputed from an opaque copy of the thread index, so the load
  // addresses do not stay live through the transforms
  {
    const int tt = opaque_v(tid);
    const unsigned base = unsigned(tt % ZB) + unsigned((tt / ZB) % 2) * NZ + unsigned(tt / (2 * ZB)) * XS;
#pragma unroll
    for (int j = 0; j < RPL; ++j)
#pragma unroll
      for (int i = 0; i < N1; ++i) byte_at(blk, (base + unsigned(2 * j) * NZ + unsigned(N2 * i) * XS) * EB) = a[j][i];
  }
}

// flags[p] = 1 when x plane p equals plane p - 1 of the same packet bit for bit, else 0
__global__ __launch_bounds__(256) void plane_equal_kernel(const unsigned long long* __restrict__ table,
                                                          long plane_elems, int nx, long first,
                                                          int* __restrict__ flags) {
  const long p = first + blockIdx.x;
  int differ = 0;
  if (p % nx == 0) {
    differ = 1;
  } else {
    const unsigned long long* a = table + p * plane_elems;
    const unsigned long long* b = a - plane_elems;
    // 16 independent pairs of loads per round, issued together (no bound test between them;
    // the early exit is tested once a round): a 256^2 plane takes 16 rounds, not 256 dependent
    // steps.  The last partial round goes one element at a time.
    constexpr int U = 16;
    const long full = plane_elems - 256L * (U - 1);   // rounds starting below this hold U elements
    long i0 = threadIdx.x;
    for (; i0 < full && !differ; i0 += 256L * U) {
      unsigned long long v[U], w[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        v[u] = a[i0 + 256L * u];
        w[u] = b[i0 + 256L * u];
      }
      unsigned long long d = 0;
#pragma unroll
      for (int u = 0; u < U; ++u) d |= v[u] ^ w[u];
      differ = d != 0;
    }
    for (; i0 < plane_elems && !differ; i0 += 256) differ = a[i0] != b[i0];
  }
  differ = __syncthreads_or(differ);
  if (threadIdx.x == 0) flags[p] = differ ? 0 : 1;
}

// src[p] = the first plane of p's run of identical planes, one block per packet: plane x of the
// packet maps to the last x' <= x whose flag is 0 (plane 0 of a packet always starts a run).
// Each thread takes a chunk of planes, the chunks' last run starts are max-scanned across the
// block, and each thread fills its chunk from the start carried in -- instead of one thread
// walking the packet's planes one dependent load at a time.
__global__ __launch_bounds__(256) void plane_runs_kernel(const int* __restrict__ flags, int nx,
                                                         int* __restrict__ src) {
  __shared__ int last[256];
  const long base = long(blockIdx.x) * nx;
  const int t = threadIdx.x;
  const int chunk = (nx + 255) / 256;
  const int x0 = t * chunk < nx ? t * chunk : nx;
  const int x1 = x0 + chunk < nx ? x0 + chunk : nx;
  int l = -1;
  for (int x = x0; x < x1; ++x)
    if (!flags[base + x]) l = x;
  last[t] = l;
  __syncthreads();
  for (int d = 1; d < 256; d <<= 1) {   // inclusive max-scan over the chunks
    const int v = t >= d ? last[t - d] : -1;
    __syncthreads();
    if (v > last[t]) last[t] = v;
    __syncthreads();
  }
  int cur = t > 0 ? last[t - 1] : -1;
  for (int x = x0; x < x1; ++x) {
    if (!flags[base + x]) cur = x;
    src[base + x] = int(base) + cur;
  }
}

// compute units of the current device (cached per device)
int cu_count() {
  static int cache[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cache[dev] > 0) return cache[dev];
  int n = 0;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
  cache[dev] = n;
  return n;
}

// persistent: one block per CU walking the row groups, else one block per group
template <int N, int MODE, bool PERSIST>
int launch_lines_m(void* data, const YzPlaneParams& p, hipStream_t s) {
  constexpr int LY = SlabShape<N>::LY, B = SlabShape<N>::B;
  static_assert(yz_lines_lds_bytes<float, N, LY, 1>() <= 160 * 1024, "rows do not fit in LDS");
  const long groups = p.planes * B;
  const long grid = PERSIST ? std::min<long>(groups, cu_count()) : groups;
  WFS_LAUNCH((yz_lines_kernel<float, N, LY, B, 1, MODE, PERSIST, true>), dim3(unsigned(grid)),
                     dim3(yz_lines_threads<N, LY, 1>()), 0, s, reinterpret_cast<float2*>(data), p);
  return int(hipGetLastError());
}

template <int N>
int launch_lines_n(int mode, int variant, void* data, const YzPlaneParams& p, hipStream_t s) {
  const bool persist = variant == 2;
  if (mode == LINES_FULL)
    return persist ? launch_lines_m<N, LINES_FULL, true>(data, p, s) : launch_lines_m<N, LINES_FULL, false>(data, p, s);
  if (mode == LINES_ENTRY)
    return persist ? launch_lines_m<N, LINES_ENTRY, true>(data, p, s) : launch_lines_m<N, LINES_ENTRY, false>(data, p, s);
  if (mode == LINES_EXIT)
    return persist ? launch_lines_m<N, LINES_EXIT, true>(data, p, s) : launch_lines_m<N, LINES_EXIT, false>(data, p, s);
  if (mode == LINES_OBSERVE) return launch_lines_m<N, LINES_OBSERVE, false>(data, p, s);
  if (mode == LINES_ENTRY_OBSERVE) return launch_lines_m<N, LINES_ENTRY_OBSERVE, false>(data, p, s);
  set_last_error("slab y-z pass: bad mode %d", mode);
  return WFS_ERR_INVALID;
}

template <int NX, int N, int ZB>
int launch_x_slab_z(void* data, const XSlabParams& p, hipStream_t s) {
  constexpr int B = SlabShape<N>::B;
  constexpr long per_packet = long(N / B) * (N / ZB);
  static_assert(per_packet % 16 == 0, "the XCD pairing needs a multiple of 16 blocks");
  WFS_LAUNCH((x_slab_kernel<float, NX, N, N, B, ZB>), dim3(unsigned(p.packets * per_packet)),
                     dim3(2 * ZB * split_of<NX>::n2), 0, s, reinterpret_cast<float2*>(data), p);
  return int(hipGetLastError());
}

// variant 2: 8 z points per block (256 threads at nx = 256), else 16
template <int N>
int launch_x_slab_ny(int nx, int variant, void* data, const XSlabParams& p, hipStream_t s) {
#define WFS_X_SLAB_CASE(NX)                                                                              \
  case NX:                                                                                               \
    return variant == 2 ? launch_x_slab_z<NX, N, 8>(data, p, s) : launch_x_slab_z<NX, N, 16>(data, p, s);
  switch (nx) {
    WFS_X_SLAB_CASE(8)
    WFS_X_SLAB_CASE(16)
    WFS_X_SLAB_CASE(32)
    WFS_X_SLAB_CASE(64)
    WFS_X_SLAB_CASE(96)
    WFS_X_SLAB_CASE(128)
    WFS_X_SLAB_CASE(192)
    WFS_X_SLAB_CASE(256)
    default:
      set_last_error("slab x pass: unsupported length %d", nx);
      return WFS_ERR_UNSUPPORTED;
  }
#undef WFS_X_SLAB_CASE
}

}  // namespace

int launch_plane_dedup(const void* table, long plane_elems, int nx, int packets, long first, long count, int* flags,
                       int* src, hipStream_t s) {
  const long planes = long(nx) * packets;
  if (planes < 1 || planes > 0x7fffffffL || plane_elems < 1 || first < 0 || count < 1 || first + count > planes) {
    set_last_error("potential plane sharing: bad sizes");
    return WFS_ERR_INVALID;
  }
  WFS_LAUNCH(plane_equal_kernel, dim3(unsigned(count)), dim3(256), 0, s,
                     static_cast<const unsigned long long*>(table), plane_elems, nx, first, flags);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return int(e);
  WFS_LAUNCH(plane_runs_kernel, dim3(unsigned(packets)), dim3(256), 0, s, flags, nx, src);
  return int(hipGetLastError());
}

// complex64, ny = nz in {128, 192, 256}, nx a supported FFT length up to 256 (at 512 the x
// lines of a lane pair do not fit in registers)
bool slab_supported(int precision, const int n[3]) {
  return precision == WFS_C64 && n[1] == n[2] && (n[1] == 128 || n[1] == 192 || n[1] == 256) && n[0] <= 256 &&
         fft_length_supported(n[0]);
}

template <int N>
static SlabRows slab_rows_n() {
  SlabRows g;
  g.b = SlabShape<N>::B;
  g.n1 = N / split_of<N>::n2;
  g.uy = SlabShape<N>::LY / g.n1;
  g.uz = split_of<N>::n2;
  return g;
}

SlabRows slab_rows(int n) {
  return n == 128 ? slab_rows_n<128>() : n == 192 ? slab_rows_n<192>() : slab_rows_n<256>();
}

template <int N>
static LinesOrder slab_lines_order_n() {
  constexpr int LY = SlabShape<N>::LY;
  LinesOrder m;
  m.ny = N;
  m.nz = N;
  m.ly = LY;
  m.b = SlabShape<N>::B;
  m.rh = LY;
  m.n2z = split_of<N>::n2;
  m.n1 = N / split_of<N>::n2;
  m.threads = yz_lines_threads<N, LY, 1>();
  return m;
}

LinesOrder slab_lines_order(int n) {
  return n == 128 ? slab_lines_order_n<128>() : n == 192 ? slab_lines_order_n<192>() : slab_lines_order_n<256>();
}

int launch_yz_slab(int n, int mode, int variant, void* data, const YzPlaneParams& p, hipStream_t s) {
  if (p.planes <= 0 || p.planes * 4 > 0x7fffffffL) {
    set_last_error("slab y-z pass: bad plane count %ld", p.planes);
    return WFS_ERR_INVALID;
  }
  if (n == 256) return launch_lines_n<256>(mode, variant, data, p, s);
  if (n == 192) return launch_lines_n<192>(mode, variant, data, p, s);
  if (n == 128) return launch_lines_n<128>(mode, variant, data, p, s);
  set_last_error("slab y-z pass: unsupported plane %d", n);
  return WFS_ERR_UNSUPPORTED;
}

int launch_x_slab(int nx, int n, int variant, void* data, const XSlabParams& p, hipStream_t s) {
  const long blocks = p.packets * long(n / 2) * (n / 8);
  if (p.packets <= 0 || blocks > 0x7fffffffL) {
    set_last_error("slab x pass: bad packet count %ld", p.packets);
    return WFS_ERR_INVALID;
  }
  if (n == 256) return launch_x_slab_ny<256>(nx, variant, data, p, s);
  if (n == 192) return launch_x_slab_ny<192>(nx, variant, data, p, s);
  if (n == 128) return launch_x_slab_ny<128>(nx, variant, data, p, s);
  set_last_error("slab x pass: unsupported plane %d", n);
  return WFS_ERR_UNSUPPORTED;
}

}  // namespace wfs
