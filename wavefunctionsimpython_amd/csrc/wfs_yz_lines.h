// Fused y-z work of the split step on a group of y lines held by one block (gfx950).
//
// A block owns LY rows y = B a + b (a < LY, b = this block's residue) of one x plane, each
// row with all NZ points, in registers: N1 points per thread for both line directions
// (NZ = N1 N2Z along z, LY = N1 N2Y along y).  The y lines (LY points each, one per z
// column) and the z lines (NZ points each, one per row) are transformed with column_dft,
// and the switch between them goes through a padded LDS transposition in H rounds.
//
// B = 1 and LY = NZ is the whole y-z plane: the full step
//   y inverse -> transpose -> z inverse -> potential phase -> z forward -> transpose -> y forward
// with the split, table twiddles and operation order of the axis passes, so its result is
// that of the passes y_inv, z_inv_v_fwd and y_fwd.
//
// B > 1 splits the y transform in two stages, NY = LY B, y = B a + b, ky = ka + LY kb:
//   X[ka + LY kb] = sum_b W_B^{b kb} W_NY^{b ka} sum_a x[B a + b] W_LY^{a ka}
// The block does the inner length-LY transforms over a (the forward one leaves
// Y[b][ka] = sum_a x[B a + b] W_LY^{a ka} at row B ka + b); the x pass of the slab scheme
// (x_slab_kernel) does the twiddles and the length-B transforms over b.  The inverse runs
// the same stages backwards.
//
// Thread t owns y lines (columns z = c0 + h NZ/H) as thread u = t / (NZ/H) of each, and z
// lines (rows c0 + r LY/H) as thread u = t / (LY/H); for NZ/H (LY/H) a multiple of 64 the
// table twiddles of the y (z) transforms are wavefront-uniform and come in by scalar loads.
// The LDS holds LY/H rows of NZ + 1 points.
#pragma once

#include "wfs_internal.h"

namespace wfs {

// split N = N1 * N2 of a length-N line (N2 threads per line, N1 points per thread)
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

// An opaque copy of v: the compiler recomputes what derives from it instead of
// keeping values from an earlier stage alive (register pressure).
__device__ __forceinline__ int opaque_v(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

__device__ __forceinline__ double lines_simpson(int i, int n) {
  if (i == 0 || i == n - 1) return 1.0;
  return (i & 1) ? 4.0 : 2.0;
}


// Box moments of the position-space row a thread holds in the z-line mapping of a
// yz_lines_kernel round (row c0 = tid mod LY of the group, z = u + N2Z i): per box
//   sum w |psi|^2 [1, x, y, z, x^2, y^2, z^2],  w = s(ix) s(iy) s(iz)  (Simpson 1-4-2-...-4-1)
// accumulated in FP64, reduced over the block and written as the group's record
// partials[group][box][7].  Same sums as box_moments_kernel (wfs_observables.hip).
template <typename R, int NZ, int LY, int B, int N2Z, int N1, int T>
__device__ void lines_moments(const cx<R> (&a)[N1], const LinesMoments& m, long plane, int bres, long group, int tid,
                              int u) {
  constexpr int NB = kBoxesPerPass;
  constexpr int kWaves = T / 64;
  __shared__ double red[kWaves][NB * 7];
  const int c0 = tid % LY;
  const int gx = int(plane % m.n[0]);
  const int gy = B * c0 + bres;
  const double rx = m.dx[0] * double(gx - m.n[0] / 2);
  const double ry = m.dx[1] * double(gy - m.n[1] / 2);
  const int lane = tid & 63, wave = tid >> 6;
  // one box at a time (few registers live besides the row's points)
#pragma unroll 1
  for (int b = 0; b < m.nb; ++b) {
    const int ix = gx - m.bot[b][0], iy = gy - m.bot[b][1];
    const bool row_in = ix >= 0 && ix < m.num[b][0] && iy >= 0 && iy < m.num[b][1];
    const double wxy = row_in ? lines_simpson(ix, m.num[b][0]) * lines_simpson(iy, m.num[b][1]) : 0.0;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    if (row_in) {
#pragma unroll
      for (int i = 0; i < N1; ++i) {
        // an opaque z index per point: the point's position and weight are formed where they
        // are used, not all sixteen ahead (FP64 registers)
        const int gz = opaque_v(u) + N2Z * i;
        const int iz = gz - m.bot[b][2];
        const double rho = double(a[i].x) * double(a[i].x) + double(a[i].y) * double(a[i].y);
        const double rz = m.dx[2] * double(gz - m.n[2] / 2);
        const double wr = (iz >= 0 && iz < m.num[b][2]) ? lines_simpson(iz, m.num[b][2]) * rho : 0.0;
        s0 += wr;
        s1 += wr * rz;
        s2 += wr * rz * rz;
      }
    }
    const double v[7] = {wxy * s0, wxy * rx * s0, wxy * ry * s0, wxy * s1, wxy * rx * rx * s0,
                         wxy * ry * ry * s0, wxy * s2};
    {
      const double t = wave_sum7(v, lane);
      const int q = (lane >> 3) & 7;
      if ((lane & 7) == 0 && q < 7) red[wave][b * 7 + q] = t;
    }
  }
  __syncthreads();
  if (tid < m.nb * 7) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) t += red[w][tid];
    m.partials[group * m.nb * 7 + tid] = t;
  }
}

template <int NZ, int LY, int H>
constexpr int yz_lines_threads() {
  return NZ * (LY / (NZ / split_of<NZ>::n2)) / H;
}

template <typename R, int NZ, int LY, int H>
constexpr long yz_lines_lds_bytes() {
  return long(LY / H) * (NZ + 1) * long(sizeof(cx<R>));
}

// PERSIST: a grid of at most one block per CU walks the groups blockIdx.x, + gridDim.x, ...
// and loads the next group's points while it transforms the current one (H = 1 only), so
// that one block per CU still overlaps its memory traffic with its arithmetic.
// PV_LINES: the potential table is in lines order (LinesOrder, wfs_internal.h), read as one
// contiguous run per wavefront, and the twiddle tables hold each thread's twiddles
// contiguously ([u][k1], u the thread's index within a line); else the potential table is
// in the state's C order and the twiddles come from the length-N axis tables.
template <typename R, int NZ, int LY, int B, int H, int MODE, bool PERSIST = false, bool PV_LINES = false>
__global__ __launch_bounds__((yz_lines_threads<NZ, LY, H>()))
__attribute__((amdgpu_waves_per_eu(H == 1 ? 1 : 4))) void yz_lines_kernel(cx<R>* __restrict__ data,
                                                                           const YzPlaneParams p) {
  constexpr int N2Z = split_of<NZ>::n2;
  constexpr int N1 = NZ / N2Z;               // points per thread and line, both directions
  constexpr int N2Y = LY / N1;
  static_assert(N1 * N2Y == LY && N1 % N2Y == 0, "bad y split");
  constexpr int NY = LY * B;
  constexpr int CH = NZ / H;                 // y lines (columns z) per round
  constexpr int RH = LY / H;                 // z lines (rows) per round
  constexpr int N1H = N1 / H;                // points of a y line per round of the transposition
  constexpr int TS = NZ + 1;                 // padded row stride of the transposition
  constexpr unsigned RS = unsigned(B) * NZ;  // element stride between the block's rows
  static_assert(N1 % H == 0 && N2Y >= H && N2Z >= H, "bad round split");
  constexpr bool UNIFORM_Y = (CH % 64) == 0;
  constexpr bool UNIFORM_Z = (RH % 64) == 0;
  using C = cx<R>;
  using CC = const __attribute__((address_space(4))) C;
  __shared__ C lds[RH * TS];

  static_assert(!PERSIST || H == 1, "the persistent form holds one round");
  const int tid = threadIdx.x;
  const long groups = p.planes * B;
  long group = blockIdx.x;
  long plane = group / B;                    // x plane, counted over all packets
  int bres = int(group % B);                 // residue b of the block's rows
  C* __restrict__ blk = data + plane * long(NY) * NZ + long(bres) * NZ;

  C a[H][N1];
  C w[N1];
  auto thread_u_y = [&]() -> int {
    if constexpr (UNIFORM_Y) return __builtin_amdgcn_readfirstlane(tid / CH);
    else return opaque_v(tid) / CH;
  };
  auto thread_u_z = [&]() -> int {
    if constexpr (UNIFORM_Z) return __builtin_amdgcn_readfirstlane(tid / RH);
    else return opaque_v(tid) / RH;
  };
  // twiddles W_L^{u k1} of a length-L line from the length-(M L) table: entry M u k1 mod (M L)
  auto load_w = [&](const void* twp, int u, auto uniform_c, auto m_c, auto len_c) {
    constexpr int M = decltype(m_c)::value, LEN = decltype(len_c)::value;
    if constexpr (PV_LINES && decltype(uniform_c)::value) {
      CC* tw = (CC*)(twp) + u * N1;
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) w[k1] = mk<R>(tw[k1].x, tw[k1].y);
    } else if constexpr (PV_LINES) {
      const C* __restrict__ tw = reinterpret_cast<const C*>(twp) + u * N1;
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) w[k1] = tw[k1];
    } else if constexpr (decltype(uniform_c)::value) {
      CC* tw = (CC*)(twp);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) w[k1] = mk<R>(tw[(M * u * k1) % LEN].x, tw[(M * u * k1) % LEN].y);
    } else {
      const C* __restrict__ tw = reinterpret_cast<const C*>(twp);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) w[k1] = tw[(M * u * k1) % LEN];
    }
  };
  auto load_wy = [&](const void* twp, int u) {
    load_w(twp, u, std::integral_constant<bool, UNIFORM_Y>{}, std::integral_constant<int, B>{},
           std::integral_constant<int, NY>{});
  };
  auto load_wz = [&](const void* twp, int u) {
    load_w(twp, u, std::integral_constant<bool, UNIFORM_Z>{}, std::integral_constant<int, 1>{},
           std::integral_constant<int, NZ>{});
  };
  // the exchanges of column_dft for the lines of one round
  auto exchange_y = [&](C (&line)[N1], int u, auto sign_c) {
    const int c0 = opaque_v(tid) % CH;
    column_dft_second<LY, N2Y, decltype(sign_c)::value>(line, lds, u,
                                                        [&](int k1, int n2) { return (k1 * N2Y + n2) * CH + c0; });
  };
  auto exchange_z = [&](C (&line)[N1], int u, auto sign_c) {
    const int c0 = opaque_v(tid) % RH;
    column_dft_second<NZ, N2Z, decltype(sign_c)::value>(line, lds, u,
                                                        [&](int k1, int n2) { return (k1 * N2Z + n2) * RH + c0; });
  };
  // one transform of every y line, round by round (one round's temporaries live at a time)
  auto transform_y = [&](const void* twp, auto sign_c) {
    const int u = thread_u_y();
#pragma unroll
    for (int h = 0; h < H; ++h) {
      if (h) __syncthreads();
      load_wy(twp, u);
      column_dft_first<LY, N2Y, decltype(sign_c)::value>(a[h], w);
      exchange_y(a[h], u, sign_c);
    }
  };
  // one transform of every z line (no potential), round by round
  auto transform_z = [&](const void* twp, auto sign_c) {
    const int u = thread_u_z();
#pragma unroll
    for (int r = 0; r < H; ++r) {
      if (r) __syncthreads();
      load_wz(twp, u);
      column_dft_first<NZ, N2Z, decltype(sign_c)::value>(a[r], w);
      exchange_z(a[r], u, sign_c);
    }
  };
  // transpositions between the y lines and the z lines, LY/H rows a round
  auto cols_to_rows = [&]() {
    const int t = opaque_v(tid);
    const int c0 = t % CH, u = t / CH;
    const int r0 = t % RH, v = t / RH;
    C bb[H][N1];
#pragma unroll
    for (int r = 0; r < H; ++r) {
      if (r) __syncthreads();
#pragma unroll
      for (int h = 0; h < H; ++h)
#pragma unroll
        for (int i = 0; i < N1H; ++i) lds[(u + N2Y * i) * TS + c0 + h * CH] = a[h][r * N1H + i];
      __syncthreads();
#pragma unroll
      for (int i = 0; i < N1; ++i) bb[r][i] = lds[r0 * TS + v + N2Z * i];
    }
#pragma unroll
    for (int r = 0; r < H; ++r)
#pragma unroll
      for (int i = 0; i < N1; ++i) a[r][i] = bb[r][i];
  };
  auto rows_to_cols = [&]() {
    const int t = opaque_v(tid);
    const int c0 = t % CH, u = t / CH;
    const int r0 = t % RH, v = t / RH;
    C bb[H][N1];
#pragma unroll
    for (int r = 0; r < H; ++r) {
      if (r) __syncthreads();
#pragma unroll
      for (int i = 0; i < N1; ++i) lds[r0 * TS + v + N2Z * i] = a[r][i];
      __syncthreads();
#pragma unroll
      for (int h = 0; h < H; ++h)
#pragma unroll
        for (int i = 0; i < N1H; ++i) bb[h][r * N1H + i] = lds[(u + N2Y * i) * TS + c0 + h * CH];
    }
#pragma unroll
    for (int h = 0; h < H; ++h)
#pragma unroll
      for (int i = 0; i < N1; ++i) a[h][i] = bb[h][i];
  };

  {
    const int c0 = tid % CH, u = tid / CH;
#pragma unroll
    for (int h = 0; h < H; ++h)
#pragma unroll
      for (int i = 0; i < N1; ++i) a[h][i] = blk[unsigned(c0 + h * CH) + unsigned(u + N2Y * i) * RS];
  }

  for (;;) {
    // persistent form: the next group's points, loaded now and used after this group's store
    C nxt[PERSIST ? N1 : 1];
    const long next = group + gridDim.x;
    const bool more = PERSIST && next < groups;
    if constexpr (PERSIST) {
      if (more) {
        const long np = next / B;
        const C* __restrict__ nb = data + np * long(NY) * NZ + long(next % B) * NZ;
        const int c0 = tid % CH, u = tid / CH;
#pragma unroll
        for (int i = 0; i < N1; ++i) nxt[i] = nb[unsigned(c0) + unsigned(u + N2Y * i) * RS];
      }
    }

    if constexpr (MODE == LINES_FULL) {
      // y inverse on columns
      transform_y(p.tw_yi, std::integral_constant<int, +1>{});
      __syncthreads();
      cols_to_rows();
      __syncthreads();

      // z inverse, potential phase and z forward on rows c0 + r LY/H (z = u + N2Z i).  Each
      // row round is finished (exchange, potential, first forward step) before the next one's
      // exchange, so that only one round's potential factors are held.
      {
        const int u = thread_u_z();
#pragma unroll
        for (int r = 0; r < H; ++r) {
          if (r) __syncthreads();
          load_wz(p.tw_z, u);
          column_dft_first<NZ, N2Z, +1>(a[r], w);
          exchange_z(a[r], u, std::integral_constant<int, +1>{});
          const int c0 = opaque_v(tid) % RH;
          if (p.spx) {
            const long pk = plane / p.nx;
            const C* __restrict__ px = reinterpret_cast<const C*>(p.spx);
            const C* __restrict__ py = reinterpret_cast<const C*>(p.spy) + pk * NY;
            const C* __restrict__ pzb = reinterpret_cast<const C*>(p.spz) + pk * NZ;
            const C pxy = cmul(px[plane], py[bres + B * (c0 + r * RH)]);
            const int ur = opaque_v(u);     // separate loads per round (not held across rounds)
#pragma unroll
            for (int i = 0; i < N1; ++i) {
              a[r][i] = cmul(a[r][i], cmul(pzb[ur + N2Z * i], pxy));
              if (H > 1 && i % 4 == 3) asm volatile("" ::: "memory");   // at most 4 factors in flight
            }
          } else if constexpr (PV_LINES) {
            constexpr unsigned T = unsigned(yz_lines_threads<NZ, LY, H>());
            // x planes whose factors equal the previous plane's read that plane's copy
            const long tg = p.pv_planes ? long(p.pv_planes[plane]) * B + bres : group;
            const C* __restrict__ pvb = reinterpret_cast<const C*>(p.pv) + tg * (long(LY) * NZ);
#pragma unroll
            for (int i = 0; i < N1; ++i) {
              a[r][i] = cmul(a[r][i], pvb[unsigned(r * N1 + i) * T + unsigned(tid)]);
              if (H > 1 && i % 4 == 3) asm volatile("" ::: "memory");
            }
          } else {
            const C* __restrict__ pvb = reinterpret_cast<const C*>(p.pv) + plane * long(NY) * NZ + long(bres) * NZ;
#pragma unroll
            for (int i = 0; i < N1; ++i) {
              a[r][i] = cmul(a[r][i], pvb[unsigned(c0 + r * RH) * RS + unsigned(u + N2Z * i)]);
              if (H > 1 && i % 4 == 3) asm volatile("" ::: "memory");
            }
          }
          load_wz(p.tw_z, u);
          column_dft_first<NZ, N2Z, -1>(a[r], w);
        }
#pragma unroll
        for (int r = 0; r < H; ++r) {
          __syncthreads();
          exchange_z(a[r], u, std::integral_constant<int, -1>{});
        }
      }
      __syncthreads();
      rows_to_cols();
      __syncthreads();

      // y forward on columns
      transform_y(p.tw_yf, std::integral_constant<int, -1>{});
    } else if constexpr (MODE == LINES_ENTRY) {
      cols_to_rows();
      __syncthreads();
      transform_z(p.tw_z, std::integral_constant<int, -1>{});
      __syncthreads();
      rows_to_cols();
      __syncthreads();
      transform_y(p.tw_yf, std::integral_constant<int, -1>{});
    } else if constexpr (MODE == LINES_ENTRY_OBSERVE) {
      // the entry launch with the box moments of the position-space rows it reads
      static_assert(H == 1 && !PERSIST, "the observing launch holds one round, one group per block");
      cols_to_rows();
      lines_moments<R, NZ, LY, B, N2Z, N1, yz_lines_threads<NZ, LY, H>()>(a[0], p.mom, plane, bres, group, tid,
                                                                          thread_u_z());
      __syncthreads();
      transform_z(p.tw_z, std::integral_constant<int, -1>{});
      __syncthreads();
      rows_to_cols();
      __syncthreads();
      transform_y(p.tw_yf, std::integral_constant<int, -1>{});
    } else if constexpr (MODE == LINES_EXIT) {
      transform_y(p.tw_yi, std::integral_constant<int, +1>{});
      __syncthreads();
      cols_to_rows();
      __syncthreads();
      transform_z(p.tw_z, std::integral_constant<int, +1>{});
      __syncthreads();
      rows_to_cols();
    } else {
      // LINES_OBSERVE: the exit launch, the box moments of the position-space rows it would
      // store, and the entry launch that would read them back, in one pass (same arithmetic)
      static_assert(H == 1 && !PERSIST, "the observing launch holds one round, one group per block");
      transform_y(p.tw_yi, std::integral_constant<int, +1>{});
      __syncthreads();
      cols_to_rows();
      __syncthreads();
      transform_z(p.tw_z, std::integral_constant<int, +1>{});
      lines_moments<R, NZ, LY, B, N2Z, N1, yz_lines_threads<NZ, LY, H>()>(a[0], p.mom, plane, bres, group, tid,
                                                                          thread_u_z());
      __syncthreads();
      transform_z(p.tw_zf, std::integral_constant<int, -1>{});
      __syncthreads();
      rows_to_cols();
      __syncthreads();
      transform_y(p.tw_yf, std::integral_constant<int, -1>{});
    }

    {
      const int t = opaque_v(tid);
      const int c0 = t % CH, u = t / CH;
#pragma unroll
      for (int h = 0; h < H; ++h)
#pragma unroll
        for (int i = 0; i < N1; ++i) blk[unsigned(c0 + h * CH) + unsigned(u + N2Y * i) * RS] = a[h][i];
    }
    if (!more) break;
    if constexpr (PERSIST) {
#pragma unroll
      for (int i = 0; i < N1; ++i) a[0][i] = nxt[i];
      group = next;
      plane = group / B;
      bres = int(group % B);
      blk = data + plane * long(NY) * NZ + long(bres) * NZ;
      __syncthreads();   // the LDS of this group's last exchange is written again next
    }
  }
}

}  // namespace wfs
