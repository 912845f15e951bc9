// Strided batched GEMM on the bf16 matrix cores of gfx950, FP32 accumulation.
//
//   C[b](m, n) = alpha * sum_k A[b](m, k) B[b](k, n) + beta * C[b](m, n)
//
// with every operand addressed by explicit element strides, so one kernel
// serves all contractions of the Fourier neural operator (truncated DFTs along
// each axis, channel mixing, spectral weights) and their backward passes
// (which are the same contractions with swapped strides).  A and B are FP32 in
// memory and rounded to bf16 (nearest even) as they are staged into LDS; the
// bf16 inputs are what the matrix cores multiply.
//
// Split operands (flags bit 0: A, bit 1: B): the operand is staged as two bf16
// parts, hi = bf16(x) and lo = bf16(x - hi), and the tile product adds
// Ahi Bhi + Alo Bhi + Ahi Blo (lo lo is dropped), so that operand carries about
// 16 significant bits.  The Fourier neural operator splits its activations and
// keeps weights and transform tables in plain bf16: rounding a fixed operand
// perturbs the map by a fixed amount, while rounding activations adds noise
// that does not shrink with the change of the input.
//
// Block tile (32*WM) x (32*WN) x 32, one wavefront per 32 x 32 sub-tile, using
// gfx950's v_mfma_f32_32x32x16_bf16 (16 k per instruction, twice the rate of the 8-k
// form): lane l gives 8 k values of row l%32 of A and the same 8 k values of column l%32
// of B; accumulator item v of lane l is C[8(v/4)+4(l/32)+v%4][l%32].  Which k a lane's
// element stands for needs no care: the instruction sums A[i][k] B[k][j] over the 16
// (lane half, element) slots, so any assignment of a step's 16 k values to those slots
// that A and B share gives the same sum.  Lane l's elements e are k = k0 + 8(e/4) +
// 4(l/32) + e%4 here (the two 4-element runs of the 8-k form's layout side by side).
//
// Split-K: gridDim.z = batch * splits; with splits > 1 each block writes its
// partial tile to a workspace and a second kernel adds the splits in a fixed
// order (deterministic), then applies alpha/beta.
#include "wfs_internal.h"

namespace wfs {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// the k values of two 4-element runs as one 8-element operand
__device__ __forceinline__ s16x8 cat8(const bf16x4& a, const bf16x4& b) {
  return s16x8{a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]};
}

// C += A B over 16 k on the bf16 matrix cores (v_mfma_f32_32x32x16_bf16)
__device__ __forceinline__ f32x16 mfma16(const s16x8& a, const s16x8& b, const f32x16& c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                  0, 0);
}

// round to nearest even (gfx950 converts in hardware: v_cvt_pk_bf16_f32)
__device__ __forceinline__ unsigned short f32_to_bf16(float f) {
  return __builtin_bit_cast(unsigned short, (__bf16)f);
}

struct GemmArgs {
  const float* A;
  const float* B;
  float* C;
  float* work;         // split-K partials [batch][splits][M][N]
  long M, N, K;
  long sAb, sAm, sAk;
  long sBb, sBm_unused, sBk, sBn;
  long sCb, sCm, sCn;
  float alpha, beta;
  int batch, splits;
  int flags;           // bit 0: split A into hi + lo; bit 1: split B; bit 2: do not store C
  const float* bias;   // per-row bias added after alpha/beta (may be null)
  float* act;          // GELU of the result, same strides as C (may be null)
  long kchunk;         // K range per split (multiple of 32)
};

constexpr int kBK = 32;

// C = alpha acc + beta C (+ bias[m]); act = GELU(C); C not stored with flag bit 2
__device__ __forceinline__ void epilogue(const GemmArgs& g, long b, long m, long n, float accv) {
  const long off = b * g.sCb + m * g.sCm + n * g.sCn;
  float val = (g.beta != 0.0f) ? g.alpha * accv + g.beta * g.C[off] : g.alpha * accv;
  if (g.bias) val = val + g.bias[m];
  if (!(g.flags & 4)) g.C[off] = val;
  if (g.act) g.act[off] = gelu_f(val);
}

template <int WM, int WN, bool SA, bool SB>
__global__ __launch_bounds__(64 * WM * WN) void gemm_bf16_kernel(GemmArgs g) {
  constexpr int BM = 32 * WM, BN = 32 * WN;
  constexpr int NT = 64 * WM * WN;
  // k-contiguous bf16 tiles, padded by 4 elements against bank conflicts
  __shared__ unsigned short As[BM][kBK + 4];
  __shared__ unsigned short Bs[BN][kBK + 4];
  __shared__ unsigned short Al[SA ? BM : 1][kBK + 4];
  __shared__ unsigned short Bl[SB ? BN : 1][kBK + 4];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int b = blockIdx.z / g.splits;
  const int split = blockIdx.z % g.splits;
  const long n0 = long(blockIdx.x) * BN;   // x: column tiles (the long dimension of the network's GEMMs)
  const long m0 = long(blockIdx.y) * BM;
  const long kbeg = long(split) * g.kchunk;
  long kend = kbeg + g.kchunk;
  if (kend > g.K) kend = g.K;

  const float* __restrict__ A = g.A + long(b) * g.sAb;
  const float* __restrict__ B = g.B + long(b) * g.sBb;
  const bool a_k_fast = (g.sAk == 1);
  const bool b_n_fast = (g.sBn == 1);
  // 16-byte loads of four consecutive elements along the contiguous axis when every
  // quad start is 16-byte aligned (k0 is a multiple of 32)
  const bool a_vec = a_k_fast && (g.sAm % 4 == 0) && (g.sAb % 4 == 0) &&
                     (reinterpret_cast<unsigned long>(g.A) % 16 == 0);
  const bool b_vec = b_n_fast && (g.sBk % 4 == 0) && (g.sBb % 4 == 0) &&
                     (reinterpret_cast<unsigned long>(g.B) % 16 == 0);

  f32x16 acc;
#pragma unroll
  for (int v = 0; v < 16; ++v) acc[v] = 0.0f;

  for (long k0 = kbeg; k0 < kend; k0 += kBK) {
    // stage A (BM x 32) and B (32 x BN) as bf16
    if (a_vec) {
      for (int e = tid; e < BM * (kBK / 4); e += NT) {
        const int mm = e / (kBK / 4), kk = (e % (kBK / 4)) * 4;
        const long m = m0 + mm, k = k0 + kk;
        float v[4];
        if (m < g.M && k + 3 < kend) {
          const float4 q = *reinterpret_cast<const float4*>(A + m * g.sAm + k);
          v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
        } else {
#pragma unroll
          for (int j = 0; j < 4; ++j) v[j] = (m < g.M && k + j < kend) ? A[m * g.sAm + k + j] : 0.0f;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const unsigned short h = f32_to_bf16(v[j]);
          As[mm][kk + j] = h;
          if constexpr (SA) Al[mm][kk + j] = f32_to_bf16(v[j] - __uint_as_float(unsigned(h) << 16));
        }
      }
    } else {
      for (int e = tid; e < BM * kBK; e += NT) {
        int mm, kk;
        if (a_k_fast) { mm = e / kBK; kk = e % kBK; } else { kk = e / BM; mm = e % BM; }
        const long m = m0 + mm, k = k0 + kk;
        const float v = (m < g.M && k < kend) ? A[m * g.sAm + k * g.sAk] : 0.0f;
        const unsigned short h = f32_to_bf16(v);
        As[mm][kk] = h;
        if constexpr (SA) Al[mm][kk] = f32_to_bf16(v - __uint_as_float(unsigned(h) << 16));
      }
    }
    if (b_vec) {
      for (int e = tid; e < kBK * (BN / 4); e += NT) {
        const int kk = e / (BN / 4), nn = (e % (BN / 4)) * 4;
        const long n = n0 + nn, k = k0 + kk;
        float v[4];
        if (k < kend && n + 3 < g.N) {
          const float4 q = *reinterpret_cast<const float4*>(B + k * g.sBk + n);
          v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
        } else {
#pragma unroll
          for (int j = 0; j < 4; ++j) v[j] = (k < kend && n + j < g.N) ? B[k * g.sBk + n + j] : 0.0f;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const unsigned short h = f32_to_bf16(v[j]);
          Bs[nn + j][kk] = h;
          if constexpr (SB) Bl[nn + j][kk] = f32_to_bf16(v[j] - __uint_as_float(unsigned(h) << 16));
        }
      }
    } else {
      for (int e = tid; e < BN * kBK; e += NT) {
        int nn, kk;
        if (b_n_fast) { kk = e / BN; nn = e % BN; } else { nn = e / kBK; kk = e % kBK; }
        const long n = n0 + nn, k = k0 + kk;
        const float v = (n < g.N && k < kend) ? B[k * g.sBk + n * g.sBn] : 0.0f;
        const unsigned short h = f32_to_bf16(v);
        Bs[nn][kk] = h;
        if constexpr (SB) Bl[nn][kk] = f32_to_bf16(v - __uint_as_float(unsigned(h) << 16));
      }
    }
    __syncthreads();
    const int row = wm * 32 + (lane & 31);
    const int col = wn * 32 + (lane & 31);
    const int kq = 4 * (lane >> 5);
#pragma unroll
    for (int kk = 0; kk < kBK; kk += 16) {
      s16x8 a, bb;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        a[j] = (short)As[row][kk + kq + j];
        a[4 + j] = (short)As[row][kk + 8 + kq + j];
        bb[j] = (short)Bs[col][kk + kq + j];
        bb[4 + j] = (short)Bs[col][kk + 8 + kq + j];
      }
      if constexpr (SA) {
        s16x8 al;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          al[j] = (short)Al[row][kk + kq + j];
          al[4 + j] = (short)Al[row][kk + 8 + kq + j];
        }
        acc = mfma16(al, bb, acc);
      }
      if constexpr (SB) {
        s16x8 bl;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          bl[j] = (short)Bl[col][kk + kq + j];
          bl[4 + j] = (short)Bl[col][kk + 8 + kq + j];
        }
        acc = mfma16(a, bl, acc);
      }
      acc = mfma16(a, bb, acc);
    }
    __syncthreads();
  }

  const long n = n0 + wn * 32 + (lane & 31);
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const long m = m0 + wm * 32 + 8 * (v / 4) + 4 * (lane >> 5) + (v % 4);
    if (m < g.M && n < g.N) {
      if (g.splits == 1) {
        epilogue(g, b, m, n, acc[v]);
      } else {
        g.work[((long(b) * g.splits + split) * g.M + m) * g.N + n] = acc[v];
      }
    }
  }
}

// one wavefront per output element: lane l adds splits l, l + 64, ... in order, then a
// fixed xor tree across the lanes (deterministic)
__global__ __launch_bounds__(256) void gemm_splitk_reduce_kernel(GemmArgs g) {
  const long mn = g.M * g.N;
  const long total = long(g.batch) * mn;
  const int lane = threadIdx.x & 63;
  const long e = long(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (e >= total) return;   // uniform per wavefront
  const long b = e / mn, r = e % mn;
  const float* __restrict__ w = g.work + long(b) * g.splits * mn + r;
  float s = 0.0f;
  for (int sp = lane; sp < g.splits; sp += 64) s += w[long(sp) * mn];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) epilogue(g, b, r / g.N, r % g.N, s);
}

// ---------------------------------------------------------------------------
// One layer of the Fourier neural operator, fused over a grid line.  For the line
// r = (x, y) (all n = 32 NZC points along z) and all channels:
//   u  = W v + Z Gz + b        (W: C x cin, v: cin channels, Z: the layer's inverse
//                               spectral path before its last (z) stage, C x tm per line)
//   v' = GELU(u) or u
//   T' = v' Fz                 (the next layer's first (z) spectral stage, C x tm per line)
// which the unfused network computes as four GEMMs and writes S = Z Gz and u to HBM in
// between.  Operand rounding is that of those GEMMs: W, Gz, Fz plain bf16, v, Z and v'
// split into bf16 hi + lo; products accumulate in FP32 (in another order).
//
// One wavefront per line, no LDS and no barriers.  The transposed product is formed,
// u^T = v^T W^T + Gz^T Z^T, so that (in 8-k runs: lane l holds A[l%32][8s+4(l/32)..+3],
// B[8s+4(l/32)..+3][l%32], two runs s = 2S, 2S + 1 per 16-k instruction, and
// accumulator item i is C[8(i/4)+4(l/32)+i%4][l%32]) every operand comes straight from
// memory or from the previous product: the A fragment of v^T is four channel rows at one
// z (each load instruction reads 2 x 128 contiguous bytes), W^T, Gz^T and Fz are per-lane
// constants held in registers, Z^T is one 16-byte load per run, and the accumulator of a
// 32-point z chunk of u^T, items 4q..4q+3, is already the A run of v' for the z run q of
// T = v' Fz.  A chunk takes 12 matrix instructions (W v, Z Gz and v' Fz, two each per
// 16 k for the hi and lo parts).  C, cin, tm <= 32.
// ---------------------------------------------------------------------------
struct LayerArgs {
  const float* W;
  const float* v;
  const float* Z;      // may be null (no spectral term)
  const float* bias;
  const float* Gz;     // [tm][n]
  const float* Fz;     // [n][tm]
  float* u;            // may be null
  float* vn;           // may be null
  float* T;            // may be null
  int cin, C, tm, act;
  long rows;           // lines (n * n)
};

__device__ __forceinline__ void split4(const float (&x)[4], bf16x4& hi, bf16x4& lo) {
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const unsigned short h = f32_to_bf16(x[j]);
    hi[j] = (short)h;
    lo[j] = (short)f32_to_bf16(x[j] - __uint_as_float(unsigned(h) << 16));
  }
}

template <int NZC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void fno_layer_kernel(LayerArgs g) {
  constexpr int n = 32 * NZC;
  constexpr int GP = kBK + 4;                        // padded rows of the transform tables
  __shared__ unsigned short GzT[n][GP];              // A = Gz^T: [z][j]
  __shared__ unsigned short FzT[32][n + 4];          // B = Fz:   [j][z]
  const int lane = threadIdx.x & 63;
  const int lc = lane & 31;
  const int kq = 4 * (lane >> 5);
  const int C = g.C, cin = g.cin, tm = g.tm;
  const long rows = g.rows;
  const unsigned P = unsigned(rows * n);             // C * P * 4 bytes < 2^32 (n <= 128, C <= 32)
  const int kin = (cin + 7) / 8;                     // k steps of W v
  const int ktm = g.Z ? (tm + 7) / 8 : 0;            // k steps of Z Gz

  for (int e = threadIdx.x; e < n * 32; e += blockDim.x) {
    const int z = e / 32, j = e % 32;
    GzT[z][j] = (g.Z && j < tm) ? f32_to_bf16(g.Gz[j * n + z]) : (unsigned short)0;
    FzT[j][z] = (g.T && j < tm) ? f32_to_bf16(g.Fz[z * tm + j]) : (unsigned short)0;
  }
  bf16x4 Wf[4];                                      // B = W^T: (k = c_in, n = c_out = lc)
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = 8 * s + kq + j;
      Wf[s][j] = (short)((lc < C && k < cin) ? f32_to_bf16(g.W[lc * cin + k]) : 0);
    }
  const float bias_c = lc < C ? g.bias[lc] : 0.0f;
  __syncthreads();

  // 32-bit byte offsets from the buffers' base addresses (C * n^3 * 4 < 2^32); loads past
  // the operand's extent read a clamped valid element and are zeroed, so no load is
  // conditional
  const char* __restrict__ vb = reinterpret_cast<const char*>(g.v);
  const char* __restrict__ zb = reinterpret_cast<const char*>(g.Z);
  char* __restrict__ ub = reinterpret_cast<char*>(g.u);
  char* __restrict__ nb = reinterpret_cast<char*>(g.vn);
  char* __restrict__ tb = reinterpret_cast<char*>(g.T);
  const unsigned lcC = unsigned(lc < C ? lc : C - 1);
  const long stride = long(gridDim.x) * (blockDim.x / 64);
  // the A fragments of v^T for one 32-point chunk; the next chunk's (or the next line's
  // first chunk's) are loaded while the current one is computed
  auto load_v = [&](unsigned ru, int ch, float (&x)[16]) {
    const unsigned zoff = ru * n + 32 * ch + lc;
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int c = 8 * s + kq + j;
        x[4 * s + j] = *reinterpret_cast<const float*>(vb + ((c < cin ? unsigned(c) : unsigned(cin - 1)) * P + zoff) * 4u);
      }
  };
  float xv[16];
  long r = long(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (r < rows) load_v(unsigned(r), 0, xv);
  for (; r < rows; r += stride) {
    const unsigned ru = unsigned(r);
    bf16x4 Zh[4], Zl[4];                             // B = Z^T: (k = j, n = c = lc)
#pragma unroll
    for (int s = 0; s < 4; ++s) { Zh[s] = bf16x4{0, 0, 0, 0}; Zl[s] = bf16x4{0, 0, 0, 0}; }
    if (zb) {
      // four consecutive j per lane: one 16-byte load (tm is a multiple of 4, so a quad lies
      // wholly inside or wholly past the row; past it, the row's last quad is read and dropped)
      const unsigned row = (lcC * unsigned(rows) + ru) * unsigned(tm);
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int k0 = 8 * s + kq;
        const float4 y = *reinterpret_cast<const float4*>(zb + (row + unsigned(k0 < tm ? k0 : tm - 4)) * 4u);
        const bool ok = k0 < tm && lc < C;
        float x[4] = {ok ? y.x : 0.0f, ok ? y.y : 0.0f, ok ? y.z : 0.0f, ok ? y.w : 0.0f};
        split4(x, Zh[s], Zl[s]);
      }
    }
    f32x16 accT;
#pragma unroll
    for (int i = 0; i < 16; ++i) accT[i] = 0.0f;
    const unsigned line = ru * n;
    const long rnext = r + stride;
    // chunk ch of this line from x (its loads were issued one chunk earlier)
    auto process = [&](int ch, float (&x)[16]) {
#pragma unroll
      for (int i = 0; i < 16; ++i)
        if (8 * (i / 4) + kq + (i % 4) >= cin) x[i] = 0.0f;   // channels past cin
      f32x16 acc;
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
      // k steps of 16: the 8-k steps 2S and 2S + 1 side by side (past cin and tm the
      // operands are zero)
#pragma unroll
      for (int S = 0; S < 2; ++S) {
        if (2 * S < kin) {
          bf16x4 ah0, al0, ah1, al1;
          float x0[4] = {x[8 * S], x[8 * S + 1], x[8 * S + 2], x[8 * S + 3]};
          float x1[4] = {x[8 * S + 4], x[8 * S + 5], x[8 * S + 6], x[8 * S + 7]};
          split4(x0, ah0, al0);
          split4(x1, ah1, al1);
          const s16x8 w = cat8(Wf[2 * S], Wf[2 * S + 1]);
          acc = mfma16(cat8(al0, al1), w, acc);
          acc = mfma16(cat8(ah0, ah1), w, acc);
        }
      }
#pragma unroll
      for (int S = 0; S < 2; ++S) {
        if (2 * S < ktm) {
          s16x8 a;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            a[j] = (short)GzT[32 * ch + lc][16 * S + kq + j];
            a[4 + j] = (short)GzT[32 * ch + lc][16 * S + 8 + kq + j];
          }
          acc = mfma16(a, cat8(Zl[2 * S], Zl[2 * S + 1]), acc);
          acc = mfma16(a, cat8(Zh[2 * S], Zh[2 * S + 1]), acc);
        }
      }
      // lane: column c = lc; items 4q..4q+3: z = 32 ch + 8 q + kq + 0..3; the items of q = 2Q
      // and 2Q + 1 are the A operand of one 16-k step of T = v' Fz
#pragma unroll
      for (int Q = 0; Q < 2; ++Q) {
        bf16x4 ahq[2], alq[2], bq[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int q = 2 * Q + h;
          float val[4], gv[4];
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            val[j] = acc[4 * q + j] + bias_c;
            gv[j] = lc < C ? (g.act ? gelu_f(val[j]) : val[j]) : 0.0f;
          }
          if (lc < C) {
            const unsigned off = (unsigned(lc) * P + line + 32 * ch + 8 * q + kq) * 4u;
            if (ub) *reinterpret_cast<float4*>(ub + off) = make_float4(val[0], val[1], val[2], val[3]);
            if (nb) *reinterpret_cast<float4*>(nb + off) = make_float4(gv[0], gv[1], gv[2], gv[3]);
          }
          if (tb) {
            split4(gv, ahq[h], alq[h]);
#pragma unroll
            for (int j = 0; j < 4; ++j) bq[h][j] = (short)FzT[lc][32 * ch + 8 * q + kq + j];
          }
        }
        if (tb) {
          const s16x8 b = cat8(bq[0], bq[1]);
          accT = mfma16(cat8(alq[0], alq[1]), b, accT);
          accT = mfma16(cat8(ahq[0], ahq[1]), b, accT);
        }
      }
    };
    // the chunk after ch: this line's ch + 1, or the next line's first
    auto prefetch = [&](int ch, float (&x)[16]) {
      if (ch + 1 < NZC) load_v(ru, ch + 1, x);
      else if (rnext < rows) load_v(unsigned(rnext), 0, x);
    };
    // chunks in pairs through two buffers, so that no register copy waits for a load in flight
    float xw[16];
#pragma unroll 1
    for (int ch = 0; ch < NZC; ch += 2) {
      prefetch(ch, xw);
      process(ch, xv);
      if (ch + 1 < NZC) {
        prefetch(ch + 1, xv);
        process(ch + 1, xw);
      } else {
#pragma unroll
        for (int i = 0; i < 16; ++i) xv[i] = xw[i];   // odd chunk count: next line's first chunk
      }
    }
    if (tb && lc < tm) {
      // accumulator: column j = lc, rows c = 8(i/4) + 4(l/32) + i%4
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int c = 8 * (i / 4) + kq + (i % 4);
        if (c < C)
          *reinterpret_cast<float*>(tb + ((unsigned(c) * unsigned(rows) + ru) * unsigned(tm) + lc) * 4u) = accT[i];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Complex per-mode channel mixing of the Fourier neural operator:
//   O[o][q] = sum_i R[q][i][o] X[i][q]     (complex: Or = Rr Xr - Ri Xi, Oi = Ri Xr + Rr Xi)
// for m3 modes q and C channels, with the GEMM path's operand rounding (R plain bf16,
// X as bf16 hi + lo, FP32 accumulation).  The GEMM path issues it as four strided
// batched GEMMs with N = 1 (one 32 x 32 tile per mode for one useful column); here four
// lanes own one (mode, output channel) pair.  Layouts: R [m3][C][C] ([q][i][o]),
// X and O [C][2][m3] (real parts, then imaginary parts).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void fno_modes_kernel(const float* __restrict__ Rr, const float* __restrict__ Ri,
                                                        const float* __restrict__ X, float* __restrict__ O,
                                                        int m3, int C) {
  // four lanes per (mode, output channel): lane h sums input channels i = h, h + 4, ...,
  // then the four partial sums are added across the lanes (xor 1, xor 2)
  const long t = long(blockIdx.x) * 256 + threadIdx.x;
  const long pair = t >> 2;
  const int h = int(t & 3);
  const bool live = pair < long(m3) * C;
  const long pp = live ? pair : 0;
  const int q = int(pp / C), o = int(pp % C);
  const float* __restrict__ rr = Rr + long(q) * C * C + o;
  const float* __restrict__ ri = Ri + long(q) * C * C + o;
  float srr = 0.0f, sii = 0.0f, sir = 0.0f, sri = 0.0f;   // Rr Xr, Ri Xi, Ri Xr, Rr Xi
  for (int i = h; i < C; i += 4) {
    const float a = __uint_as_float(unsigned(f32_to_bf16(rr[long(i) * C])) << 16);
    const float b = __uint_as_float(unsigned(f32_to_bf16(ri[long(i) * C])) << 16);
    const float xr = X[long(i) * 2 * m3 + q], xi = X[long(i) * 2 * m3 + m3 + q];
    const float xrh = __uint_as_float(unsigned(f32_to_bf16(xr)) << 16);
    const float xih = __uint_as_float(unsigned(f32_to_bf16(xi)) << 16);
    const float xrl = __uint_as_float(unsigned(f32_to_bf16(xr - xrh)) << 16);
    const float xil = __uint_as_float(unsigned(f32_to_bf16(xi - xih)) << 16);
    srr = __builtin_fmaf(a, xrh, __builtin_fmaf(a, xrl, srr));
    sii = __builtin_fmaf(b, xih, __builtin_fmaf(b, xil, sii));
    sir = __builtin_fmaf(b, xrh, __builtin_fmaf(b, xrl, sir));
    sri = __builtin_fmaf(a, xih, __builtin_fmaf(a, xil, sri));
  }
#pragma unroll
  for (int off = 1; off < 4; off <<= 1) {
    srr += __shfl_xor(srr, off, 64);
    sii += __shfl_xor(sii, off, 64);
    sir += __shfl_xor(sir, off, 64);
    sri += __shfl_xor(sri, off, 64);
  }
  if (live && h == 0) {
    O[long(o) * 2 * m3 + q] = srr - sii;
    O[long(o) * 2 * m3 + m3 + q] = sir + sri;
  }
}

}  // namespace wfs

using namespace wfs;

extern "C" int wfs_fno_modes(const float* Rr, const float* Ri, const float* X, float* O, int m3, int C,
                             void* stream) {
  if (!Rr || !Ri || !X || !O || m3 < 1 || C < 1) { set_last_error("wfs_fno_modes: bad argument"); return WFS_ERR_INVALID; }
  const long total = long(m3) * C * 4;
  hipLaunchKernelGGL(fno_modes_kernel, dim3(unsigned((total + 255) / 256)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), Rr, Ri, X, O, m3, C);
  WFS_HIP_CHECK(hipGetLastError());
  return 0;
}

// layer of the Fourier neural operator fused over grid lines (see fno_layer_kernel)
extern "C" int wfs_fno_layer(const float* W, const float* v, const float* Z, const float* bias, const float* Gz,
                             const float* Fz, float* u, float* vn, float* T, int cin, int C, int n, int tm, int act,
                             void* stream) {
  if (!W || !v || !bias) { set_last_error("wfs_fno_layer: null argument"); return WFS_ERR_INVALID; }
  if (n % 32 != 0 || n < 32 || n > 128 || C < 1 || C > 32 || cin < 1 || cin > 32 || tm < 4 || tm > 32 || tm % 4) {
    set_last_error("wfs_fno_layer: needs n in {32, 64, 96, 128}, C, cin <= 32 and tm a multiple of 4 up to 32 "
                   "(n=%d C=%d cin=%d tm=%d)", n, C, cin, tm);
    return WFS_ERR_UNSUPPORTED;
  }
  if ((Z && !Gz) || (T && (!Fz || !vn)) || (!u && !vn)) {
    set_last_error("wfs_fno_layer: Z needs Gz; T needs Fz and vn; one of u, vn is required");
    return WFS_ERR_INVALID;
  }
  if ((reinterpret_cast<unsigned long>(u) % 16) != 0 || (reinterpret_cast<unsigned long>(vn) % 16) != 0 ||
      (reinterpret_cast<unsigned long>(Z) % 16) != 0) {
    set_last_error("wfs_fno_layer: u, vn and Z must be 16-byte aligned");
    return WFS_ERR_INVALID;
  }
  LayerArgs g{W, v, Z, bias, Gz, Fz, u, vn, T, cin, C, tm, act, long(n) * n};
  // one wavefront per line, four per block
  const long want = (g.rows + 3) / 4;
  const unsigned blocks = unsigned(want < 2048 ? want : 2048);
  hipStream_t s = static_cast<hipStream_t>(stream);
  switch (n / 32) {
    case 1: hipLaunchKernelGGL(fno_layer_kernel<1>, dim3(blocks), dim3(256), 0, s, g); break;
    case 2: hipLaunchKernelGGL(fno_layer_kernel<2>, dim3(blocks), dim3(256), 0, s, g); break;
    case 3: hipLaunchKernelGGL(fno_layer_kernel<3>, dim3(blocks), dim3(256), 0, s, g); break;
    default: hipLaunchKernelGGL(fno_layer_kernel<4>, dim3(blocks), dim3(256), 0, s, g); break;
  }
  WFS_HIP_CHECK(hipGetLastError());
  return 0;
}

template <int WM, int WN>
static void launch_tile(dim3 grid, hipStream_t s, const GemmArgs& g) {
  switch (g.flags & 3) {
    case 0: hipLaunchKernelGGL((gemm_bf16_kernel<WM, WN, false, false>), grid, dim3(64 * WM * WN), 0, s, g); break;
    case 1: hipLaunchKernelGGL((gemm_bf16_kernel<WM, WN, true, false>), grid, dim3(64 * WM * WN), 0, s, g); break;
    case 2: hipLaunchKernelGGL((gemm_bf16_kernel<WM, WN, false, true>), grid, dim3(64 * WM * WN), 0, s, g); break;
    default: hipLaunchKernelGGL((gemm_bf16_kernel<WM, WN, true, true>), grid, dim3(64 * WM * WN), 0, s, g); break;
  }
}

// dims = {batch, M, N, K}; strides = {sAb, sAm, sAk, sBb, sBk, sBn, sCb, sCm, sCn} (elements)
extern "C" int wfs_gemm_bf16(const float* A, const float* B, float* C, const long dims[4], const long strides[9],
                             float alpha, float beta, int flags, int splits, float* work, long work_elems,
                             const float* bias, float* act, void* stream) {
  if (!A || !B || !C || !dims || !strides) { set_last_error("wfs_gemm_bf16: null argument"); return WFS_ERR_INVALID; }
  GemmArgs g{};
  g.A = A; g.B = B; g.C = C; g.work = work;
  g.batch = int(dims[0]); g.M = dims[1]; g.N = dims[2]; g.K = dims[3];
  g.sAb = strides[0]; g.sAm = strides[1]; g.sAk = strides[2];
  g.sBb = strides[3]; g.sBk = strides[4]; g.sBn = strides[5];
  g.sCb = strides[6]; g.sCm = strides[7]; g.sCn = strides[8];
  g.alpha = alpha; g.beta = beta; g.flags = flags; g.bias = bias; g.act = act;
  if (flags & ~7) { set_last_error("wfs_gemm_bf16: unknown flags"); return WFS_ERR_INVALID; }
  if ((flags & 4) && !act) { set_last_error("wfs_gemm_bf16: WFS_GEMM_NO_C needs an activation output"); return WFS_ERR_INVALID; }
  if (g.batch < 1 || g.M < 1 || g.N < 1 || g.K < 1 || g.batch > 65535) { set_last_error("wfs_gemm_bf16: bad sizes"); return WFS_ERR_INVALID; }
  if (splits < 1) splits = 1;
  const long kblocks = (g.K + kBK - 1) / kBK;
  if (splits > kblocks) splits = int(kblocks);
  g.kchunk = ((kblocks + splits - 1) / splits) * kBK;
  splits = int((g.K + g.kchunk - 1) / g.kchunk);
  g.splits = splits;
  if (splits > 1 && (!work || work_elems < long(g.batch) * splits * g.M * g.N)) {
    set_last_error("wfs_gemm_bf16: split-K needs a workspace of batch*splits*M*N floats");
    return WFS_ERR_INVALID;
  }
  if (long(g.batch) * splits > 65535) { set_last_error("wfs_gemm_bf16: batch*splits too large"); return WFS_ERR_INVALID; }
  hipStream_t s = static_cast<hipStream_t>(stream);
  // tile shape by output shape: 32-row outputs use one wave row (32 x 128 tiles when wide, one
  // 32 x 32 wave when small and the k range per block is short), tall-skinny outputs and long-k
  // small outputs 128 x 32 (four waves share the staging), the rest 64 x 64
  int wm, wn;
  if (g.M <= 32 && g.N > 32) { wm = 1; wn = 4; }
  else if (g.M <= 32 && g.kchunk <= 256) { wm = 1; wn = 1; }
  else if (g.N <= 32) { wm = 4; wn = 1; }
  else { wm = 2; wn = 2; }
  const long gy = (g.M + 32 * wm - 1) / (32 * wm);
  const long gx = (g.N + 32 * wn - 1) / (32 * wn);
  if (gy > 65535 || gx > 0x7fffffffL) { set_last_error("wfs_gemm_bf16: output too large for the launch grid"); return WFS_ERR_INVALID; }
  dim3 grid(unsigned(gx), unsigned(gy), unsigned(g.batch * splits));
  if (wm == 1 && wn == 1) launch_tile<1, 1>(grid, s, g);
  else if (wm == 1) launch_tile<1, 4>(grid, s, g);
  else if (wm == 4) launch_tile<4, 1>(grid, s, g);
  else launch_tile<2, 2>(grid, s, g);
  WFS_HIP_CHECK(hipGetLastError());
  if (splits > 1) {
    const long total = long(g.batch) * g.M * g.N;
    const long blocks = (total + 3) / 4;
    if (blocks > 0x7fffffffL) { set_last_error("wfs_gemm_bf16: split-K output too large"); return WFS_ERR_INVALID; }
    hipLaunchKernelGGL(gemm_splitk_reduce_kernel, dim3(unsigned(blocks)), dim3(256), 0, s, g);
    WFS_HIP_CHECK(hipGetLastError());
  }
  return 0;
}
