// Power-series propagator exp(-i H dt) ~ sum_{n<=p} (-i H dt)^n / n! (gfx950).
//
// The reference evaluates the series by the recurrence
//   s_n = (i dt / n) [ (1/2m) L s_{n-1} - V s_{n-1} ],  psi += s_n
// with one kernel launch per term (sources/sim_state.py:721-752,
// sources/cuda_kernels/power_series_operator.cu:3-65).  The same polynomial is
// evaluated here by Horner's rule,
//   u <- psi;  for n = p..1:  u <- psi + (-i dt / n) H u,
// which reads u (stencil), psi and V and writes u once per stage: 32 B per
// point per stage in complex64 instead of 40.
//
// H u = -(1/2m) L u + V u with L the 7-point Laplacian whose neighbour index
// falls back to the voxel itself at a face (edge replication) and whose sum is
// divided by dx_x^2 on every axis (power_series_operator.cu:23-58).
//
// Each thread owns one (y, z) column of a block tile and marches along x over
// a 16-plane slab, keeping u(x-1), u(x), u(x+1) in registers; the y/z
// neighbours come from the same plane, which the block's other lanes load at
// the same time (L1/L2 hits).
#include "wfs_internal.h"

namespace wfs {

// One Horner update at a point: psi + (-i c) (-(1/2m) L u + V u), with L u * dx^2 summed in the
// reference's order.  Shared by the single-stage and the two-stage kernels so that both round
// identically.
template <typename R>
__device__ __forceinline__ cx<R> horner_point(cx<R> cur, cx<R> xm, cx<R> xp, cx<R> ym, cx<R> yp, cx<R> zm,
                                              cx<R> zp, cx<R> ps, float2 v, R c, R inv_2m, R inv_dx2) {
  cx<R> lap;
  lap.x = xm.x + xp.x + ym.x - R(6) * cur.x + yp.x + zm.x + zp.x;
  lap.y = xm.y + xp.y + ym.y - R(6) * cur.y + yp.y + zm.y + zp.y;
  cx<R> hu;
  hu.x = -inv_2m * inv_dx2 * lap.x + (R(v.x) * cur.x - R(v.y) * cur.y);
  hu.y = -inv_2m * inv_dx2 * lap.y + (R(v.x) * cur.y + R(v.y) * cur.x);
  // psi + (-i c) H u = psi + c (Hu.y, -Hu.x)
  cx<R> o;
  o.x = ps.x + c * hu.y;
  o.y = ps.y - c * hu.x;
  return o;
}

template <typename R, int TZ, int TY>
__global__ __launch_bounds__(TZ* TY) void horner_stage_kernel(const cx<R>* __restrict__ u_in, cx<R>* u_out,
                                                             const cx<R>* psi, const float2* __restrict__ V,
                                                             int nx, int ny, int nz, R dt_over_n, R inv_2m,
                                                             R inv_dx2, int tile_x, int ntiles) {
  using C = cx<R>;
  const int z = blockIdx.x * TZ + threadIdx.x;
  const int y = blockIdx.y * TY + threadIdx.y;
  const int packet = blockIdx.z / ntiles;          // independent grids stacked along x
  const int x0 = (blockIdx.z % ntiles) * tile_x;
  if (z >= nz || y >= ny) return;
  const long poff = long(packet) * nx * ny * nz;
  u_in += poff;
  u_out += poff;
  psi += poff;
  V += poff;
  int x1 = x0 + tile_x;
  if (x1 > nx) x1 = nx;
  const long sx = long(ny) * nz;
  const long sy = nz;
  const int zm = z > 0 ? z - 1 : z;
  const int zp = z < nz - 1 ? z + 1 : z;
  const int ym = y > 0 ? y - 1 : y;
  const int yp = y < ny - 1 ? y + 1 : y;

  auto at = [&](int x, int yy, int zz) -> long { return long(x) * sx + long(yy) * sy + zz; };

  C prev = u_in[at(x0 > 0 ? x0 - 1 : x0, y, z)];
  C cur = u_in[at(x0, y, z)];
  for (int x = x0; x < x1; ++x) {
    const C next = u_in[at(x < nx - 1 ? x + 1 : x, y, z)];
    const C uym = u_in[at(x, ym, z)];
    const C uyp = u_in[at(x, yp, z)];
    const C uzm = u_in[at(x, y, zm)];
    const C uzp = u_in[at(x, y, zp)];
    const long i = at(x, y, z);
    u_out[i] = horner_point<R>(cur, prev, next, uym, uyp, uzm, uzp, psi[i], V[i], dt_over_n, inv_2m, inv_dx2);
    prev = cur;
    cur = next;
  }
}

// Two Horner stages in one pass (temporal blocking): w = psi + (-i c1) H u, then
// u_out = psi + (-i c2) H w.  A block owns a TY x TZ (y, z) tile and marches along x
// through a slab; the LDS holds three x planes of u with a two-point (y, z) halo and
// three planes of w with a one-point halo.  Per output point it moves about 37 B per two
// stages (u with halo, psi and V, u_out) instead of 2 x 32.  Faces replicate the edge
// voxel as in the single stage: w outside the grid is a copy of w at the clamped voxel,
// and out-of-grid u cells hold u at the clamped voxel.  w is rounded to the state
// precision exactly as the single stage stores it, so the result equals two single stages
// bit for bit.  u_out must not alias psi or u_in (other blocks read their halos).
template <typename R, int TZ, int TY, int NT>
__global__ __launch_bounds__(NT) void horner2_kernel(const cx<R>* __restrict__ u_in, cx<R>* __restrict__ u_out,
                                                     const cx<R>* __restrict__ psi, const float2* __restrict__ V,
                                                     int nx, int ny, int nz, R c1, R c2, R inv_2m, R inv_dx2,
                                                     int tile_x, int ntiles) {
  using C = cx<R>;
  constexpr int UY = TY + 4, UZ = TZ + 4;   // u window (halo 2)
  constexpr int WY = TY + 2, WZ = TZ + 2;   // w window (halo 1)
  constexpr int RU = (UY * UZ + NT - 1) / NT;   // u cells per thread
  constexpr int RW = (WY * WZ + NT - 1) / NT;   // w cells per thread
  __shared__ C su[3][UY][UZ];
  __shared__ C sw[3][WY][WZ];

  const int tid = threadIdx.x;
  // Blocks are dispatched round-robin over the 8 XCDs in linear order.  Give each XCD a
  // contiguous run of tiles (z fastest, then y, then x slab): the tiles of one XCD march the
  // same x planes together, so the y/z halo lines one tile fetches are L2 hits for its
  // neighbours instead of second HBM reads (the order changes nothing in the arithmetic).
  int bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
  {
    const long total = long(gridDim.x) * gridDim.y * gridDim.z;
    if (total % 8 == 0) {
      const long lin = blockIdx.x + long(gridDim.x) * (blockIdx.y + long(gridDim.y) * blockIdx.z);
      const long logical = (lin % 8) * (total / 8) + lin / 8;
      bx = int(logical % gridDim.x);
      by = int((logical / gridDim.x) % gridDim.y);
      bz = int(logical / (long(gridDim.x) * gridDim.y));
    }
  }
  const int z0 = bx * TZ, y0 = by * TY;
  const int packet = bz / ntiles;
  const int x0 = (bz % ntiles) * tile_x;
  int x1 = x0 + tile_x;
  if (x1 > nx) x1 = nx;
  const long poff = long(packet) * nx * ny * nz;
  u_in += poff;
  u_out += poff;
  psi += poff;
  V += poff;
  const long sx = long(ny) * nz;
  auto clampi = [](int v, int n) { return v < 0 ? 0 : (v >= n ? n - 1 : v); };
  auto slot = [](int x) { return ((x % 3) + 3) % 3; };

  // Each thread owns the u cells e = tid + NT r and the w cells e = tid + NT r of the two
  // windows; what depends only on the cell (its voxel's offset within an x plane, its LDS
  // positions, whether it is a core cell) is computed once.  An out-of-grid w cell is
  // computed at its clamped voxel from the same u cells, psi and V as that voxel's own w
  // cell, so it equals the copy the clamped Laplacian implies, with no extra pass.  Plane
  // bases are block-uniform and in-plane offsets 32-bit (ny nz <= 2^18).
  unsigned uoff[RU];                        // clamped voxel of u cell r, offset in its plane
#pragma unroll
  for (int r = 0; r < RU; ++r) {
    const int e = tid + NT * r;
    const int gy = clampi(y0 - 2 + e / UZ, ny), gz = clampi(z0 - 2 + e % UZ, nz);
    uoff[r] = unsigned(gy) * unsigned(nz) + unsigned(gz);
  }
  unsigned woff[RW];                        // clamped voxel of w cell r, offset in its plane
  int wsu[RW];                              // that voxel's position in an su plane
  bool core[RW];                            // w cell r is an output cell of this block
#pragma unroll
  for (int r = 0; r < RW; ++r) {
    const int e = tid + NT * r;
    const int j = e / WZ, k = e % WZ;
    const int gy = y0 - 1 + j, gz = z0 - 1 + k;
    const int cy = clampi(gy, ny), cz = clampi(gz, nz);
    woff[r] = unsigned(cy) * unsigned(nz) + unsigned(cz);
    wsu[r] = (cy - (y0 - 2)) * UZ + (cz - (z0 - 2));
    core[r] = e < WY * WZ && j >= 1 && j <= TY && k >= 1 && k <= TZ && gy < ny && gz < nz;
  }

  // The next planes' operands are fetched into registers ahead so their latency overlaps the
  // current plane's arithmetic (u one plane ahead; psi and V of plane x + 2 while plane x is
  // computed, before w(x + 1) needs those of x + 1); psi and V of a w cell are kept for the
  // output at the same point one plane later.
  C ureg[RU];
  C pahead[RW], pnext[RW], pcur[RW];
  float2 vahead[RW], vnext[RW], vcur[RW];
  constexpr unsigned EB = unsigned(sizeof(C));

  auto fetch_u = [&](int x) {
    const C* __restrict__ base = u_in + long(clampi(x, nx)) * sx;
#pragma unroll
    for (int r = 0; r < RU; ++r)
      if (tid + NT * r < UY * UZ) ureg[r] = byte_at(base, uoff[r] * EB);
  };
  auto store_u = [&](int x) {
    C* dst = &su[slot(x)][0][0];
#pragma unroll
    for (int r = 0; r < RU; ++r) {
      const int e = tid + NT * r;
      if (e < UY * UZ) dst[e] = ureg[r];
    }
  };
  // psi and V at the (clamped) voxels of the owned w cells of plane x (x in the grid)
  auto fetch_pv_to = [&](int x, C (&pd)[RW], float2 (&vd)[RW]) {
    const C* __restrict__ pb = psi + long(x) * sx;
    const float2* __restrict__ vb = V + long(x) * sx;
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      if (tid + NT * r < WY * WZ) {
        pd[r] = byte_at(pb, woff[r] * EB);
        vd[r] = byte_at(vb, woff[r] * unsigned(sizeof(float2)));
      }
    }
  };
  auto fetch_pv = [&](int x) { fetch_pv_to(x, pnext, vnext); };
  // w plane x from u(x-1), u(x), u(x+1) and pnext/vnext; a plane outside the grid is a copy
  // of the nearest grid plane
  auto make_w = [&](int x) {
    if (x >= 0 && x < nx) {
      C* dst = &sw[slot(x)][0][0];
      const C* um = &su[slot(x - 1)][0][0];
      const C* uc = &su[slot(x)][0][0];
      const C* up = &su[slot(x + 1)][0][0];
#pragma unroll
      for (int r = 0; r < RW; ++r) {
        const int e = tid + NT * r;
        if (e < WY * WZ) {
          const int c = wsu[r];
          dst[e] = horner_point<R>(uc[c], um[c], up[c], uc[c - UZ], uc[c + UZ], uc[c - 1], uc[c + 1], pnext[r],
                                   vnext[r], c1, inv_2m, inv_dx2);
        }
      }
    } else {
      const int ws = slot(x), fs = slot(clampi(x, nx));
      for (int e = tid; e < WY * WZ; e += NT) {
        const int j = e / WZ, k = e % WZ;
        sw[ws][j][k] = sw[fs][j][k];
      }
    }
  };

  // prologue: u(x0-2..x0+2) and w(x0-1), w(x0); pcur = psi, V of plane x0
  fetch_u(x0 - 2);
  store_u(x0 - 2);
  fetch_u(x0 - 1);
  store_u(x0 - 1);
  fetch_u(x0);
  store_u(x0);
  if (x0 - 1 >= 0) fetch_pv(x0 - 1);
  __syncthreads();
  if (x0 - 1 >= 0) {
    make_w(x0 - 1);
    __syncthreads();
  }
  fetch_u(x0 + 1);
  store_u(x0 + 1);           // replaces u(x0-2)
  fetch_pv(x0);
  __syncthreads();
  make_w(x0);
  __syncthreads();
  if (x0 - 1 < 0) {
    make_w(x0 - 1);          // copy of w(x0)
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < RW; ++r) { pcur[r] = pnext[r]; vcur[r] = vnext[r]; }
  fetch_u(x0 + 2);
  store_u(x0 + 2);           // replaces u(x0-1)
  if (x0 + 1 < nx) fetch_pv(x0 + 1);
  __syncthreads();

#pragma unroll
  for (int r = 0; r < RW; ++r) {   // defined values where no plane x + 2 is fetched
    pahead[r] = pnext[r];
    vahead[r] = vnext[r];
  }
  // state at the top of iteration x: su = u(x..x+2), sw = w(x-1), w(x), pnext = psi/V(x+1),
  // pcur = psi/V(x)
  for (int x = x0; x < x1; ++x) {
    fetch_u(x + 3);                           // into registers, stored after this plane
    if (x + 2 < nx) fetch_pv_to(x + 2, pahead, vahead);   // used by w(x + 2), next iteration
    make_w(x + 1);                            // replaces w(x-2)
    __syncthreads();
    // output plane x at the owned core cells
    {
      const C* wm = &sw[slot(x - 1)][0][0];
      const C* wc = &sw[slot(x)][0][0];
      const C* wp = &sw[slot(x + 1)][0][0];
      C* __restrict__ ob = u_out + long(x) * sx;
#pragma unroll
      for (int r = 0; r < RW; ++r) {
        const int e = tid + NT * r;
        if (core[r])
          byte_at(ob, woff[r] * EB) = horner_point<R>(wc[e], wm[e], wp[e], wc[e - WZ], wc[e + WZ], wc[e - 1],
                                                      wc[e + 1], pcur[r], vcur[r], c2, inv_2m, inv_dx2);
      }
    }
    store_u(x + 3);                           // replaces u(x), last read by w(x + 1)
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      pcur[r] = pnext[r];
      vcur[r] = vnext[r];
      pnext[r] = pahead[r];
      vnext[r] = vahead[r];
    }
    __syncthreads();
  }
}

template <int TZ, int TY, int NT = 256>
static void launch_pair_f32(int batch, int nx, int ny, int nz, const void* u_in, void* u_out, const void* psi,
                            const float2* V, double c1, double c2, double inv_2m, double inv_dx2, int tile_x,
                            int ntiles, hipStream_t s) {
  dim3 grid((nz + TZ - 1) / TZ, (ny + TY - 1) / TY, unsigned(ntiles * batch));
  WFS_LAUNCH((horner2_kernel<float, TZ, TY, NT>), grid, dim3(NT), 0, s, static_cast<const float2*>(u_in),
                     static_cast<float2*>(u_out), static_cast<const float2*>(psi), V, nx, ny, nz, float(c1),
                     float(c2), float(inv_2m), float(inv_dx2), tile_x, ntiles);
}

// variant (complex64): 1 = 16 x 64 (y, z) tile, 2 = 8 x 64, 3 = 8 x 128 with 256 threads, 4 = 8 x 128
// with 512 threads, 5 = 8 x 256 with 1024 threads.  A wider z tile wastes less of the partly used
// memory lines at its z halo, a deeper y tile less y halo; both cost LDS and so blocks per CU.
// More threads per tile give each fewer cells (fewer registers) and the CU more wavefronts to
// hide the plane loads with: 4 and 5 run 16 wavefronts per CU where the LDS of 3 holds 8.
// Measured at 256^3 (one box, alternating): 120, 120, 120-125, 113 and 111 us per launch; the
// plan takes 5 where nz >= 256, 4 where nz >= 128, else 1.
int launch_horner_pair(int precision, int batch, int nx, int ny, int nz, const void* u_in, void* u_out,
                       const void* psi, const float2* V, double c1, double c2, double inv_2m, double inv_dx2,
                       int variant, hipStream_t s) {
  const int tile_x = nx < 32 ? nx : 32;
  const int ntiles = (nx + tile_x - 1) / tile_x;
  if (precision == WFS_C128) {
    constexpr int TZ = 32, TY = 8, NT = 256;
    dim3 grid((nz + TZ - 1) / TZ, (ny + TY - 1) / TY, unsigned(ntiles * batch));
    WFS_LAUNCH((horner2_kernel<double, TZ, TY, NT>), grid, dim3(NT), 0, s,
                       static_cast<const double2*>(u_in), static_cast<double2*>(u_out),
                       static_cast<const double2*>(psi), V, nx, ny, nz, c1, c2, inv_2m, inv_dx2, tile_x, ntiles);
  } else if (variant == 1) {
    launch_pair_f32<64, 16>(batch, nx, ny, nz, u_in, u_out, psi, V, c1, c2, inv_2m, inv_dx2, tile_x, ntiles, s);
  } else if (variant == 3) {
    launch_pair_f32<128, 8>(batch, nx, ny, nz, u_in, u_out, psi, V, c1, c2, inv_2m, inv_dx2, tile_x, ntiles, s);
  } else if (variant == 4) {
    launch_pair_f32<128, 8, 512>(batch, nx, ny, nz, u_in, u_out, psi, V, c1, c2, inv_2m, inv_dx2, tile_x, ntiles, s);
  } else if (variant == 5) {
    launch_pair_f32<256, 8, 1024>(batch, nx, ny, nz, u_in, u_out, psi, V, c1, c2, inv_2m, inv_dx2, tile_x, ntiles, s);
  } else {
    launch_pair_f32<64, 8>(batch, nx, ny, nz, u_in, u_out, psi, V, c1, c2, inv_2m, inv_dx2, tile_x, ntiles, s);
  }
  return int(hipGetLastError());
}

int launch_horner_stage(int precision, int batch, int nx, int ny, int nz, const void* u_in, void* u_out,
                        const void* psi, const float2* V, double dt_over_n, double inv_2m, double inv_dx2,
                        int tile_x, hipStream_t s) {
  // 64 z x 4 y threads per block, each marching 16 x planes: 16x more blocks
  // than whole-column marching (which left one block per CU, latency bound).
  constexpr int TZ = 64, TY = 4;
  if (tile_x <= 0) tile_x = nx < 16 ? nx : 16;
  const int ntiles = (nx + tile_x - 1) / tile_x;
  dim3 grid((nz + TZ - 1) / TZ, (ny + TY - 1) / TY, unsigned(ntiles * batch));
  dim3 block(TZ, TY, 1);
  if (precision == WFS_C128) {
    WFS_LAUNCH((horner_stage_kernel<double, TZ, TY>), grid, block, 0, s,
                       static_cast<const double2*>(u_in), static_cast<double2*>(u_out),
                       static_cast<const double2*>(psi), V, nx, ny, nz, dt_over_n, inv_2m, inv_dx2, tile_x, ntiles);
  } else {
    WFS_LAUNCH((horner_stage_kernel<float, TZ, TY>), grid, block, 0, s,
                       static_cast<const float2*>(u_in), static_cast<float2*>(u_out),
                       static_cast<const float2*>(psi), V, nx, ny, nz, float(dt_over_n), float(inv_2m),
                       float(inv_dx2), tile_x, ntiles);
  }
  return int(hipGetLastError());
}

}  // namespace wfs
