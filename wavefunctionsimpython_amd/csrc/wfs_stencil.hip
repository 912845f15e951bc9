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
    const float2 v = V[i];
    const C ps = psi[i];
    // L u * dx^2 in the reference's summation order
    C lap;
    lap.x = prev.x + next.x + uym.x - R(6) * cur.x + uyp.x + uzm.x + uzp.x;
    lap.y = prev.y + next.y + uym.y - R(6) * cur.y + uyp.y + uzm.y + uzp.y;
    // H u = -(1/2m) L u + V u
    C hu;
    hu.x = -inv_2m * inv_dx2 * lap.x + (R(v.x) * cur.x - R(v.y) * cur.y);
    hu.y = -inv_2m * inv_dx2 * lap.y + (R(v.x) * cur.y + R(v.y) * cur.x);
    // u_out = psi + (-i dt/n) H u = psi + (dt/n) * (Hu.y, -Hu.x)
    C o;
    o.x = ps.x + dt_over_n * hu.y;
    o.y = ps.y - dt_over_n * hu.x;
    u_out[i] = o;
    prev = cur;
    cur = next;
  }
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
    hipLaunchKernelGGL((horner_stage_kernel<double, TZ, TY>), grid, block, 0, s,
                       static_cast<const double2*>(u_in), static_cast<double2*>(u_out),
                       static_cast<const double2*>(psi), V, nx, ny, nz, dt_over_n, inv_2m, inv_dx2, tile_x, ntiles);
  } else {
    hipLaunchKernelGGL((horner_stage_kernel<float, TZ, TY>), grid, block, 0, s,
                       static_cast<const float2*>(u_in), static_cast<float2*>(u_out),
                       static_cast<const float2*>(psi), V, nx, ny, nz, float(dt_over_n), float(inv_2m),
                       float(inv_dx2), tile_x, ntiles);
  }
  return int(hipGetLastError());
}

}  // namespace wfs
