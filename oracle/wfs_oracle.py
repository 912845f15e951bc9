"""FP64 numpy/scipy restatement of the reference's propagation path.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` as the checker; never shipped in,
called by or linked into the product path.

Every function cites the reference file:line it restates (paths relative to the
reference repository root).  How each function is pinned (closed forms,
invariants, textbook library routines, brute force) is listed in the table
below; the pins live in ``tests/test_oracle_*.py``.

  function                       pinned by
  -----------------------------  ----------------------------------------------
  voxel_coordinates              closed form r(N/2)=0, last = L/2-dx
  wave_numbers                   numpy.fft.fftfreq identity, brute-force DFT
  gaussian_wave_packet           norm, <r>=r0, sigma, phase gradient = k0
  fftn_ortho / ifftn_ortho       brute-force O(N^2) DFT on 4^3..6^3, Parseval
  kinetic_half_phase             single plane wave gains exp(-i k^2 t / 2m)
  potential_phase                pointwise closed form, norm decay exp(2 dt ImV)
  split_step                     analytic free packet (<x>, sigma), reversibility,
                                 unitarity, 2nd-order convergence, Ehrenfest
  power_series_step              dense-H brute force (matrix powers), expm limit,
                                 Horner identity
  power_series_horner            equal to power_series_step to rounding (p = 1, 3, 10)
  laplacian_clamped              dense 7-point matrix with edge-replicated faces; a
                                 per-axis-spacing "fix" is caught
  simpson_weights_1d / volume_probability / expected_location
                                 exact for x^a y^b z^c (a,b,c<=3), Gaussian moments
  standard_deviation             Gaussian sigma from its first two moments
  plane_current / plane_sample_count
                                 plane wave flux |A|^2 k/m (tilted normal, mass),
                                 exact for degree-6 polynomials, Gaussian flux at
                                 matched pitch, odd-count rule
  observation_box_voxels / volume_probability_box / expected_location_box
                                 hand-worked voxel examples
  absorbing_potential            faces: 1-D polynomial ramp closed form; corners
                                 and edges: closed form for symmetric and for
                                 per-side parameters (every mix() collapses),
                                 edge->face and corner->edge limits at box planes
                                 eps away, planted per-branch side/mix slips caught
  potential_wall                 plateau/slope closed form, slit geometry at
                                 points away from cell-boundary ties
  harmonic_potential / lattice_potential   closed form (no reference counterpart)
  predict_free_space_*           closed form (doi:10.1088/0143-0807/18/3/022)
  indefinite_simpson_integral    exact for cubics; hand-worked small n
  relative_l2                    definition (used as the comparison, not compared)
"""

from __future__ import annotations

import math
import os

import numpy as np
import scipy.fft as sfft

__all__ = [
    "voxel_coordinates", "wave_numbers", "gaussian_wave_packet",
    "fftn_ortho", "ifftn_ortho", "kinetic_half_phase", "potential_phase",
    "split_step", "laplacian_clamped", "power_series_step",
    "power_series_horner", "simpson_weights_1d", "observation_box_voxels",
    "volume_probability_box", "expected_location_box", "volume_probability",
    "expected_location", "standard_deviation", "absorbing_potential",
    "potential_wall", "harmonic_potential", "lattice_potential",
    "predict_free_space_standard_deviation",
    "predict_free_space_expected_location", "indefinite_simpson_integral",
    "relative_l2", "plane_current", "plane_sample_count",
]

_WORKERS = int(os.environ.get("WFS_ORACLE_WORKERS", "1"))


# ---------------------------------------------------------------------------
# Geometry
# ---------------------------------------------------------------------------

def voxel_coordinates(n: int, dx: float) -> np.ndarray:
    """r(i) = (i - n//2) * dx with integer n//2.

    sources/cuda_kernels/gaussian_wave_packet.cu:36-40 and
    sources/cuda_kernels/expected_location.cu:33-37 (integer N/2).
    """
    return (np.arange(n, dtype=np.int64) - n // 2).astype(np.float64) * float(dx)


def wave_numbers(n: int, dx: float) -> np.ndarray:
    """k = 2*pi*fftfreq(n, dx) -- sources/sim_state.py:275-278."""
    return 2.0 * np.pi * np.fft.fftfreq(n, d=float(dx))


def gaussian_wave_packet(shape, dx, sigma, r0, k0, dtype=np.complex128) -> np.ndarray:
    """Separable Gaussian times plane wave.

    sources/cuda_kernels/common.cu:565-605 (a = 2*sigma,
    g_d = (2/(pi a_d^2))^(1/4) exp(-(r_d-r0_d)^2/a_d^2), times exp(i k0.r)) with
    the voxel coordinates of sources/cuda_kernels/gaussian_wave_packet.cu:36-40.
    k0 = m * v (sources/wave_packet.py:76-79, sources/math_utils.py:67-68).
    Evaluated in FP64; the result is rounded to ``dtype``.
    """
    g = []
    ph = []
    for d in range(3):
        r = voxel_coordinates(shape[d], dx[d])
        a = 2.0 * float(sigma[d])
        g.append((2.0 / np.pi / a / a) ** 0.25 * np.exp(-(r - r0[d]) ** 2 / a / a))
        ph.append(float(k0[d]) * r)
    amp = g[0][:, None, None] * g[1][None, :, None] * g[2][None, None, :]
    phase = ph[0][:, None, None] + ph[1][None, :, None] + ph[2][None, None, :]
    return (amp * np.exp(1j * phase)).astype(dtype)


# ---------------------------------------------------------------------------
# Split-operator Fourier step
# ---------------------------------------------------------------------------

def fftn_ortho(a: np.ndarray, workers: int | None = None) -> np.ndarray:
    """Unitary 3-D DFT, e^{-2 pi i jk/N} / sqrt(N) per axis.

    Stands in for cupy.fft.fftn(norm="ortho") at sources/sim_state.py:646,677.
    """
    return sfft.fftn(a, norm="ortho", workers=workers or _WORKERS)


def ifftn_ortho(a: np.ndarray, workers: int | None = None) -> np.ndarray:
    """Unitary inverse 3-D DFT -- sources/sim_state.py:665,696."""
    return sfft.ifftn(a, norm="ortho", workers=workers or _WORKERS)


def kinetic_half_phase(shape, dx, dt, mass) -> np.ndarray:
    """exp(i * theta), theta = -|k|^2 * dt / (4 m): a half kinetic step.

    sources/cuda_kernels/kinetic_operator.cu:16-23 (hBar = 1).
    """
    kx, ky, kz = (wave_numbers(shape[d], dx[d]) for d in range(3))
    k2 = kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2
    return np.exp(1j * (-k2 * float(dt) / 4.0 / float(mass)))


def potential_phase(V: np.ndarray, dt: float) -> np.ndarray:
    """cexp_i(-dt*V) = (cos(-dt ReV), sin(-dt ReV)) * exp(dt ImV).

    sources/cuda_kernels/potential_operator.cu:10-15 and
    sources/cuda_kernels/common.cu:108-111.  V is complex64 in the reference
    (potential_operator.cu:7); it is widened to FP64 here.
    """
    ang = -float(dt) * V.astype(np.complex128)
    return (np.cos(ang.real) + 1j * np.sin(ang.real)) * np.exp(-ang.imag)


def split_step(psi, V, dx, dt, mass, nsteps: int = 1, workers: int | None = None, phases=None):
    """Strang step psi <- F^-1 K/2 F . V . F^-1 K/2 F psi, repeated nsteps times.

    Follows sources/sim_state.py:641-697 step for step: four unitary FFTs, two
    half-kinetic multiplies and one potential multiply per step.  ``phases``
    may carry precomputed (kinetic_half_phase, potential_phase) arrays.
    """
    psi = np.asarray(psi, dtype=np.complex128).copy()
    if phases is None:
        kh = kinetic_half_phase(psi.shape, dx, dt, mass)
        pv = potential_phase(np.asarray(V), dt)
    else:
        kh, pv = phases
    for _ in range(nsteps):
        psi = fftn_ortho(psi, workers)
        psi *= kh
        psi = ifftn_ortho(psi, workers)
        psi *= pv
        psi = fftn_ortho(psi, workers)
        psi *= kh
        psi = ifftn_ortho(psi, workers)
    return psi


# ---------------------------------------------------------------------------
# Power series exp(-i H dt)
# ---------------------------------------------------------------------------

def laplacian_clamped(s: np.ndarray, dx_x: float) -> np.ndarray:
    """7-point Laplacian with edge-replicated faces, divided by dx_x^2 on ALL axes.

    sources/cuda_kernels/power_series_operator.cu:23-46 (neighbour index falls
    back to the voxel itself at a face) and :50-58 (sum / delta_x / delta_x).
    """
    p = np.pad(s, 1, mode="edge")
    c = p[1:-1, 1:-1, 1:-1]
    lap = (p[:-2, 1:-1, 1:-1] + p[2:, 1:-1, 1:-1] + p[1:-1, :-2, 1:-1]
           - 6.0 * c + p[1:-1, 2:, 1:-1] + p[1:-1, 1:-1, :-2] + p[1:-1, 1:-1, 2:])
    return lap / float(dx_x) / float(dx_x)


def power_series_step(psi, V, dt, dx, mass, order: int = 10, nsteps: int = 1):
    """Truncated Taylor series of exp(-i H dt), the reference recurrence.

    sources/sim_state.py:721-752 and
    sources/cuda_kernels/power_series_operator.cu:60-64:
      s_0 = psi;  s_n = (i dt / n) [ (1/2m) L s_{n-1} - V s_{n-1} ];  psi += s_n.
    """
    psi = np.asarray(psi, dtype=np.complex128).copy()
    v = np.asarray(V).astype(np.complex128)
    for _ in range(nsteps):
        s = psi.copy()
        for n in range(1, order + 1):
            s = (1j * float(dt) / n) * ((1.0 / 2.0 / float(mass)) * laplacian_clamped(s, dx[0]) - v * s)
            psi += s
    return psi


def power_series_horner(psi, V, dt, dx, mass, order: int = 10):
    """Same polynomial as power_series_step, evaluated by Horner's rule.

    u = psi; for n = order..1: u = psi + (-i dt H / n) u.  Algebraically equal
    to the recurrence (pinned by tests); this is the form the GPU stencil uses.
    """
    psi = np.asarray(psi, dtype=np.complex128)
    v = np.asarray(V).astype(np.complex128)
    u = psi.copy()
    for n in range(order, 0, -1):
        hu = -(1.0 / 2.0 / float(mass)) * laplacian_clamped(u, dx[0]) + v * u
        u = psi + (-1j * float(dt) / n) * hu
    return u


# ---------------------------------------------------------------------------
# Observables
# ---------------------------------------------------------------------------

def simpson_weights_1d(count: int) -> np.ndarray:
    """Composite Simpson weights over ``count`` samples (÷3 included).

    sources/cuda_kernels/common.cu:466-492: 1 at both ends, 4 at odd and 2 at
    even interior positions (position counted from the box start).
    """
    w = np.ones(count, dtype=np.float64)
    i = np.arange(count)
    interior = (i > 0) & (i < count - 1)
    w[interior & (i % 2 == 0)] = 2.0
    w[interior & (i % 2 == 1)] = 4.0
    return w / 3.0


def _phys_to_voxel(pos, L, dx):
    """((pos + L/2) / dx).astype(uint) -- sources/sim_state.py:354-358."""
    return ((np.asarray(pos, dtype=np.float64) + np.asarray(L, dtype=np.float64) * 0.5)
            / np.asarray(dx, dtype=np.float64)).astype(np.uint64)


def observation_box_voxels(L, N, bottom_bohr, top_bohr):
    """Observation box in voxels -- sources/sim_state.py:105-161.

    Flip inverted corners, clip to +-L/2, int32 truncation of (p + L/2)/dx,
    then top >= N -> N-1 and bottom < 0 -> 0.  Returns
    (bottom_bohr, top_bohr, bottom_voxel, top_voxel).
    """
    L = np.asarray(L, dtype=np.float64)
    N = np.asarray(N)
    b = np.array(bottom_bohr, dtype=np.float64)
    t = np.array(top_bohr, dtype=np.float64)
    for i in range(3):
        if b[i] > t[i]:
            b[i], t[i] = t[i], b[i]
    for i in range(3):
        if b[i] < -L[i] * 0.5:
            b[i] = -L[i] * 0.5
        if t[i] > L[i] * 0.5:
            t[i] = L[i] * 0.5
    dx = L / N
    bv = np.array((b + L * 0.5) / dx, dtype=np.int32)
    tv = np.array((t + L * 0.5) / dx, dtype=np.int32)
    for i in range(3):
        if tv[i] >= N[i]:
            tv[i] = N[i] - 1
        if bv[i] < 0:
            bv[i] = 0
    return b, t, bv, tv


def volume_probability_box(L, N, bottom_bohr, top_bohr):
    """Voxel box of a VolumeProbability -- sources/measurement.py:86-116.

    bottom = voxel(bottom); top = voxel(top) + 1; count = top - bottom, an even
    count is reduced by one.  Not clipped to the grid (reference behaviour).
    Returns (bottom_voxel, sample_count) as int64 arrays.
    """
    b = np.array(bottom_bohr, dtype=np.float64)
    t = np.array(top_bohr, dtype=np.float64)
    for i in range(3):
        if b[i] > t[i]:
            b[i], t[i] = t[i], b[i]
    dx = np.asarray(L, dtype=np.float64) / np.asarray(N)
    bv = _phys_to_voxel(b, L, dx)
    tv = _phys_to_voxel(t, L, dx) + np.uint64(1)
    count = (tv - bv).astype(np.int64)
    for i in range(3):
        if count[i] % 2 == 0:
            count[i] -= 1
    return bv.astype(np.int64), count


def expected_location_box(obs_bottom_voxel, obs_top_voxel):
    """ExpectedLocation sample box -- sources/measurement.py:320-329.

    count = (top + 1) - bottom, reduced by one when even.  Returns
    (bottom_voxel, sample_count, mutated_top_voxel): the reference mutates the
    SimState top voxel in place (+1, then -1 when the count was even).
    """
    bv = np.array(obs_bottom_voxel, dtype=np.int64)
    tv = np.array(obs_top_voxel, dtype=np.int64) + 1
    count = tv - bv
    for i in range(3):
        if count[i] % 2 == 0:
            count[i] -= 1
            tv[i] -= 1
    return bv, count, tv


def _box_view(a, bottom, count):
    b = [int(x) for x in bottom]
    c = [int(x) for x in count]
    return a[b[0]:b[0] + c[0], b[1]:b[1] + c[1], b[2]:b[2] + c[2]]


def volume_probability(psi, dx, bottom_voxel, sample_count) -> float:
    """Simpson-weighted |psi|^2 dV over the box.

    sources/cuda_kernels/volume_probability.cu:29-33 (w * |psi|^2 * dx dy dz),
    weights from sources/cuda_kernels/common.cu:466-492.
    """
    sub = _box_view(psi, bottom_voxel, sample_count)
    if sub.shape != tuple(int(c) for c in sample_count):
        raise IndexError("box reaches outside the grid (the reference reads out of bounds here)")
    rho = (sub.real.astype(np.float64) ** 2 + sub.imag.astype(np.float64) ** 2)
    wx, wy, wz = (simpson_weights_1d(int(c)) for c in sample_count)
    dv = float(dx[0]) * float(dx[1]) * float(dx[2])
    return float(np.einsum("i,j,k,ijk->", wx, wy, wz, rho) * dv)


def expected_location(psi, dx, bottom_voxel, sample_count):
    """(E[r], E[r^2]) componentwise over the box, not divided by the norm.

    sources/cuda_kernels/expected_location.cu:33-37 (r = dr*(v + bottom - N/2)),
    :45-46 and :61-62 (r and r*r times w*|psi|^2*dV).
    """
    N = psi.shape
    sub = _box_view(psi, bottom_voxel, sample_count)
    if sub.shape != tuple(int(c) for c in sample_count):
        raise IndexError("box reaches outside the grid")
    rho = (sub.real.astype(np.float64) ** 2 + sub.imag.astype(np.float64) ** 2)
    w = [simpson_weights_1d(int(c)) for c in sample_count]
    dv = float(dx[0]) * float(dx[1]) * float(dx[2])
    r = [float(dx[d]) * (np.arange(int(sample_count[d])) + int(bottom_voxel[d]) - N[d] // 2)
         for d in range(3)]
    wr = np.einsum("i,j,k,ijk->ijk", w[0], w[1], w[2], rho) * dv
    m0 = [wr.sum(axis=(1, 2)), wr.sum(axis=(0, 2)), wr.sum(axis=(0, 1))]
    e_r = np.array([np.dot(r[d], m0[d]) for d in range(3)])
    e_r2 = np.array([np.dot(r[d] * r[d], m0[d]) for d in range(3)])
    return e_r, e_r2


def standard_deviation(e_r, e_r2):
    """sqrt(E[r^2] - E[r]^2) per axis -- sources/measurement.py:394-400."""
    return np.sqrt(np.asarray(e_r2) - np.asarray(e_r) ** 2)


def plane_sample_count(resolution):
    """resolution + 1, an even count reduced by one -- sources/measurement.py:218-223."""
    c = np.array(resolution, dtype=np.int64) + 1
    return np.where(c % 2 == 0, c - 1, c)


def plane_current(psi, dx, center, normal, size, count, mass, single_precision=True):
    """Flux of j = Im(conj(psi) grad psi)/m through a sampled plane, and the j.n map.

    sources/cuda_kernels/probability_current_density.cu: in-plane axes (:41-49,
    float32), pixel positions (:56-57, float32), voxel = trunc(r/dx + N/2)
    (:61-62, state precision), 6th-order central difference whose k-step
    neighbour falls back to the centre voxel off the grid (:68-176),
    Re[(-1/2m) i (conj(psi) dpsi - psi conj(dpsi))] (:184-192), 2-D Simpson
    weights / 9 times dW dH (:198-202, common.cu:525-543).  Pixels whose voxel
    lies outside the grid contribute 0.  Returns (flux, density[cu, cv]).
    """
    f = np.float32
    cu, cv = int(count[0]), int(count[1])
    n = np.array(normal, dtype=f)
    pu = np.array([0, 1, 0], dtype=f)
    if abs(n[0] * pu[0] + n[1] * pu[1] + n[2] * pu[2]) > 0.99:
        pu[1] = f(0.0)
        pu[2] = f(1.0)
    r = np.array([n[1] * pu[2] - n[2] * pu[1], n[2] * pu[0] - n[0] * pu[2], n[0] * pu[1] - n[1] * pu[0]], dtype=f)
    r = r / np.sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2])
    u = np.array([r[1] * n[2] - r[2] * n[1], r[2] * n[0] - r[0] * n[2], r[0] * n[1] - r[1] * n[0]], dtype=f)
    u = u / np.sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2])
    if single_precision:
        dW = f(size[0]) / f(cu - 1)
        dH = f(size[1]) / f(cv - 1)
        T = np.float32
    else:
        dW = float(size[0]) / (cu - 1)
        dH = float(size[1]) / (cv - 1)
        T = np.float64
    si = (np.arange(cu) - cu // 2).astype(f)[:, None]
    sj = (np.arange(cv) - cv // 2).astype(f)[None, :]
    c = np.array(center, dtype=f)
    N = psi.shape
    vox = []
    inside = np.ones((cu, cv), dtype=bool)
    for a in range(3):
        fr = (c[a] + r[a] * f(dW) * si) + u[a] * f(dH) * sj
        fv = fr.astype(T) / T(dx[a]) + T(0.5) * T(N[a])
        inside &= (fv >= 0) & (fv < N[a])
        vox.append(np.where(inside, fv, 0).astype(np.int64))
    vox = [np.where(inside, v, 0) for v in vox]
    p = psi.astype(np.complex128)
    p0 = p[vox[0], vox[1], vox[2]]
    dpsi = np.zeros((cu, cv), dtype=np.complex128)
    for a in range(3):
        vals = {}
        for k in (1, 2, 3):
            for sgn in (-1, 1):
                idx = [v.copy() for v in vox]
                moved = idx[a] + sgn * k
                ok = (moved >= 0) & (moved < N[a])
                idx[a] = np.where(ok, moved, idx[a])
                vals[sgn * k] = np.where(ok, p[idx[0], idx[1], idx[2]], p0)
        g = (vals[3] - 9 * vals[2] + 45 * vals[1] - 45 * vals[-1] + 9 * vals[-2] - vals[-3]) / 60.0 / float(dx[a])
        dpsi += float(n[a]) * g
    jn = np.where(inside, (np.conj(p0) * dpsi).imag / float(mass), 0.0)
    wu = simpson_weights_1d(cu) * 3
    wv = simpson_weights_1d(cv) * 3
    flux = float(np.einsum("i,j,ij->", wu, wv, jn) / 9.0 * float(dW) * float(dH))
    return flux, jn


# ---------------------------------------------------------------------------
# Potentials (float32, as the reference kernels compute them)
# ---------------------------------------------------------------------------

_f = np.float32


def _mix(xyz, u, v):
    """common.cu:394-397: z*v + (1-v)*(y*u + x*(1-u)), float32."""
    one = _f(1.0)
    return xyz[2] * v + (one - v) * (xyz[1] * u + xyz[0] * (one - u))


def _len3(a, b, c):
    """common.cu:215-218 length(): sqrtf(x*x + y*y + z*z) in float32."""
    return np.sqrt(a * a + b * b + c * c)


# Corner and edge branches of absorbing_potential.cu in the kernel's order (the first
# branch whose region holds wins).  Per entry:
#   region  per axis -1 (below the box bottom), +1 (above its top), 0 (strictly inside)
#   sides   per axis the side (-1 bottom / +1 top) whose corner, offset, exponent and
#           outer potential the branch reads; the kernel reads the region's own side
#   angle   "u": u = atan(|d_y| / |d_x|) * 2/pi blends x (u=0) with y (u=1), v = 0;
#           "v": v = atan(|d_z| / |d_other|) * 2/pi blends the other outside axis
#           (v=0) with z (v=1); "uv": corners, u as above and
#           v = atan(|d_z| / sqrt(d_x^2 + d_y^2)) * 2/pi
#   umix    for "v" branches the fixed u the kernel passes to mix() for the max
#           distance, the exponent and the potential (an inside axis enters the max
#           distance as the box half-size and the exponent and potential as 0)
_CAP_BRANCHES = [
    # corners (absorbing_potential.cu:58-155)
    ((-1, -1, -1), (-1, -1, -1), "uv", None),       # :59-70
    ((-1, -1, 1), (-1, -1, 1), "uv", None),         # :71-82
    ((1, -1, 1), (1, -1, 1), "uv", None),           # :83-94
    ((1, -1, -1), (1, -1, -1), "uv", None),         # :95-106
    ((1, 1, 1), (1, 1, 1), "uv", None),             # :108-119
    ((1, 1, -1), (1, 1, -1), "uv", None),           # :120-131
    ((-1, 1, -1), (-1, 1, -1), "uv", None),         # :132-143
    ((-1, 1, 1), (-1, 1, 1), "uv", None),           # :144-155
    # bottom edges (:156-196)
    ((-1, -1, 0), (-1, -1, 0), "u", None),          # :157-166
    ((0, -1, 1), (0, -1, 1), "v", (0.0, 1.0, 1.0)),    # :167-176
    ((1, -1, 0), (1, -1, 0), "u", None),            # :177-186
    ((0, -1, -1), (0, -1, -1), "v", (0.0, 1.0, 1.0)),  # :187-196
    # pillar edges (:197-237)
    ((-1, 0, -1), (-1, 0, -1), "v", (0.0, 0.0, 0.0)),  # :198-207
    ((-1, 0, 1), (-1, 0, 1), "v", (0.0, 0.0, 0.0)),    # :208-217
    ((1, 0, 1), (1, 0, 1), "v", (0.0, 0.0, 0.0)),      # :218-227
    ((1, 0, -1), (1, 0, -1), "v", (0.0, 0.0, 0.0)),    # :228-237
    # top plane edges (:238-278)
    ((-1, 1, 0), (-1, 1, 0), "u", None),            # :239-248
    ((0, 1, 1), (0, 1, 1), "v", (1.0, 1.0, 1.0)),      # :249-258
    ((1, 1, 0), (1, 1, 0), "u", None),              # :259-268
    ((0, 1, -1), (0, 1, -1), "v", (0.0, 0.0, 0.0)),    # :269-278
]

# faces (:279-303) in the kernel's order: (axis, side)
_CAP_FACES = [(1, -1), (2, 1), (0, 1), (2, -1), (0, -1), (1, 1)]


def absorbing_potential(shape, dx, bottom, top, offset, pot_pos, pot_neg, exp_pos, exp_neg):
    """Imaginary part added by the complex absorbing potential (float32).

    sources/cuda_kernels/absorbing_potential.cu:38-306.  r = dr*v - 0.5*N*dr
    (:52); d = r - bottom below the box, r - top above it.  Faces (:279-303) are
    1-D ramps p*(max(|d| - offset, 0) / maxD)^e with maxD = box_half - (|corner| +
    offset); corners (:58-155) and edges (:156-278) blend maxD, e and p of the
    axes involved with mix() (common.cu:394-397) at atan angles, including the
    reference's own fixed mix() arguments on the "v" edges (e.g. u=0 for maxD at
    :172, :192, :274, and u=0 for the exponent and potential at :275-276), and
    use the length of d over the outside axes.  Returns a float32 array of the
    per-voxel potential (to be added as +i*pot).
    """
    pi_f = _f(np.pi)
    two = _f(2.0)
    zero = _f(0.0)
    dr = np.array(dx, dtype=_f)
    Nf = np.array(shape, dtype=_f)
    b = np.array(bottom, dtype=_f)
    t = np.array(top, dtype=_f)
    off = _f(offset)
    pp = np.array(pot_pos, dtype=_f)
    pn = np.array(pot_neg, dtype=_f)
    ep = np.array(exp_pos, dtype=_f)
    en = np.array(exp_neg, dtype=_f)
    half = _f(0.5)
    bh = half * Nf * dr
    r = []
    for d in range(3):
        v = np.arange(shape[d], dtype=_f)
        r.append(dr[d] * v - half * Nf[d] * dr[d])
    R = np.broadcast_arrays(r[0][:, None, None], r[1][None, :, None], r[2][None, None, :])
    dB = [R[d] - b[d] for d in range(3)]
    dT = [R[d] - t[d] for d in range(3)]
    # region tests per axis and side, all strict (a voxel on a box plane matches no branch)
    cond = {(d, -1): dB[d] < 0 for d in range(3)}
    cond.update({(d, 1): dT[d] > 0 for d in range(3)})
    cond.update({(d, 0): (dB[d] > 0) & (dT[d] < 0) for d in range(3)})
    out = np.zeros(shape, dtype=_f)
    done = np.zeros(shape, dtype=bool)

    def ang(num, den):
        return np.arctan(num / den) / pi_f * two

    with np.errstate(all="ignore"):
        for region, sides, angle, umix in _CAP_BRANCHES:
            m = cond[(0, region[0])] & cond[(1, region[1])] & cond[(2, region[2])] & ~done
            if not m.any():
                continue
            # d on the outside axes (the side of the region), 0 on the inside axis
            dd = [(dB[a] if region[a] < 0 else dT[a])[m] if region[a] else np.zeros(int(m.sum()), dtype=_f)
                  for a in range(3)]
            md3 = np.array([bh[a] - (np.abs(b[a] if sides[a] < 0 else t[a]) + off) if sides[a] else bh[a]
                            for a in range(3)], dtype=_f)
            e3 = np.array([(en[a] if sides[a] < 0 else ep[a]) if sides[a] else 0.0 for a in range(3)], dtype=_f)
            p3 = np.array([(pn[a] if sides[a] < 0 else pp[a]) if sides[a] else 0.0 for a in range(3)], dtype=_f)
            if angle == "uv":
                u = ang(np.abs(dd[1]), np.abs(dd[0]))
                v = ang(np.abs(dd[2]), np.sqrt(dd[0] * dd[0] + dd[1] * dd[1]))
                q = [(u, v)] * 3
            elif angle == "u":
                u = ang(np.abs(dd[1]), np.abs(dd[0]))
                q = [(u, zero)] * 3
            else:
                other = 0 if region[0] else 1
                v = ang(np.abs(dd[2]), np.abs(dd[other]))
                q = [(_f(umix[k]), v) for k in range(3)]
            maxD = _mix(md3, *q[0])
            ex = _mix(e3, *q[1])
            mp = _mix(p3, *q[2])
            out[m] = np.power(np.maximum(_len3(dd[0], dd[1], dd[2]) - off, zero) / maxD, ex) * mp
            done |= m
        for axis, side in _CAP_FACES:
            m = ~done
            for a in range(3):
                m &= cond[(a, side if a == axis else 0)]
            if not m.any():
                continue
            dist = (-dB[axis] if side < 0 else dT[axis])[m]
            maxD = bh[axis] - (np.abs(b[axis] if side < 0 else t[axis]) + off)
            e = en[axis] if side < 0 else ep[axis]
            p = pn[axis] if side < 0 else pp[axis]
            out[m] = np.power(np.maximum(dist - off, zero) / maxD, e) * p
            done |= m
    return out


def _rotate_vector_f32(v, axis, rad):
    """common.cu:348-369: quaternion rotation matrix, float32."""
    h = _f(rad) / _f(2.0)
    q0 = np.cos(h, dtype=_f)
    s = np.sin(h, dtype=_f)
    q1, q2, q3 = s * axis[0], s * axis[1], s * axis[2]
    two = _f(2.0)
    Q = np.array([
        [q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3, two * (q1 * q2 - q0 * q3), two * (q1 * q3 + q0 * q2)],
        [two * (q2 * q1 + q0 * q3), q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3, two * (q2 * q3 - q0 * q1)],
        [two * (q3 * q1 - q0 * q2), two * (q3 * q2 + q0 * q1), q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3],
    ], dtype=_f)
    return np.array([Q[i, 0] * v[0] + Q[i, 1] * v[1] + Q[i, 2] * v[2] for i in range(3)], dtype=_f)


def potential_wall(shape, dx, potential, center, normal, plateau, slope, slope_exponent,
                   slit_count, slit_spacing, slit_width, slit_rotation):
    """Real potential added by one wall (float32).

    sources/cuda_kernels/potential_wall.cu:30-82: r = dr*(v - 0.5 N) (:37),
    d = |n.(r-c)| (:38); in-plane 'right' = normalize(n x up), up=(0,1,0) or,
    when |n.up| > 0.99, up with x=0, z=1 (:39-43); rotated about n (:45);
    tangential distance dR (:46); 1-4 slits punch holes (:47-75); plateau
    and polynomial slope (:76-82).  ``normal`` must already be unit length
    (sources/potential.py:180 normalises in FP64 on the host).
    """
    dr = np.array(dx, dtype=_f)
    Nf = np.array(shape, dtype=_f)
    c = np.array(center, dtype=_f)
    n = np.array(normal, dtype=_f)
    half = _f(0.5)
    r = [dr[d] * (np.arange(shape[d], dtype=_f) - half * Nf[d]) for d in range(3)]
    X, Y, Z = np.broadcast_arrays(r[0][:, None, None], r[1][None, :, None], r[2][None, None, :])
    ex, ey, ez = X - c[0], Y - c[1], Z - c[2]
    d = np.abs(n[0] * ex + n[1] * ey + n[2] * ez)
    up = np.array([0.0, 1.0, 0.0], dtype=_f)
    if abs(n[0] * up[0] + n[1] * up[1] + n[2] * up[2]) > _f(0.99):
        up[0] = _f(0.0)
        up[2] = _f(1.0)
    cr = np.array([n[1] * up[2] - n[2] * up[1], n[2] * up[0] - n[0] * up[2],
                   n[0] * up[1] - n[1] * up[0]], dtype=_f)
    right = cr / np.sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2])
    right = _rotate_vector_f32(right, n, slit_rotation)
    dR = np.abs(right[0] * ex + right[1] * ey + right[2] * ez)
    w = _f(slit_width)
    sp = _f(slit_spacing)
    hole = np.zeros(shape, dtype=bool)
    if slit_count == 1:
        hole = dR < w * half
    elif slit_count == 2:
        hole = (dR > (sp - w) * half) & (dR < (sp + w) * half)
    elif slit_count == 3:
        hole = (dR < w * half) | ((dR > sp - w * half) & (dR < sp + w * half))
    elif slit_count == 4:
        hole = (((dR > (sp - w) * half) & (dR < (sp + w) * half))
                | ((dR > sp * _f(1.5) - w * half) & (dR < sp * _f(1.5) + w * half)))
    pl = _f(plateau) * half
    sl = _f(slope)
    out = np.zeros(shape, dtype=_f)
    inside = (d <= pl) & ~hole
    out[inside] = _f(potential)
    with np.errstate(all="ignore"):
        sm = (d > pl) & (d <= pl + sl) & ~hole
        out[sm] = _f(potential) * np.power(_f(1.0) - (d[sm] - pl) / sl, _f(slope_exponent))
    return out


def harmonic_potential(shape, dx, omega, mass=1.0, center=(0.0, 0.0, 0.0)):
    """V = 1/2 m w^2 |r - c|^2 on the voxel coordinates (FP64).

    The formula of the uncalled helper sources/potential.py:414-423 (which
    has no config key in the reference); used by the harmonic and lattice
    workloads.
    """
    r = [voxel_coordinates(shape[d], dx[d]) - center[d] for d in range(3)]
    r2 = r[0][:, None, None] ** 2 + r[1][None, :, None] ** 2 + r[2][None, None, :] ** 2
    return 0.5 * float(mass) * float(omega) ** 2 * r2


def lattice_potential(shape, dx, depth, period):
    """V = depth * sum_d cos^2(pi r_d / period) (FP64).

    No reference counterpart (the reference's optical grid, potential.py:478-505,
    is a Gaussian-node grid); closed form.
    """
    out = np.zeros(shape, dtype=np.float64)
    for d in range(3):
        r = voxel_coordinates(shape[d], dx[d])
        c = np.cos(np.pi * r / float(period)) ** 2
        sh = [1, 1, 1]
        sh[d] = shape[d]
        out = out + float(depth) * c.reshape(sh)
    return out


# ---------------------------------------------------------------------------
# Analytic predictions and time integrals
# ---------------------------------------------------------------------------

def predict_free_space_standard_deviation(delta_t, sigma0, mass, step_count):
    """sigma(t) = sqrt(a^2 + 4 t^2 / (m^2 a^2)) / 2, a = 2 sigma0.

    sources/math_utils.py:237-251.
    """
    a = 2.0 * float(sigma0)
    t = np.arange(step_count, dtype=np.float64) * float(delta_t)
    return np.sqrt(a * a + 4.0 * (1.0 / float(mass) ** 2) * t * t / (a * a)) / 2.0


def predict_free_space_expected_location(delta_t, x0, velocity, step_count):
    """<x>(t) = x0 + v t -- sources/math_utils.py:254-267."""
    t = np.arange(step_count, dtype=np.float64) * float(delta_t)
    return float(x0) + t * float(velocity)


def indefinite_simpson_integral(array, dt):
    """Running Simpson integral -- sources/math_utils.py:199-234.

    Entry 1 is the trapezoid (a0+a1)/2; odd and even chains then advance by
    Simpson panels; the whole array is finally scaled by dt/3 (entry 1 is
    pre-multiplied by 3 to compensate).
    """
    a = np.asarray(array, dtype=np.float64)
    n = a.size
    if n == 0:
        return np.empty(0, dtype=np.float64)
    if n == 1:
        return np.array([0.0])
    if n == 2:
        return np.array([0.0, (a[0] + a[1]) / 2.0])
    odd_section = n - 1 if n % 2 == 0 else n
    even_section = n if n % 2 == 0 else n - 1
    out = np.zeros(n, dtype=np.float64)
    out[1] = 3.0 * (a[0] + a[1]) / 2.0
    for i in range(2, odd_section, 2):
        out[i] = out[i - 2] + a[i - 2] + 4.0 * a[i - 1] + a[i]
    for i in range(3, even_section, 2):
        out[i] = out[i - 2] + a[i - 2] + 4.0 * a[i - 1] + a[i]
    return out * float(dt) / 3.0


def relative_l2(a, ref) -> float:
    """||a - ref||_2 / ||ref||_2 in FP64."""
    a = np.asarray(a).astype(np.complex128)
    ref = np.asarray(ref).astype(np.complex128)
    return float(np.linalg.norm((a - ref).ravel()) / np.linalg.norm(ref.ravel()))
