"""Pins of the CPU oracle against mathematics the reference cannot contradict.

The reference has no tests, golden vectors or fixtures for this path (and its
CuPy/CUDA code cannot run here), so every oracle function is pinned by a
closed form, an invariant, a textbook routine or brute force, chosen so that a
plausible slip (dropped term, wrong sign, transposed axis, off-by-one voxel)
fails.  The ``planted`` tests show that such slips are caught.
"""

import numpy as np
import pytest
import scipy.linalg

import oracle

import cap_closed_forms

RNG = np.random.default_rng(1234)


# ------------------------------------------------------------------ geometry
def test_voxel_coordinates_convention():
    r = oracle.voxel_coordinates(64, 60.0 / 64)
    assert r[32] == 0.0
    assert r[0] == -30.0
    assert np.isclose(r[-1], 30.0 - 60.0 / 64)
    r = oracle.voxel_coordinates(7, 1.0)          # odd N: integer N/2 = 3
    assert list(r) == [-3, -2, -1, 0, 1, 2, 3]


def test_wave_numbers_closed_form():
    k = oracle.wave_numbers(8, 0.5)
    m = np.array([0, 1, 2, 3, -4, -3, -2, -1])
    assert np.allclose(k, 2 * np.pi * m / (8 * 0.5))


# ------------------------------------------------------------------ DFT
def _dft_matrix(n, sign):
    j = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(j, j) / n) / np.sqrt(n)


@pytest.mark.parametrize("shape", [(4, 6, 5), (3, 3, 8), (6, 4, 4)])
def test_fft_matches_brute_force_dft(shape):
    a = RNG.standard_normal(shape) + 1j * RNG.standard_normal(shape)
    Fx, Fy, Fz = (_dft_matrix(n, -1) for n in shape)
    ref = np.einsum("ia,jb,kc,abc->ijk", Fx, Fy, Fz, a)
    assert np.allclose(oracle.fftn_ortho(a), ref, atol=1e-13)
    Ix, Iy, Iz = (_dft_matrix(n, +1) for n in shape)
    ref_i = np.einsum("ia,jb,kc,abc->ijk", Ix, Iy, Iz, a)
    assert np.allclose(oracle.ifftn_ortho(a), ref_i, atol=1e-13)


def test_fft_parseval_delta_constant():
    a = RNG.standard_normal((8, 8, 8)) + 1j * RNG.standard_normal((8, 8, 8))
    assert np.isclose(np.linalg.norm(oracle.fftn_ortho(a)), np.linalg.norm(a))
    d = np.zeros((8, 8, 8), complex)
    d[0, 0, 0] = 1
    assert np.allclose(oracle.fftn_ortho(d), 1 / np.sqrt(512))


# ------------------------------------------------------------------ kinetic / potential phases
def test_plane_wave_gains_exact_kinetic_phase():
    shape = (16, 8, 12)
    dx = (0.5, 0.7, 0.4)
    idx = (3, -2, 5)
    k = [oracle.wave_numbers(shape[d], dx[d])[idx[d]] for d in range(3)]
    r = [oracle.voxel_coordinates(shape[d], dx[d]) for d in range(3)]
    psi = np.exp(1j * (k[0] * r[0][:, None, None] + k[1] * r[1][None, :, None] + k[2] * r[2][None, None, :]))
    dt, m, n = 0.013, 1.7, 9
    out = oracle.split_step(psi, np.zeros(shape, np.complex64), dx, dt, m, nsteps=n)
    k2 = k[0] ** 2 + k[1] ** 2 + k[2] ** 2
    assert np.allclose(out, psi * np.exp(-1j * k2 * dt * n / (2 * m)), atol=1e-12)


def test_potential_phase_closed_form():
    V = np.array([[[2.0 - 0.5j, 0.0 + 0.0j]]], dtype=np.complex64)
    P = oracle.potential_phase(V, 0.1)
    assert np.isclose(P[0, 0, 0], np.exp(-1j * 0.2) * np.exp(-0.05))
    assert np.isclose(P[0, 0, 1], 1.0)


# ------------------------------------------------------------------ split step invariants
def _free_packet(n=48, L=48.0, sigma=3.0, k0=1.0):
    dx = [L / n] * 3
    return dx, oracle.gaussian_wave_packet((n,) * 3, dx, [sigma] * 3, [0, 0, 0], [k0, 0, 0])


def _spread_error(split_fn):
    dx, psi = _free_packet()
    b, c = (4, 4, 4), (41, 41, 41)
    worst = 0.0
    for s in range(5):
        er, er2 = oracle.expected_location(psi, dx, b, c)
        sd = oracle.standard_deviation(er, er2)
        t = s * 0.2
        worst = max(worst, abs(er[0] - t), float(np.max(np.abs(sd - np.sqrt(9 + t * t / 36)))))
        psi = split_fn(psi, np.zeros(psi.shape, np.complex64), dx, 0.01, 1.0, nsteps=20)
    return worst


def test_free_packet_follows_analytic_spread():
    assert _spread_error(oracle.split_step) < 1e-7


def test_planted_sign_error_in_kinetic_phase_is_caught(monkeypatch):
    good = oracle.wfs_oracle.kinetic_half_phase
    monkeypatch.setattr(oracle.wfs_oracle, "kinetic_half_phase", lambda *a: np.conj(good(*a)))
    assert _spread_error(oracle.wfs_oracle.split_step) > 1e-2


def test_planted_half_instead_of_quarter_is_caught(monkeypatch):
    good = oracle.wfs_oracle.kinetic_half_phase
    monkeypatch.setattr(oracle.wfs_oracle, "kinetic_half_phase", lambda *a: good(*a) ** 2)
    assert _spread_error(oracle.wfs_oracle.split_step) > 1e-3


def test_split_step_reversible_and_unitary():
    shape = (16, 16, 16)
    dx = (0.5, 0.5, 0.5)
    psi = oracle.gaussian_wave_packet(shape, dx, (1.5, 1.5, 1.5), (0.5, 0, 0), (1.0, 0.5, 0))
    V = (oracle.harmonic_potential(shape, dx, 0.4) + 0.3 * oracle.lattice_potential(shape, dx, 1.0, 2.0)).astype(np.complex64)
    fwd = oracle.split_step(psi, V, dx, 0.02, 1.0, nsteps=7)
    assert np.isclose(np.linalg.norm(fwd), np.linalg.norm(psi), rtol=1e-13)
    back = oracle.split_step(fwd, V, dx, -0.02, 1.0, nsteps=7)
    assert oracle.relative_l2(back, psi) < 1e-12


def test_split_step_norm_decay_with_constant_absorber():
    shape = (8, 8, 8)
    dx = (1.0, 1.0, 1.0)
    psi = oracle.gaussian_wave_packet(shape, dx, (1.5,) * 3, (0, 0, 0), (0, 0, 0))
    V = np.full(shape, -0.5j, dtype=np.complex64)
    out = oracle.split_step(psi, V, dx, 0.1, 1.0, nsteps=4)
    assert np.isclose(np.linalg.norm(out) ** 2, np.linalg.norm(psi) ** 2 * np.exp(-2 * 0.5 * 0.4), rtol=1e-12)


def test_split_step_second_order_in_dt():
    shape = (32, 8, 8)
    dx = (0.4, 1.0, 1.0)
    psi = oracle.gaussian_wave_packet(shape, dx, (1.0, 2.0, 2.0), (0, 0, 0), (1.5, 0, 0))
    V = oracle.lattice_potential(shape, dx, 3.0, 2.5).astype(np.complex64)
    T = 0.4
    ref = oracle.split_step(psi, V, dx, T / 256, 1.0, nsteps=256)
    errs = [oracle.relative_l2(oracle.split_step(psi, V, dx, T / n, 1.0, nsteps=n), ref) for n in (8, 16, 32)]
    ratios = [errs[i] / errs[i + 1] for i in range(2)]
    assert all(3.5 < r < 4.5 for r in ratios)


def test_harmonic_coherent_state_follows_ehrenfest():
    n, L, w = 32, 24.0, 0.5
    dx = [L / n] * 3
    x0, p0 = 2.0, 0.8
    s0 = np.sqrt(1 / (2 * w))      # ground-state width: the coherent state keeps its shape
    psi = oracle.gaussian_wave_packet((n,) * 3, dx, [s0] * 3, [x0, 0, 0], [p0, 0, 0])
    V = oracle.harmonic_potential((n,) * 3, dx, w).astype(np.complex64)
    b, c = (1, 1, 1), (29, 29, 29)
    for step in range(4):
        t = step * 0.5
        er, _ = oracle.expected_location(psi, dx, b, c)
        assert abs(er[0] - (x0 * np.cos(w * t) + p0 / w * np.sin(w * t))) < 2e-3
        psi = oracle.split_step(psi, V, dx, 0.01, 1.0, nsteps=50)


# ------------------------------------------------------------------ power series
def _dense_h(shape, dx, V, mass):
    """Dense H = -(1/2m) L + V with the clamped 7-point Laplacian / dx_x^2 on all axes."""
    nx, ny, nz = shape
    n = nx * ny * nz
    H = np.zeros((n, n), complex)
    idx = lambda x, y, z: (x * ny + y) * nz + z
    for x in range(nx):
        for y in range(ny):
            for z in range(nz):
                i = idx(x, y, z)
                nb = [idx(max(x - 1, 0), y, z), idx(min(x + 1, nx - 1), y, z), idx(x, max(y - 1, 0), z),
                      idx(x, min(y + 1, ny - 1), z), idx(x, y, max(z - 1, 0)), idx(x, y, min(z + 1, nz - 1))]
                for j in nb:
                    H[i, j] += -1.0 / (2 * mass) / dx[0] ** 2
                H[i, i] += 6.0 / (2 * mass) / dx[0] ** 2 + V.reshape(-1)[i]
    return H


def test_power_series_equals_dense_taylor_polynomial():
    shape = (4, 3, 5)
    dx = (0.5, 0.7, 0.9)   # only dx_x is used, on every axis
    V = (RNG.standard_normal(shape) - 0.3j * RNG.random(shape)).astype(np.complex64)
    psi = RNG.standard_normal(shape) + 1j * RNG.standard_normal(shape)
    H = _dense_h(shape, dx, V.astype(complex), 1.3)
    A = -1j * 0.02 * H
    for p in (1, 3, 10):
        acc = np.zeros(psi.size, complex)
        term = psi.reshape(-1).copy()
        acc += term
        for k in range(1, p + 1):
            term = A @ term / k
            acc += term
        got = oracle.power_series_step(psi, V, 0.02, dx, 1.3, order=p)
        assert np.allclose(got.reshape(-1), acc, atol=1e-12)
        assert np.allclose(oracle.power_series_horner(psi, V, 0.02, dx, 1.3, order=p), got, atol=1e-12)
    exact = scipy.linalg.expm(A) @ psi.reshape(-1)
    assert np.allclose(oracle.power_series_step(psi, V, 0.02, dx, 1.3, order=30).reshape(-1), exact, atol=1e-12)


def test_planted_laplacian_axis_spacing_is_caught():
    """Using dx_y on the y axis (a 'fix' of the reference) must not match the reference's dx_x-everywhere H."""
    shape = (4, 3, 5)
    dx = (0.5, 0.7, 0.9)
    s = RNG.standard_normal(shape) + 0j
    H = _dense_h(shape, dx, np.zeros(shape), 0.5)
    lap_ref = -(H @ s.reshape(-1)).reshape(shape)     # (1/2m) L s with m = 0.5
    assert np.allclose(oracle.laplacian_clamped(s, dx[0]), lap_ref, atol=1e-12)
    p = np.pad(s, 1, mode="edge")
    wrong = ((p[:-2, 1:-1, 1:-1] + p[2:, 1:-1, 1:-1] - 2 * s) / dx[0] ** 2
             + (p[1:-1, :-2, 1:-1] + p[1:-1, 2:, 1:-1] - 2 * s) / dx[1] ** 2
             + (p[1:-1, 1:-1, :-2] + p[1:-1, 1:-1, 2:] - 2 * s) / dx[2] ** 2)
    assert not np.allclose(wrong, lap_ref, atol=1e-3)


# ------------------------------------------------------------------ Simpson observables
def test_simpson_weights():
    assert np.allclose(oracle.simpson_weights_1d(5) * 3, [1, 4, 2, 4, 1])
    assert np.allclose(oracle.simpson_weights_1d(1) * 3, [1])
    assert np.allclose(oracle.simpson_weights_1d(3) * 3, [1, 4, 1])


def test_simpson_box_integrals_exact_for_cubics():
    shape = (20, 24, 18)
    dx = (0.5, 0.25, 0.4)
    b, c = (3, 5, 2), (13, 15, 11)
    r = [oracle.voxel_coordinates(shape[d], dx[d]) for d in range(3)]
    # |psi|^2 = (x+6)^2 (y+4) (z+5)^2 is a polynomial of degree <= 2 per axis;
    # times r_d it stays <= 3, times r_d^2 it reaches 4 on that axis only
    px = (r[0] + 6) ** 2
    py = (r[1] + 4)
    pz = (r[2] + 5) ** 2
    rho = px[:, None, None] * py[None, :, None] * pz[None, None, :]
    psi = np.sqrt(rho).astype(complex)
    lo = [r[d][b[d]] for d in range(3)]
    hi = [r[d][b[d] + c[d] - 1] for d in range(3)]

    def I(f_antideriv, d):
        return f_antideriv(hi[d]) - f_antideriv(lo[d])

    Fx = lambda x: (x + 6) ** 3 / 3
    Fy = lambda y: (y + 4) ** 2 / 2
    Fz = lambda z: (z + 5) ** 3 / 3
    P = I(Fx, 0) * I(Fy, 1) * I(Fz, 2)
    assert np.isclose(oracle.volume_probability(psi, dx, b, c), P, rtol=1e-12)
    er, _ = oracle.expected_location(psi, dx, b, c)
    Fxx = lambda x: (x + 6) ** 4 / 4 - 6 * (x + 6) ** 3 / 3      # int x (x+6)^2
    Fyy = lambda y: y ** 3 / 3 + 2 * y ** 2                         # int y (y+4)
    assert np.isclose(er[0], I(Fxx, 0) * I(Fy, 1) * I(Fz, 2), rtol=1e-12)
    assert np.isclose(er[1], I(Fx, 0) * I(Fyy, 1) * I(Fz, 2), rtol=1e-12)


def test_gaussian_moments_and_norm():
    shape = (64, 48, 40)
    dx = (0.25, 0.3, 0.35)
    psi = oracle.gaussian_wave_packet(shape, dx, (1.0, 1.1, 1.0), (0.5, -0.3, 0.2), (1.0, 2.0, -0.5))
    b, c = (0, 0, 0), (63, 47, 39)
    P = oracle.volume_probability(psi, dx, b, c)
    er, er2 = oracle.expected_location(psi, dx, b, c)
    assert np.isclose(P, 1.0, atol=1e-9)
    assert np.allclose(er, [0.5, -0.3, 0.2], atol=1e-9)
    assert np.allclose(oracle.standard_deviation(er, er2), [1.0, 1.1, 1.0], atol=1e-8)
    # phase gradient = k0
    ph = np.angle(psi[33, 24, 20] * np.conj(psi[32, 24, 20]))
    assert np.isclose(ph, 1.0 * 0.25)


def test_box_voxel_rules_hand_worked():
    # the 64^3 free packet: L = 60, N = 64, dx = 0.9375; observation box +-25
    b, t, bv, tv = oracle.observation_box_voxels([60.0] * 3, [64] * 3, [-25.0] * 3, [25.0] * 3)
    assert list(bv) == [5, 5, 5] and list(tv) == [58, 58, 58]
    bE, cnt, tv2 = oracle.expected_location_box(bv, tv)
    assert list(cnt) == [53, 53, 53] and list(tv2) == [58, 58, 58]   # 54 is even -> 53, top +1 -1
    # flipped and clipped corners
    b, t, bv, tv = oracle.observation_box_voxels([60.0] * 3, [64] * 3, [40.0, -25.0, 0.0], [-40.0, 25.0, 1.0])
    assert list(bv) == [0, 5, 32] and list(tv) == [63, 58, 33]
    # volume probability box: bottom -10 -> (20/0.9375)=21.33 -> 21; top 10 -> 42.67 -> 42, +1 -> 43, count 22 -> 21
    bv, cnt = oracle.volume_probability_box([60.0] * 3, [64] * 3, [-10.0] * 3, [10.0] * 3)
    assert list(bv) == [21] * 3 and list(cnt) == [21] * 3


# ------------------------------------------------------------------ potentials
def test_cap_faces_closed_form_and_interior():
    shape = (40, 40, 40)
    dx = (0.5, 0.5, 0.5)
    bottom, top = (-6.0, -6.0, -6.0), (6.0, 6.0, 6.0)
    pp, pn = (-100.0, -20.0, -30.0), (-50.0, -40.0, -10.0)
    ep, en = (3.0, 2.0, 1.5), (2.0, 3.0, 2.5)
    off = 0.5
    cap = oracle.absorbing_potential(shape, dx, bottom, top, off, pp, pn, ep, en)
    r = oracle.voxel_coordinates(40, 0.5)
    inside = np.where((r > -6) & (r < 6))[0]
    sub = np.ix_(inside, inside, inside)
    assert np.all(cap[sub] == 0)
    halfL = 10.0
    for ix in range(40):
        x = r[ix]
        if x > 6:     # +x face at y = z = 0
            want = (max(x - 6 - off, 0) / (halfL - (6 + off))) ** 3 * -100.0
            assert np.isclose(cap[ix, 20, 20], want, rtol=1e-5, atol=1e-5)
        if x < -6:    # -x face
            want = (max(-6 - x - off, 0) / (halfL - (6 + off))) ** 2 * -50.0
            assert np.isclose(cap[ix, 20, 20], want, rtol=1e-5, atol=1e-5)
    # voxels exactly on a box plane get nothing (every branch needs strict inequalities)
    on = np.where(r == -6.0)[0][0]
    assert np.all(cap[on, :, :] == 0) and np.all(cap[:, on, :] == 0)


def test_cap_corner_is_symmetric_for_symmetric_parameters():
    shape = (32, 32, 32)
    dx = (0.5, 0.5, 0.5)
    cap = oracle.absorbing_potential(shape, dx, (-4.0,) * 3, (4.0,) * 3, 0.0, (-10.0,) * 3, (-10.0,) * 3,
                                     (2.0,) * 3, (2.0,) * 3)
    # r(i) = (i - 16) * 0.5: index 16 + k <-> 16 - k mirrors r -> -r
    for (i, j, k) in [(2, 3, 5), (1, 1, 1), (4, 2, 1)]:
        a = cap[16 + 8 + i, 16 + 8 + j, 16 + 8 + k]
        b = cap[16 - 8 - i, 16 - 8 - j, 16 - 8 - k]
        assert a < 0 and np.isclose(a, b, rtol=1e-6)


# The corner and edge blends (absorbing_potential.cu:58-278): closed forms under symmetric
# parameters, one value per side, and the limits at the box planes (tests/cap_closed_forms.py).
_CAP_SYM = dict(shape=(24, 20, 28), dx=(0.5, 0.6, 0.45), depth=3.1, offset=0.4, p=-37.5, e=2.5)
_CAP_SIDE = dict(depth_neg=3.1, depth_pos=4.3, p_neg=-20.0, p_pos=-70.0, e_neg=1.5, e_pos=3.0)
_CAP_EPS = dict(shape=(26, 22, 30), dx=(0.5, 0.55, 0.45), k=6, eps=2e-3, offset=0.3,
                pp=(-60.0, -25.0, -90.0), pn=(-35.0, -80.0, -15.0), ep=(2.0, 3.0, 1.5), en=(2.5, 1.5, 3.5))


def _cap_symmetric_deviation():
    g = _CAP_SYM
    bh = cap_closed_forms.box_half(g["shape"], g["dx"])
    bottom, top = tuple(-(bh - g["depth"])), tuple(bh - g["depth"])
    cap = oracle.absorbing_potential(g["shape"], g["dx"], bottom, top, g["offset"], (g["p"],) * 3, (g["p"],) * 3,
                                     (g["e"],) * 3, (g["e"],) * 3)
    want = cap_closed_forms.symmetric_closed_form(g["shape"], g["dx"], bottom, top, g["offset"], g["p"], g["e"])
    return float(np.max(np.abs(cap - want)) / abs(g["p"])), cap, (bottom, top)


def _cap_per_side_deviation():
    g, q = _CAP_SYM, _CAP_SIDE
    bh = cap_closed_forms.box_half(g["shape"], g["dx"])
    bottom, top = tuple(-(bh - q["depth_neg"])), tuple(bh - q["depth_pos"])
    cap = oracle.absorbing_potential(g["shape"], g["dx"], bottom, top, g["offset"], (q["p_pos"],) * 3,
                                     (q["p_neg"],) * 3, (q["e_pos"],) * 3, (q["e_neg"],) * 3)
    want, mask = cap_closed_forms.per_side_closed_form(g["shape"], g["dx"], bottom, top, g["offset"], q["p_neg"],
                                                       q["p_pos"], q["e_neg"], q["e_pos"])
    return float(np.max(np.abs(cap - want)[mask]) / abs(q["p_pos"]))


def _cap_limit_deviations():
    g = _CAP_EPS
    bottom, top = cap_closed_forms.eps_planes(g["shape"], g["dx"], g["k"], g["eps"])
    cap = oracle.absorbing_potential(g["shape"], g["dx"], bottom, top, g["offset"], g["pp"], g["pn"], g["ep"], g["en"])
    lim = cap_closed_forms.edge_face_limits(cap, g["shape"], g["dx"], bottom, top, g["offset"], g["pp"], g["pn"],
                                            g["ep"], g["en"], g["k"])
    cont = cap_closed_forms.corner_edge_continuity(cap, g["shape"], g["k"])
    return lim, cont


def test_cap_symmetric_parameters_closed_form():
    """Every voxel, including all 8 corners and 12 edges, equals p (max(|d| - off, 0) / maxD)^e with
    the collapsed blends (and the box-half-size / zero entries of the (0,-1,+-1), (0,+1,-1) edges)."""
    dev, cap, (bottom, top) = _cap_symmetric_deviation()
    assert dev < 1e-5
    sides, _ = cap_closed_forms.side_distances(_CAP_SYM["shape"], _CAP_SYM["dx"], bottom, top)
    S = np.stack(np.meshgrid(*sides, indexing="ij")).reshape(3, -1)
    regions = {tuple(c) for c in S.T if 0 < np.count_nonzero(c) and 9 not in c}
    assert len(regions) == 26                       # 6 faces, 12 edges, 8 corners all sampled
    assert np.all(cap[np.all(S == 0, axis=0).reshape(cap.shape)] == 0)


def test_cap_per_side_parameters_closed_form():
    """One potential / exponent / layer depth per side: voxels whose outside axes share a side
    read that side's parameters (a pos/neg swap in a same-side corner or edge fails)."""
    assert _cap_per_side_deviation() < 1e-5


def test_cap_edges_meet_faces_and_corners_meet_edges():
    """Box planes eps = 2e-3 Bohr beyond a voxel plane, distinct parameters per axis and side:
    each edge voxel at eps meets the face ramp its mix() arguments select (or the kernel's own
    limit on the (0,-1,+-1) / (0,+1,-1) edges), each corner voxel at eps meets its edge."""
    lim, cont = _cap_limit_deviations()
    assert len(lim) == 24 and len(cont) == 8
    bad = {k: v for k, v in {**lim, **cont}.items() if v > 1e-2}
    assert not bad, bad


def _plant(branch_region, sides=None, umix=None):
    br = list(oracle.wfs_oracle._CAP_BRANCHES)
    i = [b[0] for b in br].index(branch_region)
    r, s, a, u = br[i]
    br[i] = (r, sides or s, a, umix or u)
    return br


@pytest.mark.parametrize("region,sides,umix", [
    ((1, -1, 1), (1, 1, 1), None),          # corner reads the top plate's y parameters
    ((-1, -1, -1), (-1, -1, 1), None),      # same-side corner, z swapped
    ((-1, 1, 1), (1, 1, 1), None),          # corner, x swapped
    ((1, -1, 0), (-1, -1, 0), None),        # "u" edge, x swapped
    ((-1, 0, 1), (-1, 0, -1), None),        # pillar edge, z swapped
    ((0, -1, 1), (0, 1, 1), None),          # (0,-1,+1) edge, y swapped
    ((0, 1, 1), None, (0.0, 1.0, 1.0)),     # wrong fixed u for the max distance
    ((0, -1, -1), None, (1.0, 1.0, 1.0)),   # the reference's u = 0 "fixed"
    ((0, 1, -1), None, (1.0, 1.0, 1.0)),    # the reference's u = 0 "fixed"
])
def test_planted_cap_branch_slips_are_caught(monkeypatch, region, sides, umix):
    monkeypatch.setattr(oracle.wfs_oracle, "_CAP_BRANCHES", _plant(region, sides, umix))
    lim, cont = _cap_limit_deviations()
    caught = (_cap_symmetric_deviation()[0] > 1e-5 or _cap_per_side_deviation() > 1e-5
              or max(lim.values()) > 1e-2 or max(cont.values()) > 1e-2)
    assert caught


def test_wall_plateau_slope_and_slits():
    shape = (64, 64, 64)
    dx = (0.25, 0.25, 0.25)
    r = oracle.voxel_coordinates(64, 0.25)
    w = oracle.potential_wall(shape, dx, 5.0, (0, 0, 0), (1, 0, 0), 2.0, 1.0, 2.0, 0, 1.0, 0.5, 0.0)
    for ix in range(64):
        d = abs(r[ix])
        want = 5.0 if d <= 1.0 else (5.0 * (1 - (d - 1.0)) ** 2 if d <= 2.0 else 0.0)
        assert np.isclose(w[ix, 10, 10], want, rtol=1e-6)
    # two slits of width 1 centred at y = +-1.5 (spacing 3); right = normalize(n x up) = (0, 0, 1) for n = x,
    # so the slits are cut along z
    s = oracle.potential_wall(shape, dx, 5.0, (0, 0, 0), (1, 0, 0), 2.0, 0.0, 1.0, 2, 3.0, 1.0, 0.0)
    iz = lambda z: int(round(z / 0.25)) + 32
    assert s[32, 10, iz(1.5)] == 0 and s[32, 10, iz(-1.5)] == 0
    assert s[32, 10, iz(0.0)] == np.float32(5.0) and s[32, 10, iz(3.0)] == np.float32(5.0)
    # rotating the slits by pi/2 about the normal moves them to the y axis
    q = oracle.potential_wall(shape, dx, 5.0, (0, 0, 0), (1, 0, 0), 2.0, 0.0, 1.0, 2, 3.0, 1.0, np.pi / 2)
    assert q[32, iz(1.5), 10] == 0 and q[32, iz(0.0), 10] == np.float32(5.0)


def test_harmonic_and_lattice_closed_form():
    shape = (8, 8, 8)
    dx = (0.5, 0.5, 0.5)
    h = oracle.harmonic_potential(shape, dx, 2.0, mass=1.5)
    assert np.isclose(h[6, 4, 2], 0.5 * 1.5 * 4.0 * (1.0 ** 2 + 0 + 1.0 ** 2))
    lat = oracle.lattice_potential(shape, dx, 3.0, 2.0)
    assert np.isclose(lat[4, 4, 4], 9.0) and np.isclose(lat[6, 4, 4], 3.0 * (np.cos(np.pi / 2) ** 2 + 2))


# ------------------------------------------------------------------ analytic and time integrals
def test_analytic_predictions():
    s = oracle.predict_free_space_standard_deviation(0.01, 3.0, 1.0, 201)
    assert np.isclose(s[0], 3.0) and np.isclose(s[200], np.sqrt(9 + 4 / 36))
    x = oracle.predict_free_space_expected_location(0.01, -2.0, 1.5, 101)
    assert np.isclose(x[100], -0.5)


def test_indefinite_simpson_integral_hand_worked_and_exact():
    a = np.array([1.0, 2.0, 5.0])
    assert np.allclose(oracle.indefinite_simpson_integral(a, 0.5), [0, 0.75, 0.5 / 3 * (1 + 8 + 5)])
    assert oracle.indefinite_simpson_integral(np.array([]), 1.0).size == 0
    assert np.allclose(oracle.indefinite_simpson_integral(np.array([3.0]), 1.0), [0.0])
    t = np.arange(21) * 0.1
    f = 2 * t ** 3 - t + 1
    F = lambda x: x ** 4 / 2 - x ** 2 / 2 + x
    got = oracle.indefinite_simpson_integral(f, 0.1)
    assert np.allclose(got[::2], F(t[::2]), atol=1e-12)      # even entries: pure Simpson from 0


# ------------------------------------------------------------------ probability current
def test_plane_current_plane_wave_flux():
    N, h, k, A = 64, 0.25, 0.9, 0.3
    r = oracle.voxel_coordinates(N, h)
    psi = (A * np.exp(1j * k * r))[:, None, None] * np.ones((N, N, N))
    flux, jn = oracle.plane_current(psi, [h] * 3, (0.1, 0.0, 0.0), (1, 0, 0), (6.0, 6.0), (31, 31), 1.0,
                                    single_precision=False)
    # the 6th-order central difference of exp(ikx) is i k_eff exp(ikx), k_eff -> k + O((kh)^6)
    k_eff = (np.sin(3 * k * h) - 9 * np.sin(2 * k * h) + 45 * np.sin(k * h)) / (30 * h)
    assert abs(flux - A * A * k_eff * 36.0) < 1e-12
    assert np.allclose(jn, A * A * k_eff, rtol=1e-12)
    assert abs(k_eff - k) < 1e-6


def test_plane_current_tilted_normal_and_mass():
    N, h = 64, 0.25
    kx, ky, A, m = 0.7, -0.4, 0.5, 1.7
    r = oracle.voxel_coordinates(N, h)
    psi = A * np.exp(1j * (kx * r[:, None, None] + ky * r[None, :, None])) * np.ones((1, 1, N))
    n = np.array([0.6, 0.8, 0.0])
    flux, jn = oracle.plane_current(psi, [h] * 3, (0.0, 0.0, 0.0), n, (4.0, 4.0), (17, 17), m,
                                    single_precision=False)
    assert np.allclose(jn, A * A * (kx * n[0] + ky * n[1]) / m, rtol=1e-5)
    assert abs(flux - A * A * (kx * n[0] + ky * n[1]) / m * 16.0) < 1e-5


def test_plane_current_exact_for_degree_six_polynomials():
    N, h = 40, 0.1
    x = oracle.voxel_coordinates(N, h)
    a = 1.0 + 0.3 * x ** 2 - 0.05 * x ** 6
    b = 0.2 * x + 0.01 * x ** 3 + 0.002 * x ** 5
    psi = (a + 1j * b)[:, None, None] * np.ones((N, N, N))
    da = 0.6 * x - 0.3 * x ** 5
    db = 0.2 + 0.03 * x ** 2 + 0.01 * x ** 4
    want = (a * db - b * da)          # Im(conj(psi) psi') for m = 1
    flux, jn = oracle.plane_current(psi, [h] * 3, (0.3, 0.0, 0.0), (1, 0, 0), (1.0, 1.0), (5, 5), 1.0,
                                    single_precision=False)
    ix = int(round(0.3 / h)) + N // 2
    assert np.allclose(jn, want[ix], rtol=1e-10)


def test_plane_sample_count_rule():
    assert list(oracle.plane_sample_count([255, 512])) == [255, 513]


def test_plane_current_gaussian_flux_at_matched_pitch():
    """With the pixel pitch equal to the voxel pitch the flux of a Gaussian packet through x = 0
    is rho_x(0) * v; a larger pitch makes the truncated-voxel sampling skip the centre voxel
    (reference behaviour) and the integral falls short."""
    N, L = 64, 20.0
    h = L / N
    psi = oracle.gaussian_wave_packet((N,) * 3, [h] * 3, (1.5,) * 3, (-2.0, 0, 0), (2.0, 0, 0))
    rho_x0 = np.exp(-4.0 / (2 * 1.5 ** 2)) / (np.sqrt(2 * np.pi) * 1.5)
    matched, _ = oracle.plane_current(psi, [h] * 3, (0, 0, 0), (1, 0, 0), (19.375, 19.375), (63, 63), 1.0, False)
    assert abs(matched - 2.0 * rho_x0) < 2e-4
    coarse, _ = oracle.plane_current(psi, [h] * 3, (0, 0, 0), (1, 0, 0), (19.0, 19.0), (61, 61), 1.0, False)
    assert coarse < 0.9 * matched
