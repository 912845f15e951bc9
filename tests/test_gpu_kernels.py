"""Parity of the HIP library (through its C ABI) against the CPU oracle.

Tolerances: relative L2 <= 1e-12 in complex128 mode and <= 1e-5 in complex64
mode for the propagators (the FP32 horizon is stated per test), 1e-12 for the
FP64-accumulated observables, a few float32 ulps for the float32 generators.
"""

import numpy as np
import pytest
import torch

import oracle
from wavefunctionsimpython_amd import _abi

import cap_closed_forms

pytestmark = pytest.mark.gpu


def _rand_state(shape, seed, dtype):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return a.astype(dtype)


def _to_dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _plan(shape, dx, dt, mass, prec, method=_abi.WFS_METHOD_FOURIER, order=10):
    return _abi.Plan(shape, dx, dt, mass, prec, method, order)


FFT_SHAPES = [(8, 8, 8), (16, 32, 8), (64, 64, 64), (96, 96, 96), (32, 192, 16), (128, 8, 16), (256, 16, 8),
              (512, 8, 8), (8, 8, 512)]


@pytest.mark.parametrize("shape", FFT_SHAPES)
@pytest.mark.parametrize("prec", [_abi.WFS_C128, _abi.WFS_C64])
def test_fft3d_matches_scipy(dev, shape, prec):
    dtype = np.complex128 if prec == _abi.WFS_C128 else np.complex64
    a = _rand_state(shape, 1, dtype)
    p = _plan(shape, (0.5, 0.5, 0.5), 0.01, 1.0, prec)
    t = _to_dev(a, dev)
    p.fft3d(t, -1)
    got = t.cpu().numpy()
    ref = oracle.fftn_ortho(a.astype(np.complex128))
    tol = 1e-13 if prec == _abi.WFS_C128 else 2e-6
    assert oracle.relative_l2(got, ref) < tol
    p.fft3d(t, +1)
    back = t.cpu().numpy()
    assert oracle.relative_l2(back, a) < 2 * tol


def test_fft3d_delta_and_constant(dev):
    shape = (16, 16, 16)
    a = np.zeros(shape, np.complex128)
    a[0, 0, 0] = 1.0
    p = _plan(shape, (1, 1, 1), 0.01, 1.0, _abi.WFS_C128)
    t = _to_dev(a, dev)
    p.fft3d(t, -1)
    got = t.cpu().numpy()
    assert np.allclose(got, 1.0 / np.sqrt(16 ** 3), atol=1e-15)


def _free_packet_64_state(prec):
    N, L = 64, 60.0
    dx = [L / N] * 3
    dtype = np.complex128 if prec == _abi.WFS_C128 else np.complex64
    psi = oracle.gaussian_wave_packet((N,) * 3, dx, [3, 3, 3], [0, 0, 0], [1, 0, 0], dtype)
    return (N,) * 3, dx, psi


def test_split_step_free_packet_c128(dev):
    """64^3 complex128 free packet, 200 steps of dt = 0.01 (the free-packet workload)."""
    shape, dx, psi0 = _free_packet_64_state(_abi.WFS_C128)
    V = np.zeros(shape, np.complex64)
    p = _plan(shape, dx, 0.01, 1.0, _abi.WFS_C128)
    t = _to_dev(psi0, dev)
    Vd = _to_dev(V, dev)
    p.step(t, Vd, 200)
    got = t.cpu().numpy()
    ref = oracle.split_step(psi0, V, dx, 0.01, 1.0, nsteps=200)
    assert oracle.relative_l2(got, ref) < 1e-12


def _harmonic_cap_potential(shape, dx):
    L = np.array(shape) * np.array(dx)
    V = oracle.harmonic_potential(shape, dx, 0.3).astype(np.complex64)
    cap = oracle.absorbing_potential(shape, dx, -0.35 * L, 0.35 * L, 1.0, [-5.0] * 3, [-5.0] * 3, [2.0] * 3,
                                     [2.0] * 3)
    V.imag += cap
    return V


@pytest.mark.parametrize("prec,nsteps,tol", [(_abi.WFS_C128, 40, 1e-12), (_abi.WFS_C64, 100, 1e-5)])
def test_split_step_potential(dev, prec, nsteps, tol):
    """Harmonic + CAP potential; FP32 horizon 100 steps at 48^3... here 64^3."""
    shape = (64, 64, 64)
    dx = (0.5, 0.5, 0.5)
    dtype = np.complex128 if prec == _abi.WFS_C128 else np.complex64
    psi0 = oracle.gaussian_wave_packet(shape, dx, [2.0, 2.5, 3.0], [1.0, -0.5, 0.0], [0.7, 0.0, -0.4], dtype)
    V = _harmonic_cap_potential(shape, dx)
    p = _plan(shape, dx, 0.02, 1.0, prec)
    t = _to_dev(psi0, dev)
    p.step(t, _to_dev(V, dev), nsteps)
    got = t.cpu().numpy()
    ref = oracle.split_step(psi0, V, dx, 0.02, 1.0, nsteps=nsteps)
    assert oracle.relative_l2(got, ref) < tol


@pytest.mark.parametrize("shape", [(32, 16, 8), (16, 96, 32), (8, 8, 8)])
def test_split_step_one_step_shapes(dev, shape):
    dx = (0.4, 0.5, 0.6)
    psi0 = _rand_state(shape, 3, np.complex128)
    rng = np.random.default_rng(4)
    V = (rng.standard_normal(shape) + 1j * -np.abs(rng.standard_normal(shape))).astype(np.complex64)
    for nsteps in (1, 2, 3):
        p = _plan(shape, dx, 0.03, 1.3, _abi.WFS_C128)
        t = _to_dev(psi0, dev)
        p.step(t, _to_dev(V, dev), nsteps)
        ref = oracle.split_step(psi0, V, dx, 0.03, 1.3, nsteps=nsteps)
        assert oracle.relative_l2(t.cpu().numpy(), ref) < 1e-12


@pytest.mark.parametrize("prec,tol", [(_abi.WFS_C128, 1e-12), (_abi.WFS_C64, 1e-5)])
@pytest.mark.parametrize("order", [1, 2, 10])
def test_power_series_matches_recurrence(dev, prec, tol, order):
    shape = (32, 16, 16)
    dx = (0.5, 0.45, 0.6)
    dtype = np.complex128 if prec == _abi.WFS_C128 else np.complex64
    psi0 = oracle.gaussian_wave_packet(shape, dx, [2.0, 2.0, 2.0], [0.5, 0.0, 0.0], [1.0, 0.5, 0.0], dtype)
    V = _harmonic_cap_potential(shape, dx)
    p = _plan(shape, dx, 0.005, 1.0, prec, _abi.WFS_METHOD_POWER_SERIES, order)
    t = _to_dev(psi0, dev)
    p.step(t, _to_dev(V, dev), 5)
    got = t.cpu().numpy()
    ref = oracle.power_series_step(psi0, V, 0.005, dx, 1.0, order=order, nsteps=5)
    assert oracle.relative_l2(got, ref) < tol


def test_observables_match_oracle(dev):
    shape = (32, 32, 64)
    dx = (0.3, 0.4, 0.5)
    psi = _rand_state(shape, 7, np.complex128)
    p = _plan(shape, dx, 0.01, 1.0, _abi.WFS_C128)
    boxes = [((1, 2, 3), (29, 27, 61)), ((0, 0, 0), (32, 32, 64)), ((5, 5, 5), (1, 3, 5))]
    out = p.observables(_to_dev(psi, dev), boxes).cpu().numpy()
    for i, (b, c) in enumerate(boxes):
        P = oracle.volume_probability(psi, dx, b, c)
        er, er2 = oracle.expected_location(psi, dx, b, c)
        assert abs(out[i, 0] - P) <= 1e-12 * abs(P)
        assert np.allclose(out[i, 1:4], er, rtol=1e-11, atol=1e-12 * abs(P))
        assert np.allclose(out[i, 4:7], er2, rtol=1e-11, atol=1e-12 * abs(P))


def test_observables_complex64(dev):
    shape = (16, 16, 16)
    dx = (0.5, 0.5, 0.5)
    psi = _rand_state(shape, 8, np.complex64)
    p = _plan(shape, dx, 0.01, 1.0, _abi.WFS_C64)
    out = p.observables(_to_dev(psi, dev), [((0, 0, 0), (15, 15, 15))]).cpu().numpy()
    P = oracle.volume_probability(psi, dx, (0, 0, 0), (15, 15, 15))
    assert abs(out[0, 0] - P) <= 1e-12 * P


def test_observables_reject_box_outside(dev):
    p = _plan((8, 8, 8), (1, 1, 1), 0.01, 1.0, _abi.WFS_C64)
    t = torch.zeros((8, 8, 8), dtype=torch.complex64, device=dev)
    with pytest.raises(ValueError):
        p.observables(t, [((1, 0, 0), (8, 1, 1))])


@pytest.mark.parametrize("prec", [_abi.WFS_C128, _abi.WFS_C64])
def test_gaussian_generator(dev, prec):
    shape = (32, 16, 64)
    dx = (0.5, 0.7, 0.3)
    dtype = np.complex128 if prec == _abi.WFS_C128 else np.complex64
    t = torch.empty(shape, dtype=torch.complex128 if prec == _abi.WFS_C128 else torch.complex64, device=dev)
    _abi.init_gaussian(shape, dx, prec, (2.0, 3.0, 1.5), (1.0, -2.0, 0.5), (0.5, -1.0, 2.0), t)
    ref = oracle.gaussian_wave_packet(shape, dx, (2.0, 3.0, 1.5), (1.0, -2.0, 0.5), (0.5, -1.0, 2.0), np.complex128)
    tol = 1e-14 if prec == _abi.WFS_C128 else 1e-7
    assert oracle.relative_l2(t.cpu().numpy(), ref) < tol


@pytest.mark.parametrize("shape,bottom,top,offset", [
    ((64, 64, 64), (-10.0, -10.0, -10.0), (10.0, 10.0, 10.0), 0.0),
    ((32, 32, 16), (-4.0, -3.0, -2.5), (5.0, 2.0, 1.0), 1.0),
])
def test_cap_generator(dev, shape, bottom, top, offset):
    dx = (0.5, 0.5, 0.5)
    V = torch.zeros(shape, dtype=torch.complex64, device=dev)
    args = (bottom, top, offset, (-100.0, -10.0, -20.0), (-50.0, -30.0, -10.0), (3.0, 2.0, 2.5), (2.0, 3.0, 1.5))
    _abi.add_cap(shape, dx, *args, V)
    got = V.cpu().numpy()
    ref = oracle.absorbing_potential(shape, dx, *args)
    assert np.all(got.real == 0)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(got.imag - ref)) <= 4e-6 * scale


def test_cap_kernel_corner_and_edge_pins(dev):
    """The GPU CAP at the oracle's pinning geometries: symmetric-parameter closed form for every
    voxel, per-side closed form, and the edge/face and corner/edge limits at eps = 2e-3 Bohr
    (the same bounds as the oracle pins), plus agreement with the oracle itself."""
    g = dict(shape=(24, 20, 28), dx=(0.5, 0.6, 0.45), offset=0.4)
    bh = cap_closed_forms.box_half(g["shape"], g["dx"])

    def gpu_cap(shape, dx, bottom, top, offset, pp, pn, ep, en):
        V = torch.zeros(shape, dtype=torch.complex64, device=dev)
        _abi.add_cap(shape, dx, bottom, top, offset, pp, pn, ep, en, V)
        out = V.cpu().numpy()
        assert np.all(out.real == 0)
        ref = oracle.absorbing_potential(shape, dx, bottom, top, offset, pp, pn, ep, en)
        assert np.max(np.abs(out.imag - ref)) <= 4e-6 * np.max(np.abs(ref))
        return out.imag

    bottom, top = tuple(-(bh - 3.1)), tuple(bh - 3.1)
    cap = gpu_cap(g["shape"], g["dx"], bottom, top, g["offset"], (-37.5,) * 3, (-37.5,) * 3, (2.5,) * 3, (2.5,) * 3)
    want = cap_closed_forms.symmetric_closed_form(g["shape"], g["dx"], bottom, top, g["offset"], -37.5, 2.5)
    assert np.max(np.abs(cap - want)) / 37.5 < 1e-5

    bottom, top = tuple(-(bh - 3.1)), tuple(bh - 4.3)
    cap = gpu_cap(g["shape"], g["dx"], bottom, top, g["offset"], (-70.0,) * 3, (-20.0,) * 3, (3.0,) * 3, (1.5,) * 3)
    want, mask = cap_closed_forms.per_side_closed_form(g["shape"], g["dx"], bottom, top, g["offset"], -20.0, -70.0,
                                                       1.5, 3.0)
    assert np.max(np.abs(cap - want)[mask]) / 70.0 < 1e-5

    shape, dx, k, off = (26, 22, 30), (0.5, 0.55, 0.45), 6, 0.3
    pp, pn, ep, en = (-60.0, -25.0, -90.0), (-35.0, -80.0, -15.0), (2.0, 3.0, 1.5), (2.5, 1.5, 3.5)
    bottom, top = cap_closed_forms.eps_planes(shape, dx, k, 2e-3)
    cap = gpu_cap(shape, dx, bottom, top, off, pp, pn, ep, en)
    lim = cap_closed_forms.edge_face_limits(cap, shape, dx, bottom, top, off, pp, pn, ep, en, k)
    cont = cap_closed_forms.corner_edge_continuity(cap, shape, k)
    bad = {key: v for key, v in {**lim, **cont}.items() if v > 1e-2}
    assert not bad, bad


@pytest.mark.parametrize("slits,normal,rot", [(0, (1, 0, 0), 0.0), (1, (1, 0, 0), 0.0), (2, (1, 0, 0), 0.0),
                                              (3, (0, 1, 0), 0.3), (4, (0.6, 0.8, 0.0), 0.0)])
def test_wall_generator(dev, slits, normal, rot):
    shape = (64, 64, 64)
    dx = (0.3125, 0.3125, 0.3125)
    V = torch.zeros(shape, dtype=torch.complex64, device=dev)
    args = (5.0, (0.3, 0.1, 0.0), normal, 2.0, 1.0, 1.5, slits, 6.0, 1.5, rot)
    _abi.add_wall(shape, dx, *args, V)
    got = V.cpu().numpy()
    ref = oracle.potential_wall(shape, dx, *args)
    assert np.all(got.imag == 0)
    mism = np.abs(got.real - ref) > 1e-5
    # float32 ties at cell boundaries may flip a handful of voxels
    assert mism.sum() <= 1e-4 * got.size


def test_harmonic_and_lattice(dev):
    shape = (32, 32, 32)
    dx = (0.5, 0.5, 0.5)
    V = torch.zeros(shape, dtype=torch.complex64, device=dev)
    _abi.add_harmonic(shape, dx, 0.2, 1.0, (0, 0, 0), V)
    _abi.add_lattice(shape, dx, 2.0, 4.0, V)
    ref = oracle.harmonic_potential(shape, dx, 0.2) + oracle.lattice_potential(shape, dx, 2.0, 4.0)
    assert np.allclose(V.cpu().numpy().real, ref, rtol=2e-7, atol=2e-6)


def test_operator_tables(dev):
    shape = (16, 32, 8)
    dx = (0.5, 0.3, 0.7)
    kin = torch.empty(shape, dtype=torch.complex128, device=dev)
    _abi.kinetic_operator(shape, dx, 0.05, 1.5, _abi.WFS_C128, kin)
    ref = oracle.kinetic_half_phase(shape, dx, 0.05, 1.5)
    assert np.max(np.abs(kin.cpu().numpy() - ref)) < 1e-13
    rng = np.random.default_rng(2)
    V = (rng.standard_normal(shape) + 1j * -np.abs(rng.standard_normal(shape))).astype(np.complex64)
    out = torch.empty(shape, dtype=torch.complex128, device=dev)
    _abi.potential_operator(_to_dev(V, dev), 0.05, _abi.WFS_C128, out)
    assert np.max(np.abs(out.cpu().numpy() - oracle.potential_phase(V, 0.05))) < 1e-14


@pytest.mark.parametrize("prec,tol", [(_abi.WFS_C128, 1e-11), (_abi.WFS_C64, 2e-5)])
@pytest.mark.parametrize("normal,center,size,count", [
    ((1.0, 0.0, 0.0), (0.3, 0.0, 0.0), (10.0, 8.0), (41, 33)),
    ((0.6, 0.8, 0.0), (-0.5, 0.4, 0.2), (6.0, 6.0), (25, 25)),
    ((0.0, 1.0, 0.0), (0.0, 0.2, 0.0), (30.0, 30.0), (63, 63)),   # reaches outside the grid
])
def test_plane_current_matches_oracle(dev, prec, tol, normal, center, size, count):
    shape = (64, 32, 64)
    dx = (0.3, 0.45, 0.3)
    dtype = np.complex128 if prec == _abi.WFS_C128 else np.complex64
    psi = oracle.gaussian_wave_packet(shape, dx, (2.0, 2.0, 2.5), (-0.5, 0.3, 0.0), (1.2, 0.8, -0.4), dtype)
    n = np.array(normal) / np.linalg.norm(normal)
    p = _plan(shape, dx, 0.01, 1.3, prec)
    dens = torch.empty(count, dtype=torch.float64 if prec == _abi.WFS_C128 else torch.float32, device=dev)
    out = p.plane_current(_to_dev(psi, dev), center, n, size, count, density=dens)
    flux_ref, jn_ref = oracle.plane_current(psi, dx, center, n, size, count, 1.3,
                                            single_precision=(prec == _abi.WFS_C64))
    got = float(out.cpu().numpy()[0])
    assert abs(got - flux_ref) <= tol * max(abs(flux_ref), 1e-3)
    assert np.allclose(dens.cpu().numpy(), jn_ref, rtol=tol * 10, atol=tol * np.max(np.abs(jn_ref)))
