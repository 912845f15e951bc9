"""Input paths of SimState: pre-initialised potential files, raw wave-function files and walls
that move in dynamic-potential mode.

  potential.pre_initialized_potential  sources/potential.py:14-55 (float32, float64, complex64,
                                       complex128 raw files, reshaped to the grid, widened or
                                       rounded to complex64; dynamic mode off, sim_state.py:171-172)
  wave_packet.path                     sources/wave_packet.py:103-105 (raw complex64 / complex128)
  moving walls                         sources/sim_state.py:466-509: every step each wall centre
                                       advects by v dt (FP64), an x centre below 0 stops at 0,
                                       and V is rebuilt from the walls and the absorber
"""

import os

import numpy as np
import pytest
import torch

import oracle
from wavefunctionsimpython_amd import core_sim, initialisation
from wavefunctionsimpython_amd.sim_state import SimState

pytestmark = pytest.mark.gpu

N = 16


def _base_config(tmp, n=N, L=8.0, double=False):
    return {
        "simulation": {"method": "fourier", "delta_time_h_bar_per_hartree": 0.005, "total_iteration_count": 10,
                       "double_precision_calculation": double},
        "wave_packet": {"initial_position_bohr_radii_3": [0.0, 0.0, 0.0],
                        "initial_velocity_bohr_radii_hartree_per_h_bar_3": [0.5, 0.0, 0.0],
                        "initial_standard_deviation_bohr_radii_3": [1.0, 1.0, 1.0]},
        "volume": {"simulated_volume_dimensions_bohr_radii_3": [L, L, L], "number_of_voxels_3": [n, n, n],
                   "observation_box_bottom_corner_bohr_radii_3": [-3.0, -3.0, -3.0],
                   "observation_box_top_corner_bohr_radii_3": [3.0, 3.0, 3.0]},
        "paths": {"cache_dir": os.path.join(tmp, "cache"), "output_dir": os.path.join(tmp, "output")},
    }


@pytest.mark.parametrize("float_type", ["float32", "float64", "complex64", "complex128"])
@pytest.mark.parametrize("cap", [False, True])
def test_preinitialized_potential_file(dev, tmp_path, float_type, cap):
    rng = np.random.default_rng(3)
    dtype = np.dtype(float_type)
    data = rng.standard_normal((N, N, N)) * 3.0
    if dtype.kind == "c":
        data = data - 1j * np.abs(rng.standard_normal((N, N, N)))
    data = data.astype(dtype)
    path = tmp_path / f"potential_{float_type}.bin"
    data.tofile(path)
    cfg = _base_config(str(tmp_path))
    cfg["potential"] = {"pre_initialized_potential": {"path": str(path), "enable": True, "float_type": float_type}}
    cfg["absorbing_boundary_condition"] = {"enable": cap, "start_offset": 0.0}
    sim = SimState(cfg)
    sim.set_use_cache(False)
    sim.initialize_state()
    assert not sim.is_dynamic_potential_mode()          # a pre-initialised V turns dynamic mode off
    V = sim.get_potential().cpu().numpy()
    # the file widened (real types) or rounded (complex128) to complex64, then the absorber added
    want = data.astype(np.complex64)
    if cap:
        dx = sim.get_delta_x_bohr_radii_3()
        imag = oracle.absorbing_potential((N, N, N), dx, [-3.0] * 3, [3.0] * 3, 0.0, [-100.0] * 3, [-100.0] * 3,
                                          [2.0] * 3, [2.0] * 3)
        assert np.array_equal(V.real, want.real)
        assert np.max(np.abs(V.imag - (want.imag + imag))) <= 4e-6 * np.max(np.abs(imag))
    else:
        assert np.array_equal(V, want)
    # the step uses it: one complex64 step against the oracle with the same V
    psi0 = sim.get_wave_function().cpu().numpy()
    sim.evolve_state()
    ref = oracle.split_step(psi0, V, sim.get_delta_x_bohr_radii_3(), 0.005, 1.0, nsteps=1)
    assert oracle.relative_l2(sim.get_wave_function().cpu().numpy(), ref) < 1e-5


@pytest.mark.parametrize("double", [False, True])
def test_wave_packet_raw_file(dev, tmp_path, double):
    rng = np.random.default_rng(4)
    dtype = np.complex128 if double else np.complex64
    data = (rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))).astype(dtype)
    path = tmp_path / "psi.bin"
    data.tofile(path)
    cfg = _base_config(str(tmp_path), double=double)
    cfg["wave_packet"]["path"] = str(path)
    cfg["absorbing_boundary_condition"] = {"enable": False}
    sim = SimState(cfg)
    sim.set_use_cache(False)
    sim.initialize_state()
    got = sim.get_wave_function()
    assert got.dtype == (torch.complex128 if double else torch.complex64)
    assert np.array_equal(got.cpu().numpy(), data)


def _moving_wall_config(tmp):
    cfg = _base_config(tmp, n=64, L=16.0, double=True)
    cfg["simulation"]["total_iteration_count"] = 12
    cfg["wave_packet"]["initial_position_bohr_radii_3"] = [-3.0, 0.0, 0.0]
    cfg["wave_packet"]["initial_velocity_bohr_radii_hartree_per_h_bar_3"] = [2.0, 0.0, 0.0]
    cfg["volume"]["observation_box_bottom_corner_bohr_radii_3"] = [-6.0, -6.0, -6.0]
    cfg["volume"]["observation_box_top_corner_bohr_radii_3"] = [6.0, 6.0, 6.0]
    cfg["absorbing_boundary_condition"] = {"enable": True, "start_offset": 0.0,
                                           "outer_potential_in_positive_xyz_direction_hartree_3": [-30.0] * 3,
                                           "outer_potential_in_negative_xyz_direction_hartree_3": [-30.0] * 3}
    cfg["potential"] = {"walls": [{"potential_hartree": 4.0, "center_bohr_radii_3": [1.0, 0.5, 0.0],
                                   "normal_vector_3": [1.0, 0.0, 0.0], "plateau_thickness_bohr_radii": 1.0,
                                   "slope_thickness_bohr_radii": 0.5, "slope_exponent": 2.0, "slit_count": 1,
                                   "slit_width_bohr_radii": 1.0,
                                   "velocity_bohr_radius_hartree_per_h_bar_3": [-40.0, 10.0, 0.0]}]}
    cfg["measurement"] = {"expected_location": {"enable_image": True}}
    return cfg


def test_moving_wall_advects_clamps_and_regenerates(dev, tmp_path):
    """Step by step the potential equals the wall at the FP64-advected centre (x stopped at 0 from
    step 5 on) plus the absorber, and the complex128 state follows the oracle's steps with those
    potentials; run_iteration (one iteration at a time while the wall moves) gives the same state
    as the reference's time_step loop."""
    cfg = _moving_wall_config(str(tmp_path / "a"))
    sim, tools, it = initialisation.initialize(use_cache=False, config=cfg)
    assert sim.is_dynamic_potential_mode() and not sim.potential_is_static()
    shape = (64, 64, 64)
    dx = sim.get_delta_x_bohr_radii_3()
    dt = sim.get_delta_time_h_bar_per_hartree()
    cap = oracle.absorbing_potential(shape, dx, [-6.0] * 3, [6.0] * 3, 0.0, [-30.0] * 3, [-30.0] * 3, [2.0] * 3,
                                     [2.0] * 3)
    center = np.array([1.0, 0.5, 0.0])
    velocity = np.array([-40.0, 10.0, 0.0])
    psi = sim.get_wave_function().cpu().numpy()
    clamped_at = None
    for k in range(it.total_iteration_count):
        V = sim.get_potential().cpu().numpy()
        wall = oracle.potential_wall(shape, dx, 4.0, center, (1.0, 0.0, 0.0), 1.0, 0.5, 2.0, 1, 1.0, 1.0, 0.0)
        mism = np.count_nonzero(np.abs(V.real - wall) > 1e-5)
        assert mism <= 1e-4 * V.size, f"step {k}: {mism} wall voxels differ (centre {center})"
        assert np.max(np.abs(V.imag - cap)) <= 4e-6 * np.max(np.abs(cap))
        psi = oracle.split_step(psi, V, dx, dt, 1.0, nsteps=1)
        it.i = k
        core_sim.time_step(sim, tools, it)
        # the reference's advection: FP64 centre + v dt, an x centre below zero stops at zero
        center = center + velocity * dt
        if center[0] < 0.0:
            center[0] = 0.0
            clamped_at = k if clamped_at is None else clamped_at
    assert clamped_at is not None and clamped_at < 8
    err = oracle.relative_l2(sim.get_wave_function().cpu().numpy(), psi)
    print(f"moving wall, 12 steps complex128 vs oracle: {err:.2e}; x centre clamped from step {clamped_at + 1}")
    assert err < 1e-12
    # run_iteration keeps one iteration per call while the wall moves: same state bit for bit
    cfg_b = _moving_wall_config(str(tmp_path / "b"))
    sim_b, tools_b, it_b = initialisation.initialize(use_cache=False, config=cfg_b)
    core_sim.run_iteration(sim_b, tools_b, it_b, observe_every=1, exit_on_quit=False)
    assert torch.equal(sim_b.get_wave_function(), sim.get_wave_function())
    ea = tools.get_expected_location().get_expected_location_evolution()
    eb = tools_b.get_expected_location().get_expected_location_evolution()
    assert np.array_equal(ea, eb)


# --- plane-wise regeneration of a moving wall (no reference counterpart: the reference rebuilds
# V whole every step, sources/sim_state.py:466-509; the plane-wise path must give the same V,
# table and state bit for bit) ---

_WALL_ARGS = (2.5, (0.4, 0.2, -0.1), (1.0, 0.0, 0.0), 1.2, 0.6, 2.0, 2, 1.5, 0.4, 0.3)
_CAP_ARGS = ((-3.0, -3.0, -3.0), (3.0, 2.5, 3.5), 0.5, (-20.0, -30.0, -25.0), (-22.0, -18.0, -26.0),
             (2.0, 3.0, 2.5), (1.5, 2.0, 3.0))


def test_wall_and_absorber_on_plane_windows(dev):
    """wfs_add_wall_planes / wfs_add_cap_planes: the window's planes get the whole-grid values bit
    for bit, the other planes are not touched; windows outside the grid are refused."""
    from wavefunctionsimpython_amd import _abi
    shape, dx = (40, 24, 32), (0.3, 0.4, 0.35)
    rng = np.random.default_rng(5)
    V0 = torch.from_numpy((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)).to(dev)
    full = V0.clone()
    _abi.add_wall(shape, dx, *_WALL_ARGS, full)
    _abi.add_cap(shape, dx, *_CAP_ARGS, full)
    assert not torch.equal(full[15:30].real, V0[15:30].real)      # the wall lies in planes 15..29
    for lo, hi in ((0, 40), (15, 30), (7, 19), (0, 1), (39, 40), (12, 12)):
        part = V0.clone()
        _abi.add_wall(shape, dx, *_WALL_ARGS, part, planes=(lo, hi))
        _abi.add_cap(shape, dx, *_CAP_ARGS, part, planes=(lo, hi))
        torch.cuda.synchronize()
        assert torch.equal(part[lo:hi], full[lo:hi]), (lo, hi)
        assert torch.equal(part[:lo], V0[:lo]) and torch.equal(part[hi:], V0[hi:]), (lo, hi)
    for bad in ((-1, 3), (5, 4), (0, 41)):
        with pytest.raises(_abi.WfsError):
            _abi.add_wall(shape, dx, *_WALL_ARGS, V0.clone(), planes=bad)
        with pytest.raises(_abi.WfsError):
            _abi.add_cap(shape, dx, *_CAP_ARGS, V0.clone(), planes=bad)


@pytest.mark.parametrize("shape,precision,slab", [((64, 128, 128), "c64", 1), ((64, 128, 128), "c64", 0),
                                                  ((32, 32, 32), "c128", 0)])
def test_table_planes_rebuilt_equal_the_whole_rebuild(dev, shape, precision, slab):
    """wfs_set_potential_planes after a change of planes [10, 26) steps like a plan whose table was
    built from the new V whole -- including the map of repeated planes (runs that split, merge
    and end at the window's edges)."""
    from wavefunctionsimpython_amd import _abi
    prec = _abi.WFS_C64 if precision == "c64" else _abi.WFS_C128
    dt = torch.complex64 if precision == "c64" else torch.complex128
    rng = np.random.default_rng(11)
    base = (rng.uniform(-1, 1, shape[1:]) - 0.1j * rng.uniform(0, 1, shape[1:])).astype(np.complex64)
    V1 = np.broadcast_to(base, shape).copy()                       # one run of identical planes
    V1[10:14] = base + 0.5                                         # a second run, inside the window
    V1[30] = base - 0.25
    V2 = V1.copy()
    V2[10:14] = base                                               # merges back into the first run
    V2[20:26] = base * 1.5 + 0.1j                                  # a new run ending at the window's edge
    V2[23] = base * 0.5
    V1, V2 = torch.from_numpy(V1).to(dev), torch.from_numpy(V2).to(dev)
    psi0 = torch.from_numpy((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * 0.1).to(dev, dt)
    dx = (0.3, 0.25, 0.25)
    plans = []
    for _ in range(2):
        p = _abi.Plan(shape, dx, 0.01, 1.0, prec, _abi.WFS_METHOD_FOURIER)
        if precision == "c64":
            p.set_variant("yz_slab", slab)
        plans.append(p)
    a, b = plans
    a.set_potential(V1)
    a.set_potential_planes(V2, 10, 26)
    b.set_potential(V2)
    pa, pb = psi0.clone(), psi0.clone()
    a.step(pa, None, 4)
    b.step(pb, None, 4)
    torch.cuda.synchronize()
    assert torch.equal(pa, pb)
    # and again after a second window, on a table the first window already changed
    V3 = V2.clone()
    V3[26:28] = V3[0] * 2.0
    a.set_potential_planes(V3, 26, 28)
    b.set_potential(V3)
    a.step(pa, None, 3)
    b.step(pb, None, 3)
    torch.cuda.synchronize()
    assert torch.equal(pa, pb)
    with pytest.raises(_abi.WfsError):
        a.set_potential_planes(V3, 5, shape[0] + 1)


def _sweeping_wall_config(tmp, n, double):
    L = 16.0 if n <= 64 else 32.0
    cfg = _base_config(tmp, n=n, L=L, double=double)
    cfg["simulation"]["total_iteration_count"] = 10
    cfg["wave_packet"]["initial_position_bohr_radii_3"] = [-2.0, 0.0, 0.0]
    cfg["wave_packet"]["initial_velocity_bohr_radii_hartree_per_h_bar_3"] = [2.0, 0.0, 0.0]
    cfg["volume"]["observation_box_bottom_corner_bohr_radii_3"] = [-0.35 * L] * 3
    cfg["volume"]["observation_box_top_corner_bohr_radii_3"] = [0.35 * L] * 3
    cfg["absorbing_boundary_condition"] = {"enable": True, "start_offset": 0.0}
    cfg["potential"] = {"walls": [
        {"potential_hartree": 5.0, "center_bohr_radii_3": [1.0, 0.2, 0.0], "normal_vector_3": [1.0, 0.0, 0.0],
         "plateau_thickness_bohr_radii": 0.8, "slope_thickness_bohr_radii": 0.4, "slope_exponent": 2.0,
         "slit_count": 2, "slit_spacing_bohr_radii": 1.5, "slit_width_bohr_radii": 0.5,
         "velocity_bohr_radius_hartree_per_h_bar_3": [30.0, 5.0, 0.0]},
        {"potential_hartree": 2.0, "center_bohr_radii_3": [-4.0, 0.0, 0.0], "normal_vector_3": [-1.0, 0.0, 0.0],
         "plateau_thickness_bohr_radii": 0.5, "slope_thickness_bohr_radii": 0.0, "slit_count": 1,
         "slit_width_bohr_radii": 1.0}]}
    return cfg


@pytest.mark.parametrize("n,double", [(64, True), (128, False)])
def test_moving_wall_rebuilds_only_the_planes_it_sweeps(dev, tmp_path, monkeypatch, n, double):
    """A wall moving along its x normal (beside a static one): after the first whole rebuild only
    the planes it covered before or covers after each move are rebuilt, and V and the state equal
    those of whole rebuilds bit for bit (complex128 on the plane pass, complex64 on the slab scheme)."""
    from wavefunctionsimpython_amd import _abi
    calls = []
    original = _abi.Plan.set_potential_planes

    def spy(self, V, lo, hi, stream=None):
        calls.append((lo, hi))
        return original(self, V, lo, hi, stream)

    monkeypatch.setattr(_abi.Plan, "set_potential_planes", spy)
    out = []
    for partial in (True, False):
        sim, tools, it = initialisation.initialize(use_cache=False,
                                                   config=_sweeping_wall_config(str(tmp_path / str(partial)), n, double))
        sim.partial_regeneration = partial
        for k in range(it.total_iteration_count):
            it.i = k
            core_sim.time_step(sim, tools, it)
        out.append((sim.get_wave_function().clone(), sim.get_potential().clone()))
        if partial:
            # iteration 0 rebuilds whole (the first update); iterations 2.. refresh the planes the
            # previous update swept (each evolve_state refreshes before stepping)
            assert len(calls) == it.total_iteration_count - 2
            assert all(0 < hi - lo <= n // 4 for lo, hi in calls), calls
            calls.clear()
    assert not calls
    assert torch.equal(out[0][1], out[1][1])
    assert torch.equal(out[0][0], out[1][0])


@pytest.mark.parametrize("precision,shape", [("c64", (64, 128, 128)), ("c128", (16, 16, 16))])
def test_between_steps_callback_order_and_errors(dev, precision, shape):
    """wfs_step_observe_between calls back after every step in order (slab scheme and the
    per-step form), a callback that queues nothing leaves the result of wfs_step_observe bit for
    bit, and an exception raised in it stops the call and reaches the caller."""
    from wavefunctionsimpython_amd import _abi
    prec = _abi.WFS_C64 if precision == "c64" else _abi.WFS_C128
    dt = torch.complex64 if precision == "c64" else torch.complex128
    rng = np.random.default_rng(2)
    V = torch.from_numpy((rng.uniform(0, 1, shape) - 0.05j).astype(np.complex64)).to(dev)
    psi0 = torch.from_numpy((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * 0.1).to(dev, dt)
    box = [((2, 4, 4), (10, 8, 8))]                              # (bottom voxel, voxel counts)
    p = _abi.Plan(shape, (0.3, 0.25, 0.25), 0.01, 1.0, prec, _abi.WFS_METHOD_FOURIER)
    if precision == "c64":
        p.set_variant("yz_slab", 2)                              # the fused stretch
    p.set_potential(V)
    seen = []
    p.step_observe(psi0.clone(), None, 7, 3, box, between=seen.append)
    assert seen == list(range(7))
    # sampled every step, a callback that queues nothing changes no bit (without the slab scheme
    # the callback form makes one call per step, so the comparison samples every step)
    seen.clear()
    a = psi0.clone()
    p.set_step_index(0)
    oa = p.step_observe(a, None, 7, 1, box, between=seen.append)
    assert seen == list(range(7))
    b = psi0.clone()
    p.set_step_index(0)
    ob = p.step_observe(b, None, 7, 1, box)
    torch.cuda.synchronize()
    assert torch.equal(a, b) and torch.equal(oa, ob)

    class Stop(Exception):
        pass

    def stop_at_2(step):
        if step == 2:
            raise Stop()

    with pytest.raises(Stop):
        p.step_observe(psi0.clone(), None, 5, 1, box, between=stop_at_2)
    torch.cuda.synchronize()


def test_moving_wall_stretches_match_the_reference_loop(dev, tmp_path, monkeypatch):
    """run_iteration steps a wall that moves every step in stretches with update_potential after
    every step inside the call (128^3 complex64, slab scheme): sampled every step, the state and
    the series equal the reference's time_step loop (sources/core_sim.py:19-24, 42-55) bit for
    bit; sampled every 5 steps, the merged kinetic half steps leave it within float rounding."""
    from wavefunctionsimpython_amd import _abi
    used = []
    original = _abi.Plan.step_observe

    def spy(self, *args, between=None, **kw):
        used.append(between is not None)
        return original(self, *args, between=between, **kw)

    monkeypatch.setattr(_abi.Plan, "step_observe", spy)
    runs = {}
    for mode in ("loop", "stretch1", "stretch5"):
        cfg = _sweeping_wall_config(str(tmp_path / mode), 128, False)
        cfg["simulation"]["total_iteration_count"] = 12
        cfg["measurement"] = {"expected_location": {"enable_image": True}}
        sim, tools, it = initialisation.initialize(use_cache=False, config=cfg)
        if mode == "loop":
            for k in range(it.total_iteration_count):
                it.i = k
                core_sim.time_step(sim, tools, it)
        else:
            core_sim.run_iteration(sim, tools, it, observe_every=1 if mode == "stretch1" else 5,
                                   exit_on_quit=False)
        torch.cuda.synchronize()
        runs[mode] = (sim.get_wave_function().clone(), sim.get_potential().clone(),
                      tools.get_expected_location().get_expected_location_evolution())
    assert any(used), "no stretch with a callback ran"
    psi_l, V_l, e_l = runs["loop"]
    psi_1, V_1, e_1 = runs["stretch1"]
    assert torch.equal(V_1, V_l)
    assert torch.equal(psi_1, psi_l)
    assert e_1.shape == e_l.shape and np.allclose(e_1, e_l, rtol=1e-12, atol=1e-13)
    psi_5, V_5, e_5 = runs["stretch5"]
    assert torch.equal(V_5, V_l)
    err = oracle.relative_l2(psi_5.cpu().numpy(), psi_l.cpu().numpy())
    print(f"moving wall, stretches sampled every 5 steps vs one call per step: {err:.2e}")
    assert err < 1e-5
    assert np.allclose(e_5, e_l[::5], atol=1e-5)
