"""The reference's call surface (SimState, MeasurementTools, core_sim) on the GPU."""

import copy
import os

import numpy as np
import pytest
import torch

import oracle
from wavefunctionsimpython_amd import _abi, core_sim, initialisation, math_utils, operators
from wavefunctionsimpython_amd.iter_data import IterData
from wavefunctionsimpython_amd.measurement import MeasurementTools
from wavefunctionsimpython_amd.sim_state import SimState, SimulationMethod

pytestmark = pytest.mark.gpu


def free_packet_64_config(tmp):
    """64^3 complex128 free packet (the free-packet workload)."""
    return {
        "simulation": {"method": "fourier", "delta_time_h_bar_per_hartree": 0.01, "total_iteration_count": 200,
                       "double_precision_calculation": True, "is_dynamic_potential_mode": False},
        "wave_packet": {"particle_mass_electron_rest_mass": 1.0, "initial_position_bohr_radii_3": [0.0, 0.0, 0.0],
                        "initial_velocity_bohr_radii_hartree_per_h_bar_3": [1.0, 0.0, 0.0],
                        "initial_standard_deviation_bohr_radii_3": [3.0, 3.0, 3.0]},
        "volume": {"simulated_volume_dimensions_bohr_radii_3": [60.0, 60.0, 60.0], "number_of_voxels_3": [64, 64, 64],
                   "observation_box_bottom_corner_bohr_radii_3": [-25.0, -25.0, -25.0],
                   "observation_box_top_corner_bohr_radii_3": [25.0, 25.0, 25.0]},
        "absorbing_boundary_condition": {"enable": False},
        "measurement": {"expected_location": {"enable_image": True},
                        "volume_probabilities": [{"name": "left", "bottom_corner_bohr_radii_3": [-20.0, -20.0, -20.0],
                                                  "top_corner_bohr_radii_3": [0.0, 20.0, 20.0]},
                                                 {"name": "right", "bottom_corner_bohr_radii_3": [0.0, -20.0, -20.0],
                                                  "top_corner_bohr_radii_3": [20.0, 20.0, 20.0]}]},
        "paths": {"cache_dir": os.path.join(tmp, "cache"), "output_dir": os.path.join(tmp, "output")},
    }


def test_free_packet_64_matches_analytic_spread(dev, tmp_path):
    cfg = free_packet_64_config(str(tmp_path))
    sim, tools, it = initialisation.initialize(use_cache=False, config=cfg)
    assert sim.get_simulation_method() == SimulationMethod.FOURIER
    core_sim.run_iteration(sim, tools, it, observe_every=10)
    el = tools.get_expected_location()
    ex = el.get_expected_location_evolution()
    sd = el.get_standard_deviation_evolution()
    assert ex.shape == (20, 3)
    t_idx = np.arange(20) * 10
    pred_x = math_utils.predict_free_space_expected_location(0.01, 0.0, 1.0, 200)[t_idx]
    pred_s = math_utils.predict_free_space_standard_devation(0.01, 3.0, 1.0, 200)[t_idx]
    assert np.max(np.abs(ex[:, 0] - pred_x)) < 1e-6
    assert np.max(np.abs(ex[:, 1:])) < 1e-9
    assert np.max(np.abs(sd - pred_s[:, None])) < 1e-6
    # the two half boxes add up to the norm
    lp, _ = tools.get_volume_probabilities()[0].get_probability_evolution_with_name()
    rp, _ = tools.get_volume_probabilities()[1].get_probability_evolution_with_name()
    assert 0.45 < lp[0] < 0.56 and 0.45 < rp[0] < 0.56 and rp[-1] > rp[0]
    tools.finish(sim)
    out = os.path.join(str(tmp_path), "output", "measurable_npy")
    saved = np.load(os.path.join(out, "expected_location_evolution.npy"))
    assert np.array_equal(saved, ex)
    raw = np.fromfile(os.path.join(str(tmp_path), "output", "measurable_raw", "standard_deviation_evolution.bin"),
                      dtype=np.float64)
    assert np.array_equal(raw.reshape(20, 3), sd)


def test_free_packet_64_state_matches_oracle(dev, tmp_path):
    cfg = free_packet_64_config(str(tmp_path))
    cfg["simulation"]["total_iteration_count"] = 30
    sim, tools, it = initialisation.initialize(use_cache=False, config=cfg)
    psi0 = sim.get_wave_function().cpu().numpy()
    core_sim.run_iteration(sim, tools, it, observe_every=1)
    ref = oracle.split_step(psi0, np.zeros(psi0.shape, np.complex64), sim.get_delta_x_bohr_radii_3(), 0.01, 1.0,
                            nsteps=30)
    assert oracle.relative_l2(sim.get_wave_function().cpu().numpy(), ref) < 1e-12
    # the sample taken before step 7 is the oracle's observable of the state after 7 steps
    b, c = tools.get_expected_location().get_box()
    ref7 = oracle.split_step(psi0, np.zeros(psi0.shape, np.complex64), sim.get_delta_x_bohr_radii_3(), 0.01, 1.0,
                             nsteps=7)
    er, _ = oracle.expected_location(ref7, sim.get_delta_x_bohr_radii_3(), b, c)
    assert np.allclose(tools.get_expected_location().get_expected_location_evolution()[7], er, atol=1e-12)


def double_slit_config(tmp, n=64, method="fourier"):
    L = 20.0
    return {
        "simulation": {"method": method, "delta_time_h_bar_per_hartree": 0.005, "total_iteration_count": 40,
                       "double_precision_calculation": False, "order_of_approximation": 10},
        "wave_packet": {"initial_position_bohr_radii_3": [-4.0, 0.0, 0.0],
                        "initial_velocity_bohr_radii_hartree_per_h_bar_3": [2.0, 0.0, 0.0],
                        "initial_standard_deviation_bohr_radii_3": [1.5, 1.5, 1.5]},
        "volume": {"simulated_volume_dimensions_bohr_radii_3": [L, L, L], "number_of_voxels_3": [n, n, n],
                   "observation_box_bottom_corner_bohr_radii_3": [-8.0, -8.0, -8.0],
                   "observation_box_top_corner_bohr_radii_3": [8.0, 8.0, 8.0]},
        "absorbing_boundary_condition": {"outer_potential_in_positive_xyz_direction_hartree_3": [-50.0] * 3,
                                         "outer_potential_in_negative_xyz_direction_hartree_3": [-50.0] * 3,
                                         "start_offset": 0.0, "slope_exponent_in_positive_xyz_direction_3": [3.0] * 3,
                                         "slope_exponent_in_negative_xyz_direction_3": [3.0] * 3, "enable": True},
        "potential": {"walls": [{"potential_hartree": 5.0, "center_bohr_radii_3": [0.0, 0.0, 0.0],
                                 "normal_vector_3": [1.0, 0.0, 0.0], "plateau_thickness_bohr_radii": 1.0,
                                 "slope_thickness_bohr_radii": 0.0, "slit_count": 2, "slit_spacing_bohr_radii": 3.0,
                                 "slit_width_bohr_radii": 0.8}]},
        "measurement": {"expected_location": {"enable_image": True}},
        "paths": {"cache_dir": os.path.join(tmp, "cache"), "output_dir": os.path.join(tmp, "output")},
    }


@pytest.mark.parametrize("method", ["fourier", "power_series"])
def test_double_slit_state_matches_oracle(dev, tmp_path, method):
    cfg = double_slit_config(str(tmp_path), 64, method)
    sim, tools, it = initialisation.initialize(use_cache=False, config=cfg)
    psi0 = sim.get_wave_function().cpu().numpy()
    V = sim.get_potential().cpu().numpy()
    dx = sim.get_delta_x_bohr_radii_3()
    # V equals the oracle's wall + CAP on the same grid
    shape = psi0.shape
    cap = oracle.absorbing_potential(shape, dx, [-8.0] * 3, [8.0] * 3, 0.0, [-50.0] * 3, [-50.0] * 3, [3.0] * 3,
                                     [3.0] * 3)
    wall = oracle.potential_wall(shape, dx, 5.0, (0, 0, 0), (1, 0, 0), 1.0, 0.0, 1.0, 2, 3.0, 0.8, 0.0)
    assert np.max(np.abs(V.imag - cap)) <= 4e-6 * np.max(np.abs(cap))
    assert np.count_nonzero(np.abs(V.real - wall) > 1e-5) <= 1e-4 * V.size
    core_sim.run_iteration(sim, tools, it, observe_every=10)
    if method == "fourier":
        ref = oracle.split_step(psi0, V, dx, 0.005, 1.0, nsteps=40)
    else:
        ref = oracle.power_series_step(psi0, V, 0.005, dx, 1.0, order=10, nsteps=40)
    assert oracle.relative_l2(sim.get_wave_function().cpu().numpy(), ref) < 1e-5


@pytest.mark.parametrize("n", [64, 128])
def test_run_iteration_stretches_match_the_reference_loop(dev, tmp_path, n):
    """run_iteration hands stretches of iterations to one wfs_step_observe call (at 128^3 the
    slab scheme fuses every sample into one launch); the reference's loop calls time_step --
    measure, evolve one step, update the potential -- once per iteration
    (sources/core_sim.py:19-24, 42-55).  Same final state bit for bit, same series."""
    a_cfg = double_slit_config(str(tmp_path / "a"), n)
    b_cfg = double_slit_config(str(tmp_path / "b"), n)
    sa, ta, ia = initialisation.initialize(use_cache=False, config=a_cfg)
    sb, tb, ib = initialisation.initialize(use_cache=False, config=b_cfg)
    if n == 128:
        assert sa.get_plan().variants()["yz_slab"] != 0
    core_sim.run_iteration(sa, ta, ia, observe_every=1)
    for i in range(ib.total_iteration_count):
        ib.i = i
        core_sim.time_step(sb, tb, ib)
    torch.cuda.synchronize()
    assert torch.equal(sa.get_wave_function(), sb.get_wave_function())
    ea = ta.get_expected_location().get_expected_location_evolution()
    eb = tb.get_expected_location().get_expected_location_evolution()
    sda = ta.get_expected_location().get_standard_deviation_evolution()
    sdb = tb.get_expected_location().get_standard_deviation_evolution()
    assert ea.shape == eb.shape == (40, 3)
    assert np.allclose(ea, eb, rtol=1e-12, atol=1e-13) and np.allclose(sda, sdb, rtol=1e-12, atol=1e-13)


def test_observe_every_matches_single_steps(dev, tmp_path):
    a_cfg = double_slit_config(str(tmp_path / "a"))
    b_cfg = double_slit_config(str(tmp_path / "b"))
    sa, ta, ia = initialisation.initialize(use_cache=False, config=a_cfg)
    sb, tb, ib = initialisation.initialize(use_cache=False, config=b_cfg)
    core_sim.run_iteration(sa, ta, ia, observe_every=1)
    core_sim.run_iteration(sb, tb, ib, observe_every=10)
    ea = ta.get_expected_location().get_expected_location_evolution()
    eb = tb.get_expected_location().get_expected_location_evolution()
    assert ea.shape == (40, 3) and eb.shape == (4, 3)
    assert np.allclose(ea[::10], eb, atol=1e-5)
    assert oracle.relative_l2(sb.get_wave_function().cpu().numpy(), sa.get_wave_function().cpu().numpy()) < 1e-5


def test_tile_variants_bit_identical(dev):
    shape = (256, 64, 128)
    dx = (0.3, 0.3, 0.3)
    rng = np.random.default_rng(5)
    a = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    V = torch.from_numpy((rng.standard_normal(shape) * 0.1).astype(np.complex64)).to(dev)
    results = []
    for v in (0, 1, 2):
        p = _abi.Plan(shape, dx, 0.01, 1.0, _abi.WFS_C64, _abi.WFS_METHOD_FOURIER)
        for k in _abi.KERNEL_CLASSES[:6]:
            p.set_variant(k, v)
        t = torch.from_numpy(a).to(dev)
        p.step(t, V, 3)
        results.append(t.cpu().numpy())
    assert np.array_equal(results[0], results[1]) and np.array_equal(results[1], results[2])
    p = _abi.Plan(shape, dx, 0.01, 1.0, _abi.WFS_C64, _abi.WFS_METHOD_FOURIER)
    chosen = p.autotune()
    # the power-series class has variants 0..5 (this Fourier plan keeps its default); others 0..2
    assert all(0 <= v <= (5 if k == "horner_stage" else 2) for k, v in chosen.items())
    t = torch.from_numpy(a).to(dev)
    p.step(t, V, 3)
    assert np.array_equal(t.cpu().numpy(), results[0])


def test_set_potential_then_step_without_v(dev):
    shape = (32, 32, 32)
    dx = (0.5, 0.5, 0.5)
    psi0 = oracle.gaussian_wave_packet(shape, dx, (2, 2, 2), (0, 0, 0), (1, 0, 0), np.complex64)
    V = torch.from_numpy(oracle.harmonic_potential(shape, dx, 0.3).astype(np.complex64)).to(dev)
    p1 = _abi.Plan(shape, dx, 0.01, 1.0, _abi.WFS_C64, _abi.WFS_METHOD_FOURIER)
    p2 = _abi.Plan(shape, dx, 0.01, 1.0, _abi.WFS_C64, _abi.WFS_METHOD_FOURIER)
    a = torch.from_numpy(psi0).to(dev)
    b = torch.from_numpy(psi0).to(dev)
    p1.step(a, V, 5)
    p2.set_potential(V)
    p2.step(b, None, 5)
    assert torch.equal(a, b)
    p3 = _abi.Plan(shape, dx, 0.01, 1.0, _abi.WFS_C64, _abi.WFS_METHOD_FOURIER)
    with pytest.raises(ValueError):
        p3.step(b, None, 1)


def test_operators_tables(dev):
    kin = operators.init_kinetic_operator((0.5, 0.5, 0.5), 0.02, 1.0, (16, 16, 16), double_precision=True)
    assert np.max(np.abs(kin.cpu().numpy() - oracle.kinetic_half_phase((16, 16, 16), (0.5,) * 3, 0.02, 1.0))) < 1e-13
    V = torch.full((8, 8, 8), 2.0 - 0.5j, dtype=torch.complex64, device=dev)
    P = operators.init_potential_operator(None, V, 0.1)
    assert np.allclose(P.cpu().numpy(), oracle.potential_phase(V.cpu().numpy(), 0.1), atol=1e-7)


def test_snapshot_roundtrip_resumes_psi(dev, tmp_path):
    cfg = double_slit_config(str(tmp_path))
    sim, tools, it = initialisation.initialize(use_cache=False, config=cfg)
    sim.evolve_state(5)
    it.i = 5
    from wavefunctionsimpython_amd import snapshot_io
    snapshot_io.write_snapshot(sim, it)
    psi5 = sim.get_wave_function().cpu().numpy()
    sim2, tools2, it2 = initialisation.initialize(use_cache=True, config=copy.deepcopy(cfg))
    assert it2.i == 5
    assert np.array_equal(sim2.get_wave_function().cpu().numpy(), psi5)


def test_expected_location_mutates_top_voxel_like_reference(dev, tmp_path):
    cfg = free_packet_64_config(str(tmp_path))
    sim = SimState(cfg)
    top_before = sim.get_observation_box_top_corner_voxel_3().copy()
    sim.initialize_state()
    MeasurementTools(cfg, sim)
    _, _, _, tv = oracle.observation_box_voxels([60.0] * 3, [64] * 3, [-25.0] * 3, [25.0] * 3)
    _, _, mutated = oracle.expected_location_box(sim.get_observation_box_bottom_corner_voxel_3(), top_before)
    assert np.array_equal(top_before, tv)
    assert np.array_equal(sim.get_observation_box_top_corner_voxel_3(), mutated)


def test_plane_current_measurement_obeys_continuity(dev, tmp_path):
    """Probability gained by the half box x > 0 equals the time integral of the flux through x = 0.

    The pixel pitch equals the voxel pitch (19.375 / 62 = 0.3125): the reference samples each
    pixel at its truncated voxel, so a pitch that differs from the voxel pitch skips or repeats
    voxels and biases the integral (reference behaviour, reproduced by the oracle)."""
    cfg = double_slit_config(str(tmp_path))
    cfg["potential"]["walls"] = []
    cfg["absorbing_boundary_condition"]["enable"] = False
    cfg["simulation"]["total_iteration_count"] = 120
    cfg["simulation"]["double_precision_calculation"] = True
    cfg["wave_packet"]["initial_position_bohr_radii_3"] = [-2.0, 0.0, 0.0]
    cfg["measurement"]["volume_probabilities"] = [{"name": "right", "bottom_corner_bohr_radii_3": [0.0, -9.9, -9.9],
                                                   "top_corner_bohr_radii_3": [9.6, 9.6, 9.6]}]
    cfg["measurement"]["plane_probability_currents"] = [{"name": "x0", "center_bohr_radii_3": [0.0, 0.0, 0.0],
                                                         "normal_vector_3": [1.0, 0.0, 0.0],
                                                         "size_bohr_radii_2": [19.375, 19.375],
                                                         "resolution_2": [62, 62]}]
    sim, tools, it = initialisation.initialize(use_cache=False, config=cfg)
    core_sim.run_iteration(sim, tools, it, observe_every=2)
    right, _ = tools.get_volume_probabilities()[0].get_probability_evolution_with_name()
    flux, _ = tools.get_plane_probability_currents()[0].get_probability_current_evolution_with_name()
    transfer, _ = tools.get_plane_probability_currents()[0].get_integrated_probability_current_evolution_with_name()
    assert flux.shape == (60,) and flux.max() > 0.1
    # Simpson-in-space boxes start on the x = 0 plane; compare increments
    gained = right - right[0]
    assert np.max(np.abs(gained[::2] - transfer[::2])) < 5e-3
    tools.finish(sim)
    npy = os.path.join(str(tmp_path), "output", "measurable_npy", "integrated_probability_current_x0.npy")
    assert np.allclose(np.load(npy), transfer)
