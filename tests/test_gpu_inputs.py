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
