"""CPU-only checks: the C ABI library loads and exports what include/wfsim.h
declares, and host-side logic (config lookup, signals, time integrals) behaves
like the reference."""

import os
import re
import signal
import subprocess
import sys

import numpy as np
import pytest
import tomli

import oracle
from wavefunctionsimpython_amd import _abi, math_utils
from wavefunctionsimpython_amd.config_read_helper import try_read_param
from wavefunctionsimpython_amd.iter_data import IterData

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    text = open(os.path.join(ROOT, "include", "wfsim.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wfs_[a-z0-9_]+)\s*\(", text)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = _abi.lib()
    names = _header_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
        assert name in _abi.SIGNATURES, f"{name} has no ctypes signature"
    assert set(_abi.SIGNATURES) == set(names)
    assert lib.wfs_version().startswith(b"wfsim-gfx950")


def test_library_is_built_for_gfx950():
    """The fat binary embeds gfx950 code objects (our own build)."""
    data = open(_abi.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data


def test_bad_arguments_fail_loudly_without_a_gpu_call():
    lib = _abi.lib()
    import ctypes
    h = ctypes.c_void_p()
    rc = lib.wfs_plan_create(_abi.i3((100, 64, 64)), _abi.d3((1, 1, 1)), 0.01, 1.0, 0, 0, 10, ctypes.byref(h))
    assert rc == 1001 and b"length 100 not supported" in lib.wfs_last_error()
    rc = lib.wfs_step(None, None, None, 1, None)
    assert rc == 1000
    with pytest.raises(ValueError):
        _abi.Plan((100, 64, 64), (1, 1, 1), 0.01, 1.0, 0, 0)


def test_try_read_param_semantics(capsys):
    cfg = {"a": {"b": {"c": 3}}, "x": 1}
    assert try_read_param(cfg, "a.b.c", 0) == 3
    assert try_read_param(cfg, "a.q.c", 7) == 7
    out = capsys.readouterr().out
    assert 'Did not found token "q" under "a"' in out and "default value of 7" in out
    assert try_read_param(cfg, "x", 5) == 1
    assert IterData({}).total_iteration_count == 1000
    assert IterData({"simulation": {"total_iteration_count": 12}}).total_iteration_count == 12


def test_delta_time_is_required():
    from wavefunctionsimpython_amd.sim_state import SimState
    with pytest.raises(KeyError):
        SimState({"simulation": {}})


def test_sim_state_geometry_cpu_only():
    from wavefunctionsimpython_amd.sim_state import SimState, SimulationMethod
    cfg = {"simulation": {"delta_time_h_bar_per_hartree": 0.01, "method": "FDTD"},
           "volume": {"simulated_volume_dimensions_bohr_radii_3": [60.0] * 3, "number_of_voxels_3": [64] * 3,
                      "observation_box_bottom_corner_bohr_radii_3": [25.0, -40.0, -25.0],
                      "observation_box_top_corner_bohr_radii_3": [-25.0, 40.0, 25.0]}}
    s = SimState(cfg)
    assert s.get_simulation_method() == SimulationMethod.POWER_SERIES
    _, _, bv, tv = oracle.observation_box_voxels([60.0] * 3, [64] * 3, [25.0, -40.0, -25.0], [-25.0, 40.0, 25.0])
    assert np.array_equal(s.get_observation_box_bottom_corner_voxel_3(), bv)
    assert np.array_equal(s.get_observation_box_top_corner_voxel_3(), tv)
    assert np.allclose(s.get_delta_x_bohr_radii_3(), 60.0 / 64)
    s2 = SimState({"simulation": {"delta_time_h_bar_per_hartree": 0.01, "method": "something"}})
    assert s2.get_simulation_method() == SimulationMethod.FOURIER


@pytest.mark.parametrize("n", [0, 1, 2, 3, 4, 5, 10, 11, 64, 101])
def test_indefinite_simpson_integral_matches_oracle(n):
    a = np.random.default_rng(n).standard_normal(n)
    assert np.allclose(math_utils.indefinite_simpson_integral(a, 0.37), oracle.indefinite_simpson_integral(a, 0.37),
                       rtol=1e-13, atol=1e-13)


def test_analytic_predictions_match_oracle():
    assert np.allclose(math_utils.predict_free_space_standard_devation(0.01, 3.0, 1.0, 50),
                       oracle.predict_free_space_standard_deviation(0.01, 3.0, 1.0, 50))
    assert np.allclose(math_utils.predict_free_space_expected_location(0.01, 1.0, 2.0, 50),
                       oracle.predict_free_space_expected_location(0.01, 1.0, 2.0, 50))


def test_sigint_sets_quit_flag():
    from wavefunctionsimpython_amd import core_sim
    it = IterData({})
    prev = core_sim.register_signal_handler(it)
    try:
        os.kill(os.getpid(), signal.SIGINT)
        assert it.is_quit
    finally:
        signal.signal(signal.SIGINT, prev)


def test_toml_decode_error_exits_with_status_1(tmp_path):
    bad = tmp_path / "bad.toml"
    bad.write_text("[simulation\nx = ")
    code = ("import sys; sys.path.insert(0, %r); from wavefunctionsimpython_amd import initialisation; "
            "initialisation.load_config(%r)" % (ROOT, str(bad)))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True)
    assert r.returncode == 1 and "Failed to decode TOML" in r.stdout


def test_reference_toml_schema_parses(tmp_path):
    from wavefunctionsimpython_amd import initialisation
    p = tmp_path / "c.toml"
    p.write_text("""
[simulation]
method = "fourier"
delta_time_h_bar_per_hartree = 0.2
total_iteration_count = 541
double_precision_calculation = true
[volume]
number_of_voxels_3 = [512, 256, 256]
[[potential.walls]]
potential_hartree = 7.0
slit_count = 2
[[measurement.volume_probabilities]]
name = "First half"
""")
    cfg = initialisation.load_config(str(p))
    assert try_read_param(cfg, "simulation.total_iteration_count", 0) == 541
    assert cfg["potential"]["walls"][0]["slit_count"] == 2
    # the cache copy written by initialize reads back as the same dict
    assert tomli.loads(initialisation.toml_dumps(cfg)) == cfg


def test_config_file_pick(tmp_path, monkeypatch):
    """initialize(use_cache) without a config reads the one TOML file under config/, or asks for
    an index when there are several (sources/initialisation.py:16-40)."""
    from wavefunctionsimpython_amd import initialisation
    d = tmp_path / "config"
    d.mkdir()
    (d / "only.toml").write_text("[simulation]\ndelta_time_h_bar_per_hartree = 0.1\n")
    assert initialisation.select_config_file(str(d)).endswith("only.toml")
    (d / "another.toml").write_text("[simulation]\ndelta_time_h_bar_per_hartree = 0.2\n")
    answers = iter(["x", "7", "1"])
    monkeypatch.setattr("builtins.input", lambda: next(answers))
    assert initialisation.select_config_file(str(d)).endswith("only.toml")     # sorted: another, only
    with pytest.raises(SystemExit):
        initialisation.select_config_file(str(tmp_path / "missing"))


def test_bench_distinct_plane_share_and_grouped_observable_error():
    """bench.py's byte model reads the share of x planes that start a run of bitwise-equal
    planes; its observable error scales each moment by its group (P, first moments, second
    moments), so a moment that vanishes by symmetry does not blow the error up."""
    import torch
    sys.path.insert(0, ROOT)
    import bench
    V = torch.zeros((8, 4, 4), dtype=torch.complex64)
    V[3] += 1.0
    V[4] += 1.0                      # same bits as plane 3: no new run
    V[6, 0, 0] = 1j                  # one voxel differs from plane 5
    V[7] = V[6]
    V[7, 3, 3] = complex(-0.0, 0.0)  # -0.0 differs from 0.0 in its bits
    assert torch.signbit(V[7, 3, 3].real)
    assert bench.distinct_plane_fraction(V) == 5 / 8      # runs start at planes 0, 3, 5, 6 and 7
    ref = torch.zeros((5, 3, 7), dtype=torch.float64)
    ref[..., 0], ref[..., 1], ref[..., 2] = 1.0, 16.0, 1e-17
    ref[..., 4], ref[..., 5], ref[..., 6] = 300.0, 20.0, 20.0
    obs = ref.clone()
    obs[..., 2] += 1e-9                                   # noise on a vanishing first moment
    assert bench.observables_rel_err(obs, ref) == pytest.approx(1e-9 / 16.0)
    obs[..., 4] += 3e-4
    assert bench.observables_rel_err(obs, ref) == pytest.approx(1e-6)


def test_duplicate_measurement_names_take_the_first_free_suffix(capsys):
    """The rename rule for two measurement tools of one kind with equal names: the later one
    gets the first suffix _1, _2, ... that no item has (sources/measurement.py:502-518)."""
    from wavefunctionsimpython_amd.measurement import MeasurementTools

    class Named:
        def __init__(self, name):
            self.name = name

        def get_name(self):
            return self.name

        def set_name(self, name):
            self.name = name

    items = [Named("slit"), Named("slit_1"), Named("slit_3"), Named("screen")]
    dup = Named("slit")
    MeasurementTools._resolve_naming_conflicts(items, dup)
    assert dup.get_name() == "slit_2"
    out = capsys.readouterr().out
    assert 'equal name: "slit"' in out and 'Renamed one to "slit_2".' in out
    fresh = Named("detector")
    MeasurementTools._resolve_naming_conflicts(items, fresh)
    assert fresh.get_name() == "detector" and capsys.readouterr().out == ""
    single = Named("screen")
    MeasurementTools._resolve_naming_conflicts(items, single)
    assert single.get_name() == "screen_1"


def test_run_iteration_restores_a_sigint_handler_set_outside_python(monkeypatch):
    """signal.getsignal() reads None for a handler not installed from Python (a profiler's, say);
    run_iteration then restores the default handler instead of failing on the way out."""
    import signal
    from wavefunctionsimpython_amd import core_sim
    calls = []
    real = signal.signal

    def fake_signal(sig, handler):
        calls.append(handler)
        return None if len(calls) == 1 else real(sig, handler)

    class _Sim:
        def potential_is_static(self):
            return True

        def get_delta_time_h_bar_per_hartree(self):
            return 0.01

    class _Tools:
        def set_observe_every(self, n):
            pass

    it = IterData({})
    it.i = it.total_iteration_count = 0
    monkeypatch.setattr(signal, "signal", fake_signal)
    try:
        core_sim.run_iteration(_Sim(), _Tools(), it, observe_every=1, exit_on_quit=False)
    finally:
        real(signal.SIGINT, signal.default_int_handler)
    assert calls[-1] is signal.SIG_DFL
