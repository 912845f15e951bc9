"""Two-launch split step for complex64 grids with ny = nz = 256, 192 or 128 (the slab scheme).

The length-n y transform is split in a length-n/B stage on the rows y = B a + b (fused with
the z inverse / potential / z forward work) and a length-B stage fused with the x pass
(B = 4; B = 2 at 128).  Its
arithmetic differs from the four axis passes only in how the y transform is factored, so the
two paths agree to float rounding; both are pinned to the FP64 oracle at the complex64 bar.
"""

import numpy as np
import pytest
import torch

import oracle
from wavefunctionsimpython_amd import _abi

pytestmark = pytest.mark.gpu

C64, C128 = _abi.WFS_C64, _abi.WFS_C128
DX = (0.3, 0.25, 0.35)
DT = 0.01
TOL_ORACLE = 1e-5     # complex64 state vs the FP64 oracle, relative L2
TOL_PASSES = 2e-6     # slab scheme vs the four axis passes, relative L2


def _inputs(shape, seed):
    """a smooth packet with a random phase ripple and a complex potential (CAP-like imaginary part)"""
    rng = np.random.default_rng(seed)
    psi = oracle.gaussian_wave_packet(shape, DX, (1.2, 6.0, 8.0), (0.0, 1.0, -2.0), (1.0, 0.5, -0.8),
                                      np.complex128)
    psi = psi * np.exp(1j * 0.3 * rng.standard_normal(shape))
    V = (rng.standard_normal(shape) * 0.5 + 1j * rng.uniform(-0.05, 0.0, shape)).astype(np.complex64)
    return psi, V


# (yz_slab, x_slab) variants: one block per row group / persistent; 16 / 8 z points per x block
FORMS = [(1, 1), (2, 2)]


def _run(shape, slab, psi0, V=None, axes=None, batch=1, steps=5, dev="cuda:0"):
    """slab: False (four axis passes) or a (yz_slab, x_slab) variant pair"""
    p = _abi.Plan(shape, DX, DT, 1.0, C64, _abi.WFS_METHOD_FOURIER, batch=batch)
    yz, xv = slab if slab else (0, 1)
    p.set_variant("yz_slab", yz)
    p.set_variant("x_slab", xv)
    psi = torch.as_tensor(psi0).to(dev).to(torch.complex64)
    if axes is not None:
        p.set_potential_axes(*axes)
        p.step(psi, None, steps)
    else:
        p.step(psi, torch.as_tensor(V).to(dev), steps)
    torch.cuda.synchronize()
    return psi


def _rel(a, b):
    a = np.asarray(a, dtype=np.complex128)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("nx,n", [(8, 256), (16, 256), (32, 256), (8, 192), (16, 192), (96, 192), (8, 128),
                                  (128, 128)])
def test_slab_matches_oracle_table_potential(dev, nx, n, form):
    shape = (nx, n, n)
    psi0, V = _inputs(shape, 100 + nx)
    print(f"case: slab {form}, table potential, shape {shape}, 5 steps", flush=True)
    got = _run(shape, form, psi0, V).cpu().numpy()
    ref = oracle.split_step(psi0, V, DX, DT, 1.0, nsteps=5)
    err = _rel(got, ref)
    ref_passes = _run(shape, False, psi0, V).cpu().numpy()
    err_passes = _rel(got, ref_passes)
    print(f"  rel L2 vs oracle {err:.3e}, vs four passes {err_passes:.3e}, "
          f"four passes vs oracle {_rel(ref_passes, ref):.3e}")
    assert err < TOL_ORACLE
    assert err_passes < TOL_PASSES


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("n", [128, 192, 256])
def test_slab_matches_oracle_separable_potential(dev, n, form):
    shape = (16, n, n)
    psi0, _ = _inputs(shape, 7)
    rng = np.random.default_rng(8)
    axes = [rng.uniform(-1.0, 1.0, (1, m)) for m in shape]
    V = (axes[0][0][:, None, None] + axes[1][0][None, :, None] + axes[2][0][None, None, :])
    print(f"case: slab {form}, separable potential, shape {shape}, 6 steps", flush=True)
    got = _run(shape, form, psi0, axes=axes, steps=6).cpu().numpy()
    ref = oracle.split_step(psi0, V, DX, DT, 1.0, nsteps=6)
    err = _rel(got, ref)
    print(f"  rel L2 vs oracle {err:.3e}")
    assert err < TOL_ORACLE


@pytest.mark.parametrize("form", FORMS)
def test_slab_single_step_and_entry_exit(dev, form):
    """one step: entry, one y-z launch with its x launch, exit"""
    shape = (8, 256, 256)
    psi0, V = _inputs(shape, 3)
    got = _run(shape, form, psi0, V, steps=1).cpu().numpy()
    ref = oracle.split_step(psi0, V, DX, DT, 1.0, nsteps=1)
    err = _rel(got, ref)
    print(f"case: slab, 1 step, shape {shape}: rel L2 vs oracle {err:.3e}")
    assert err < TOL_ORACLE


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("n", [128, 192, 256])
def test_slab_batched_bit_identical_to_single_runs(dev, n, form):
    batch = 3
    shape = (8, n, n)
    rng = np.random.default_rng(21)
    axes = [rng.uniform(-1.0, 1.0, (batch, m)) for m in shape]
    psis = [_inputs(shape, 30 + i)[0] for i in range(batch)]
    stacked = np.stack(psis)
    print(f"case: slab {form}, separable batched, batch {batch}, shape {shape}", flush=True)
    both = _run(shape, form, stacked, axes=axes, batch=batch)
    for i in range(batch):
        one = _run(shape, form, psis[i], axes=[a[i:i + 1] for a in axes])
        assert torch.equal(both[i], one), f"packet {i}"


@pytest.mark.parametrize("n", [128, 192, 256])
def test_slab_full_grid_against_passes_and_norm(dev, n):
    """n^3 (256^3 is the benchmark grid): no absorber, so the norm is kept; slab vs four passes"""
    shape = (n, n, n)
    psi0 = oracle.gaussian_wave_packet(shape, (80 / n,) * 3, (4.0, 4.0, 4.0), (-16.0, 0.0, 0.0),
                                       (2.0, 0.0, 0.0), np.complex64)
    rng = np.random.default_rng(5)
    V = (rng.uniform(0.0, 1.0, shape)).astype(np.complex64)
    p_dx = (80 / n,) * 3
    out = {}
    for form in [(0, 1)] + FORMS:
        p = _abi.Plan(shape, p_dx, 0.005, 1.0, C64, _abi.WFS_METHOD_FOURIER)
        p.set_variant("yz_slab", form[0])
        p.set_variant("x_slab", form[1])
        psi = torch.from_numpy(psi0).to(dev)
        p.step(psi, torch.from_numpy(V).to(dev), 20)
        torch.cuda.synchronize()
        out[form] = psi
    a = out[(0, 1)]
    n0 = float(np.sum(np.abs(psi0.astype(np.complex128)) ** 2))
    for form in FORMS:
        b = out[form]
        rel = float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(a))
        n1 = float(torch.sum(torch.abs(b.to(torch.complex128)) ** 2))
        print(f"case: {n}^3 x 20 steps, slab {form}: vs passes rel L2 {rel:.3e}, norm drift {abs(n1 / n0 - 1):.3e}")
        assert rel < TOL_PASSES
        assert abs(n1 / n0 - 1) < 1e-5
    # the forms differ only in how blocks are scheduled: identical results
    assert torch.equal(out[FORMS[0]], out[FORMS[1]])


def test_slab_defaults_and_limits(dev):
    def variant(shape, precision=C64, batch=1):
        return _abi.Plan(shape, DX, DT, 1.0, precision, _abi.WFS_METHOD_FOURIER, batch=batch).variants()["yz_slab"]
    assert variant((256, 256, 256)) == 2
    assert variant((16, 256, 256), batch=4) == 2
    assert variant((192, 192, 192)) == 1
    assert variant((64, 192, 192), batch=3) == 1
    assert variant((128, 128, 128)) == 2                 # 128 x planes: the plane pass leaves CUs idle
    assert variant((64, 128, 128), batch=8) == 0         # 512 planes: the plane pass
    assert variant((128, 128, 128), C128) == 0
    assert variant((256, 256, 256), C128) == 0
    assert variant((512, 256, 256)) == 0
    assert variant((256, 128, 128)) == 0                 # 256 x 128 planes of 128^2
    assert variant((256, 256, 128)) == 0
    p = _abi.Plan((64, 64, 64), DX, DT, 1.0, C64, _abi.WFS_METHOD_FOURIER)
    with pytest.raises(_abi.WfsError):
        p.set_variant("yz_slab", 1)
    _abi.Plan((64, 128, 128), DX, DT, 1.0, C64, _abi.WFS_METHOD_FOURIER).set_variant("yz_slab", 1)
    q = _abi.Plan((16, 256, 256), DX, DT, 1.0, C64, _abi.WFS_METHOD_FOURIER)
    with pytest.raises(_abi.WfsError):
        q.set_variant("yz_slab", 3)
    q.set_variant("yz_slab", 0)
    assert q.variants()["yz_slab"] == 0


def test_slab_autotune_and_profile_classes(dev):
    shape = (32, 256, 256)
    psi0, V = _inputs(shape, 44)
    p = _abi.Plan(shape, DX, DT, 1.0, C64, _abi.WFS_METHOD_FOURIER)
    chosen = p.autotune()
    print(f"autotuned yz_slab at {shape}: {chosen['yz_slab']}")
    assert chosen["yz_slab"] in (0, 1, 2) and chosen["x_slab"] in (1, 2)
    p.set_variant("yz_slab", 1)
    p.set_profiling(True)
    psi = torch.from_numpy(psi0).to(dev).to(torch.complex64)
    p.step(psi, torch.from_numpy(V).to(dev), 4)
    prof = p.profile(reset=True)
    p.set_profiling(False)
    assert prof["yz_slab"][0] == 4 and prof["x_slab"][0] == 5 and prof["slab_edge"][0] == 2
    for k in ("y_inv", "y_fwd", "z_inv_v_fwd", "x_fwd_k_inv"):
        assert prof.get(k, (0, 0.0))[0] == 0


def test_profile_times_lie_within_the_calls(dev):
    """Per-kernel timing comes from start and end events recorded by the launches themselves:
    every class's summed kernel time is positive, the classes of a call together take no more
    than the call's own wall time on the stream (events around the whole call), and the two
    launches of a step account for most of a long call.  The observables class (moments and
    their sum, two launches) counts one timed pair per call."""
    shape = (128, 256, 256)
    psi0, V = _inputs(shape, 46)
    p = _abi.Plan(shape, DX, DT, 1.0, C64, _abi.WFS_METHOD_FOURIER)
    psi = torch.from_numpy(psi0).to(dev).to(torch.complex64)
    p.set_potential(torch.from_numpy(V).to(dev))
    p.step(psi, None, 2)
    boxes = [((0, 0, 0), (127, 255, 255)), ((4, 10, 20), (9, 31, 63))]
    out = p.observables(psi, boxes)
    torch.cuda.synchronize()
    p.set_profiling(True)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    p.step(psi, None, 20)
    p.observables(psi, boxes, out=out)
    e1.record()
    torch.cuda.synchronize()
    prof = p.profile(reset=True)
    p.set_profiling(False)
    wall = e0.elapsed_time(e1)
    print(f"case: profile vs wall, {shape}: {prof}, wall {wall:.3f} ms", flush=True)
    assert prof["yz_slab"][0] == 20 and prof["x_slab"][0] == 21 and prof["slab_edge"][0] == 2
    assert prof["observables"][0] == 1
    assert all(ms > 0.0 for _, ms in prof.values())
    total = sum(ms for _, ms in prof.values())
    assert total <= wall * 1.001
    assert prof["yz_slab"][1] + prof["x_slab"][1] >= 0.6 * wall


@pytest.mark.parametrize("form", FORMS)
def test_slab_graph_replays_match_eager(dev, form):
    shape = (16, 256, 256)
    psi0, V = _inputs(shape, 55)
    Vd = torch.from_numpy(V).to(dev)
    eager_plan = _abi.Plan(shape, DX, DT, 1.0, C64, _abi.WFS_METHOD_FOURIER)
    graph_plan = _abi.Plan(shape, DX, DT, 1.0, C64, _abi.WFS_METHOD_FOURIER)
    assert eager_plan.variants()["yz_slab"] == 2
    for pl in (eager_plan, graph_plan):
        pl.set_variant("yz_slab", form[0])
        pl.set_variant("x_slab", form[1])
    eager_plan.set_potential(Vd)
    graph_plan.set_potential(Vd)
    eager = torch.from_numpy(psi0).to(dev).to(torch.complex64)
    graphed = eager.clone()
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    with _abi.Graph(s) as g:
        graph_plan.step(graphed, None, 3, stream=s)
    for replay in range(3):
        eager_plan.set_step_index(0)      # a replay uses the table copies of its capture
        eager_plan.step(eager, None, 3)
        torch.cuda.synchronize()
        g.launch()
        s.synchronize()
        junk = torch.ones(1000, device=dev).sum().item()
        assert junk == 1000.0
        assert torch.equal(eager, graphed), f"replay {replay}"


def test_potential_table_order_follows_the_scheme(dev):
    """The plan stores the phase table in the order its scheme reads (C order for the axis
    passes, lines order for the slab launch) and rebuilds the other order on a switch: a run
    that switches after set_potential equals, bit for bit, a run that never switched."""
    shape = (8, 256, 256)
    psi0, V = _inputs(shape, 66)
    Vd = torch.from_numpy(V).to(dev)
    for first, then in ((0, 1), (1, 0), (2, 0), (0, 2)):
        switched = _abi.Plan(shape, DX, DT, 1.0, C64, _abi.WFS_METHOD_FOURIER)
        switched.set_variant("yz_slab", first)
        switched.set_potential(Vd)
        switched.set_variant("yz_slab", then)
        direct = _abi.Plan(shape, DX, DT, 1.0, C64, _abi.WFS_METHOD_FOURIER)
        direct.set_variant("yz_slab", then)
        direct.set_potential(Vd)
        a = torch.from_numpy(psi0).to(dev).to(torch.complex64)
        b = a.clone()
        switched.step(a, None, 3)
        direct.step(b, None, 3)
        torch.cuda.synchronize()
        print(f"case: table set under yz_slab {first}, stepped under {then}", flush=True)
        assert torch.equal(a, b)


def _repeating_planes(shape, batch, seed):
    """Potential whose x planes repeat in runs, as walls and absorbing layers leave them: per packet
    three distinct leading planes, a run of one plane, two distinct wall planes, then the same run
    broken by planes that differ from their neighbours in one voxel (the last, the first and a
    middle one of the plane)."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    out = []
    for _ in range(batch):
        def plane():
            return (rng.standard_normal((ny, nz)) * 0.5 + 1j * rng.uniform(-0.05, 0.0, (ny, nz))).astype(np.complex64)
        base = plane()
        V = np.empty(shape, np.complex64)
        V[:] = base
        V[0], V[1], V[2] = plane(), plane(), plane()
        V[9], V[10] = plane(), plane()
        V[13, -1, -1] += np.float32(0.25)
        V[14:, :, :] = V[13]             # the changed plane starts a new run
        V[16, 0, 0] -= np.float32(0.5)
        V[17, ny // 2, nz // 3] += np.complex64(0.01j)
        out.append(V)
    return np.stack(out) if batch > 1 else out[0]


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("batch", [1, 2])
def test_shared_potential_planes_bit_identical(dev, form, batch):
    """Reading one copy of each run of bitwise-equal x planes (kernel class pv_planes, variant 1)
    gives the state, bit for bit, of reading every plane's own copy (variant 0); single-voxel
    differences at a plane's first, last and a middle voxel end a run."""
    shape = (32, 256, 256)
    V = _repeating_planes(shape, batch, 77)
    psi1, _ = _inputs(shape, 78)
    psi0 = np.stack([psi1, np.roll(psi1, 5, axis=0)]) if batch > 1 else psi1
    Vd = torch.from_numpy(V).to(dev)
    out = {}
    for share in (1, 0):
        p = _abi.Plan(shape, DX, DT, 1.0, C64, _abi.WFS_METHOD_FOURIER, batch=batch)
        assert p.variants()["pv_planes"] == 1
        p.set_variant("yz_slab", form[0])
        p.set_variant("x_slab", form[1])
        p.set_variant("pv_planes", share)
        p.set_potential(Vd)
        psi = torch.from_numpy(psi0).to(dev).to(torch.complex64)
        p.step(psi, None, 4)
        obs = p.step_observe(psi, None, 6, 3, [((0, 0, 0), (shape[0], 128, 128))])
        torch.cuda.synchronize()
        out[share] = (psi, obs)
    print(f"case: slab {form}, batch {batch}, {shape}, shared planes against own copies", flush=True)
    assert torch.equal(out[1][0], out[0][0])
    assert torch.equal(out[1][1], out[0][1])
    if batch == 1:
        ref = oracle.split_step(psi0, V, DX, DT, 1.0, nsteps=10)
        assert _rel(out[1][0].cpu().numpy(), ref) < TOL_ORACLE
    with pytest.raises(_abi.WfsError):
        p.set_variant("pv_planes", 2)


def test_shared_planes_map_is_not_built_under_capture(dev):
    """Plane sharing switched on after the table was built, then a capture: the captured steps
    read every plane's own copy (the map is not built inside a capture, where its kernels would
    run only at replay), the first uncaptured call after the capture builds it, and every
    result equals the eager runs of a plan that shared from the start, bit for bit."""
    shape = (32, 256, 256)
    V = torch.from_numpy(_repeating_planes(shape, 1, 79)).to(dev)
    psi0, _ = _inputs(shape, 80)
    start = torch.from_numpy(psi0).to(dev).to(torch.complex64)
    p = _abi.Plan(shape, DX, DT, 1.0, C64, _abi.WFS_METHOD_FOURIER)
    p.set_variant("pv_planes", 0)
    p.set_potential(V)
    p.set_variant("pv_planes", 1)
    ref = _abi.Plan(shape, DX, DT, 1.0, C64, _abi.WFS_METHOD_FOURIER)
    ref.set_potential(V)
    graphed = start.clone()
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    with _abi.Graph(s) as g:
        p.step(graphed, None, 3, stream=s)           # table copies 0..2
    eager = start.clone()
    p.step(eager, None, 3)                           # uncaptured: builds the map; copies 3..5
    want = start.clone()
    ref.set_step_index(3)
    ref.step(want, None, 3)
    torch.cuda.synchronize()
    print("case: shared planes switched on, capture, then an uncaptured call", flush=True)
    assert torch.equal(eager, want)
    want = start.clone()
    for replay in range(2):
        g.launch()
        s.synchronize()
        ref.set_step_index(0)                        # a replay uses the copies of its capture
        ref.step(want, None, 3)
        torch.cuda.synchronize()
        assert torch.equal(graphed, want), f"replay {replay}"


def test_table_reorder_inside_a_capture(dev):
    """The lines-ordered table holds an older potential, the new one is set under the four
    passes, and the plan goes back to the slab scheme inside a capture: the reorder is captured
    (it runs at every replay) and not taken as done, so the first uncaptured call reorders
    again instead of stepping with the older potential; replays and that call equal a plan
    that used the slab scheme from the start, bit for bit."""
    shape = (8, 256, 256)
    psi0, V = _inputs(shape, 81)
    _, V_old = _inputs(shape, 82)
    Vd = torch.from_numpy(V).to(dev)
    start = torch.from_numpy(psi0).to(dev).to(torch.complex64)
    p = _abi.Plan(shape, DX, DT, 1.0, C64, _abi.WFS_METHOD_FOURIER)
    p.set_variant("yz_slab", 2)
    p.set_potential(torch.from_numpy(V_old).to(dev))       # the lines table exists, older V
    p.set_variant("yz_slab", 0)
    p.set_potential(Vd)
    p.set_variant("yz_slab", 2)
    direct = _abi.Plan(shape, DX, DT, 1.0, C64, _abi.WFS_METHOD_FOURIER)
    direct.set_variant("yz_slab", 2)
    direct.set_potential(Vd)
    graphed = start.clone()
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    with _abi.Graph(s) as g:
        p.step(graphed, None, 3, stream=s)
    eager = start.clone()
    p.step(eager, None, 3)
    want = start.clone()
    direct.set_step_index(3)
    direct.step(want, None, 3)
    torch.cuda.synchronize()
    print("case: table reorder captured, then an uncaptured call and two replays", flush=True)
    assert torch.equal(eager, want)
    want = start.clone()
    for replay in range(2):
        g.launch()
        s.synchronize()
        direct.set_step_index(0)
        direct.step(want, None, 3)
        torch.cuda.synchronize()
        assert torch.equal(graphed, want), f"replay {replay}"


def test_potential_set_inside_a_capture_retires_the_plane_map(dev):
    """A potential whose planes repeat is set uncaptured (the plan builds its plane map), then a
    potential with no repeats is set inside a capture.  An uncaptured call before the first
    replay still finds the older table; after the replay the table holds the new potential,
    which the plan cannot see, so it must not describe the table by a map built before: every
    result equals plans that hold each potential from the start, bit for bit."""
    shape = (32, 256, 256)
    V0 = torch.from_numpy(_repeating_planes(shape, 1, 83)).to(dev)
    psi0, V1n = _inputs(shape, 84)
    V1 = torch.from_numpy(V1n).to(dev)
    start = torch.from_numpy(psi0).to(dev).to(torch.complex64)
    p = _abi.Plan(shape, DX, DT, 1.0, C64, _abi.WFS_METHOD_FOURIER)
    p.set_potential(V0)
    assert p.variants()["pv_planes"] == 1
    graphed = start.clone()
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    with _abi.Graph(s) as g:
        p.step(graphed, V1, 3, stream=s)             # the new table is written at each replay
    before = start.clone()
    p.step(before, None, 3)                          # no replay yet: the older table
    torch.cuda.synchronize()                         # the replay rewrites the table this call reads
    g.launch()
    s.synchronize()
    after = start.clone()
    p.step(after, None, 3)                           # the replay's table
    torch.cuda.synchronize()
    r0 = _abi.Plan(shape, DX, DT, 1.0, C64, _abi.WFS_METHOD_FOURIER)
    r0.set_potential(V0)
    r1 = _abi.Plan(shape, DX, DT, 1.0, C64, _abi.WFS_METHOD_FOURIER)
    r1.set_potential(V1)
    want = start.clone()
    r0.set_step_index(3)
    r0.step(want, None, 3)
    torch.cuda.synchronize()
    print("case: potential set inside a capture, uncaptured calls before and after a replay", flush=True)
    assert torch.equal(before, want)
    want = start.clone()
    r1.set_step_index(6)
    r1.step(want, None, 3)
    torch.cuda.synchronize()
    assert torch.equal(after, want)
    want = start.clone()
    r1.set_step_index(0)
    r1.step(want, None, 3)
    torch.cuda.synchronize()
    assert torch.equal(graphed, want)
