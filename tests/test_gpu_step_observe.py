"""wfs_step_observe: steps with the box observables taken every observe_every steps.

The reference measures before each step (sources/core_sim.py:19-65,
sources/measurement.py:644-704); wfs_step_observe(n, every) must give what alternating
wfs_observables and wfs_step(every) calls give.  With the slab scheme the exit of one stretch,
the observation and the entry of the next run as one launch: the state must still be bit for
bit that of the alternating calls and the observables equal to FP64 rounding (their partial
sums are added in another order).
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


def _inputs(shape, seed, dtype=np.complex64):
    rng = np.random.default_rng(seed)
    psi = oracle.gaussian_wave_packet(shape[-3:], DX, (1.2, 6.0, 8.0), (0.0, 1.0, -2.0), (1.0, 0.5, -0.8),
                                      np.complex128)
    psi = np.broadcast_to(psi, shape) * np.exp(1j * 0.3 * rng.standard_normal(shape))
    V = (rng.standard_normal(shape) * 0.5 + 1j * rng.uniform(-0.05, 0.0, shape)).astype(np.complex64)
    return psi.astype(dtype), V


def _boxes(n):
    nx, ny, nz = n
    return [((1, 3, 5), (nx - 3 - (nx - 3 + 1) % 2, ny - 9, nz - 11)),        # most of the grid
            ((0, 0, 0), (nx // 2 + 1, ny // 2 + 1, nz // 4 + 1)),           # a corner
            ((nx // 2, ny // 4, nz // 3), (nx - nx // 2, min(9, ny // 2), min(33, nz - nz // 3)))]   # reaches the x end


def _pair(shape, precision, method, batch, slab, seed, V_in_call=True):
    n = shape[-3:]
    psi0, V = _inputs(shape, seed, np.complex64 if precision == C64 else np.complex128)
    plans = []
    for _ in range(2):
        p = _abi.Plan(n, DX, DT, 1.0, precision, method, batch=batch)
        if slab is not None:
            p.set_variant("yz_slab", slab)
        plans.append(p)
    dev = torch.device("cuda:0")
    return plans, torch.from_numpy(np.ascontiguousarray(psi0)).to(dev), torch.from_numpy(V).to(dev)


def _check(plans, psi, V, nsteps, every, boxes, bit_state=True, tol=1e-13):
    fused_plan, ref_plan = plans
    a = psi.clone()
    b = psi.clone()
    got = fused_plan.step_observe(a, V, nsteps, every, boxes)
    refs = []
    k = 0
    while k < nsteps:
        refs.append(ref_plan.observables(b, boxes).clone())
        m = min(every, nsteps - k)
        ref_plan.step(b, V if k == 0 else None, m)
        k += m
    torch.cuda.synchronize()
    ref = torch.stack(refs)
    assert got.shape == ref.shape
    rel = float((got - ref).abs().max() / ref.abs().max())
    print(f"  observables max rel diff {rel:.2e}; state max |diff| {float((a - b).abs().max()):.2e}", flush=True)
    if bit_state:
        assert torch.equal(a, b)
    else:
        assert float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b)) < 1e-6
    assert rel < tol


@pytest.mark.parametrize("slab", [1, 2])
def test_slab_fused_observations_match_alternating_calls(dev, slab):
    shape = (16, 256, 256)
    print(f"case: slab variant {slab}, shape {shape}, 25 steps every 10, 3 boxes", flush=True)
    plans, psi, V = _pair(shape, C64, _abi.WFS_METHOD_FOURIER, 1, slab, 3)
    _check(plans, psi, V, 25, 10, _boxes(shape))


def test_slab_fused_observations_batched_and_192(dev):
    shape = (2, 8, 192, 192)
    print(f"case: slab, batched {shape}, 12 steps every 4", flush=True)
    plans, psi, V = _pair(shape, C64, _abi.WFS_METHOD_FOURIER, 2, None, 4)
    assert plans[0].variants()["yz_slab"] != 0
    _check(plans, psi, V, 12, 4, _boxes(shape[1:]))


@pytest.mark.parametrize("shape", [(16, 128, 128), (128, 128, 128)])
def test_slab_fused_observations_at_128(dev, shape):
    """the 128^2 form of the observing launch (64-row groups of rows y = 2a + b, 512 threads):
    a single 128^3 grid runs the slab scheme, so step_observe takes this path there"""
    print(f"case: slab at 128^2, shape {shape}, 12 steps every 5, 3 boxes", flush=True)
    plans, psi, V = _pair(shape, C64, _abi.WFS_METHOD_FOURIER, 1, None, 9)
    assert plans[0].variants()["yz_slab"] != 0
    _check(plans, psi, V, 12, 5, _boxes(shape))


def test_slab_fused_observations_four_and_five_boxes(dev):
    shape = (8, 256, 256)
    boxes = _boxes(shape) + [((2, 2, 2), (3, 5, 7)), ((4, 10, 20), (1, 1, 1))]
    print(f"case: slab, {shape}, 4 boxes then 5 (two passes)", flush=True)
    plans, psi, V = _pair(shape, C64, _abi.WFS_METHOD_FOURIER, 1, None, 5)
    _check(plans, psi, V, 6, 3, boxes[:4])
    plans, psi, V = _pair(shape, C64, _abi.WFS_METHOD_FOURIER, 1, None, 5)
    _check(plans, psi, V, 6, 3, boxes)


@pytest.mark.parametrize("precision,n,method", [(C64, 64, _abi.WFS_METHOD_FOURIER), (C128, 32, _abi.WFS_METHOD_FOURIER),
                                                (C64, 32, _abi.WFS_METHOD_POWER_SERIES)])
def test_other_paths_match_alternating_calls(dev, precision, n, method):
    shape = (n, n, n)
    print(f"case: precision {precision}, method {method}, {shape}", flush=True)
    plans, psi, V = _pair(shape, precision, method, 1, None, 6)
    _check(plans, psi, V, 7, 3, _boxes(shape))


def test_edge_cases(dev):
    shape = (8, 256, 256)
    plans, psi, V = _pair(shape, C64, _abi.WFS_METHOD_FOURIER, 1, None, 7)
    # one observation (every > nsteps), and no boxes at all
    _check(plans, psi, V, 3, 10, _boxes(shape))
    a = psi.clone()
    b = psi.clone()
    out = plans[0].step_observe(a, None, 4, 2, [])
    assert out.numel() == 0
    plans[1].step(b, None, 2)      # two stretches of 2 steps, as the call without boxes runs them
    plans[1].step(b, None, 2)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    with pytest.raises(ValueError):
        plans[0].step_observe(a, None, 4, 0, _boxes(shape))


def test_slab_fused_observations_boxes_sharing_z_ranges(dev):
    """boxes with one z range (the observation box and its x halves, as the benchmark measures),
    and a run of them broken by a box with another z range"""
    shape = (16, 256, 256)
    b, c = 3, 201
    boxes = [((b, b, b), (13, c, c)), ((b, b, b), (7, c, c)), ((b + 6, b, b), (7, c, c)),
             ((0, 5, 7), (9, 9, 9)), ((2, 5, b), (5, 9, c))]
    print(f"case: slab, {shape}, boxes sharing z ranges", flush=True)
    plans, psi, V = _pair(shape, C64, _abi.WFS_METHOD_FOURIER, 1, None, 8)
    _check(plans, psi, V, 9, 3, boxes[:3])
    plans, psi, V = _pair(shape, C64, _abi.WFS_METHOD_FOURIER, 1, None, 8)
    _check(plans, psi, V, 9, 3, boxes[1:])
