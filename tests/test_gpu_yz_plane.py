"""Fused y-z plane pass of the split step.

Where ny == nz fits a plane in LDS (complex64 up to 128, complex128 up to 64) each step runs
the y inverse, z inverse / potential / z forward and y forward transforms as one kernel on
whole x planes.  It performs the same arithmetic as the three axis passes (same splits,
twiddle table copies and operation order), so the states after several steps must be equal
bit for bit; the complex128 plane path is also pinned to the FP64 oracle.
"""

import numpy as np
import pytest
import torch

import oracle
from wavefunctionsimpython_amd import _abi

pytestmark = pytest.mark.gpu

C64, C128 = _abi.WFS_C64, _abi.WFS_C128
DX = (0.3, 0.25, 0.35)
CASES = [(C64, 16), (C64, 32), (C64, 64), (C64, 96), (C64, 128), (C128, 16), (C128, 32), (C128, 64)]


def _cdtype(precision):
    return torch.complex64 if precision == C64 else torch.complex128


def _inputs(shape, precision, seed, dev):
    rng = np.random.default_rng(seed)
    psi = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex128)
    V = (rng.standard_normal(shape) * 0.5 + 1j * rng.uniform(-0.05, 0.0, shape)).astype(np.complex64)
    return torch.from_numpy(psi).to(dev).to(_cdtype(precision)), torch.from_numpy(V).to(dev)


def _axes(rng, batch, n):
    return [rng.uniform(-1.0, 1.0, (batch, m)) for m in n]


def _run(shape, precision, variant, psi0, V=None, axes=None, batch=1, steps=5):
    p = _abi.Plan(shape, DX, 0.01, 1.0, precision, _abi.WFS_METHOD_FOURIER, batch=batch)
    p.set_variant("yz_slab", 0)      # the plane pass itself, not the two-launch form at 128
    p.set_variant("yz_plane", variant)
    psi = psi0.clone()
    if axes is not None:
        p.set_potential_axes(*axes)
        p.step(psi, None, steps)
    else:
        p.step(psi, V, steps)
    torch.cuda.synchronize()
    return psi


def _variants(precision, n):
    """plane variants: lines per thread, the plane passing through LDS in that many rounds"""
    return (1, 2) if (precision == C64 and n in (64, 128)) or (precision == C128 and n == 64) else (1,)


@pytest.mark.parametrize("precision,n", CASES)
def test_plane_pass_equals_axis_passes_table_potential(dev, precision, n):
    shape = (16, n, n)
    psi0, V = _inputs(shape, precision, 7 + n, dev)
    a = _run(shape, precision, 0, psi0, V)
    for v in _variants(precision, n):
        print(f"case: table potential, precision {precision}, shape {shape}, variant {v}", flush=True)
        b = _run(shape, precision, v, psi0, V)
        print(f"  max |fused - passes| = {float((a - b).abs().max()):.3e}")
        assert torch.equal(a, b)


@pytest.mark.parametrize("precision,n", [(C64, 64), (C64, 96), (C64, 128), (C128, 32), (C128, 64)])
def test_plane_pass_equals_axis_passes_separable_batched(dev, precision, n):
    batch = 3
    shape = (8, n, n)
    psi0, _ = _inputs((batch,) + shape, precision, 11 + n, dev)
    axes = _axes(np.random.default_rng(n), batch, shape)
    a = _run(shape, precision, 0, psi0, axes=axes, batch=batch)
    for v in _variants(precision, n):
        print(f"case: separable batched, precision {precision}, batch {batch}, shape {shape}, variant {v}",
              flush=True)
        b = _run(shape, precision, v, psi0, axes=axes, batch=batch)
        assert torch.equal(a, b)


def test_plane_pass_c128_matches_oracle(dev):
    shape = (16, 32, 32)
    dx = DX
    psi0, V = _inputs(shape, C128, 3, dev)
    p = _abi.Plan(shape, dx, 0.01, 1.0, C128, _abi.WFS_METHOD_FOURIER)
    assert p.variants()["yz_plane"] == 1
    psi = psi0.clone()
    p.step(psi, V, 4)
    ref = oracle.split_step(psi0.cpu().numpy(), V.cpu().numpy(), dx, 0.01, 1.0, nsteps=4)
    err = oracle.relative_l2(psi.cpu().numpy(), ref)
    print(f"fused plane c128 vs oracle: {err:.2e}")
    assert err < 1e-12


def test_plane_pass_defaults_and_limits(dev):
    def variant(shape, precision, batch=1):
        return _abi.Plan(shape, (0.3,) * 3, 0.01, 1.0, precision, _abi.WFS_METHOD_FOURIER,
                         batch=batch).variants()["yz_plane"]
    assert variant((128, 128, 128), C64) == 1          # 128 planes: one line per thread
    assert variant((128, 128, 128), C64, batch=4) == 2  # 512 planes: two blocks per CU
    assert variant((32, 96, 96), C64) == 1
    assert variant((32, 128, 128), C128) == 0         # 256 KiB plane: three passes
    assert variant((32, 192, 192), C64) == 0            # 288 KiB plane: three passes
    assert variant((32, 256, 256), C64) == 0
    assert variant((32, 64, 128), C64) == 0           # ny != nz
    p = _abi.Plan((16, 64, 32), (0.3,) * 3, 0.01, 1.0, C64, _abi.WFS_METHOD_FOURIER)
    with pytest.raises(_abi.WfsError):
        p.set_variant("yz_plane", 1)
    p = _abi.Plan((16, 96, 96), (0.3,) * 3, 0.01, 1.0, C64, _abi.WFS_METHOD_FOURIER)
    with pytest.raises(_abi.WfsError):
        p.set_variant("yz_plane", 2)
    p.set_variant("yz_plane", 0)


def test_autotune_picks_a_plane_variant_with_identical_result(dev):
    shape = (64, 64, 64)
    psi0, V = _inputs(shape, C64, 5, dev)
    p = _abi.Plan(shape, DX, 0.01, 1.0, C64, _abi.WFS_METHOD_FOURIER)
    p.set_potential(V)
    chosen = p.autotune()
    print(f"autotuned yz_plane variant at 64^3: {chosen['yz_plane']}")
    assert chosen["yz_plane"] in (0, 1, 2)
    psi = psi0.clone()
    p.step(psi, None, 5)
    ref = _run(shape, C64, 0, psi0, V)
    assert torch.equal(psi, ref)
