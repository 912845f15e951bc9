"""Separable potentials: per-axis phase factors in the potential pass."""

import numpy as np
import pytest
import torch

import oracle
from wavefunctionsimpython_amd import _abi

pytestmark = pytest.mark.gpu

DX = (0.4, 0.5, 0.3)


def _setup(precision, dev):
    n = (64, 32, 96)
    axes = _abi.separable_axes(n, DX, omega=0.3, mass=1.0, center=(0.5, -0.3, 0.2), depth=1.2, period=3.0)
    psi = torch.empty(n, dtype=torch.complex128 if precision == _abi.WFS_C128 else torch.complex64, device=dev)
    _abi.init_gaussian(n, DX, precision, (2.0, 1.5, 2.5), (1.0, 0.0, -1.0), (0.5, -0.4, 0.8), psi)
    V64 = axes[0][:, None, None] + axes[1][None, :, None] + axes[2][None, None, :]
    return n, axes, psi, V64


@pytest.mark.parametrize("precision,tol", [(_abi.WFS_C128, 1e-12), (_abi.WFS_C64, 1e-5)])
def test_separable_matches_oracle_with_exact_potential(dev, precision, tol):
    n, axes, psi, V64 = _setup(precision, dev)
    plan = _abi.Plan(n, DX, 0.01, 1.0, precision, _abi.WFS_METHOD_FOURIER)
    plan.set_potential_axes(*axes)
    ref = oracle.split_step(psi.cpu().numpy().astype(np.complex128), V64.astype(np.complex128), DX, 0.01, 1.0,
                            nsteps=60)
    plan.step(psi, None, 60)
    err = oracle.relative_l2(psi.cpu().numpy(), ref)
    print(f"separable, 60 steps: rel L2 {err:.2e}")
    assert err < tol


def test_separable_close_to_table_and_table_restored(dev):
    n, axes, psi, V64 = _setup(_abi.WFS_C64, dev)
    V = torch.from_numpy(V64.astype(np.complex64)).to(dev)
    sep = _abi.Plan(n, DX, 0.01, 1.0, _abi.WFS_C64, _abi.WFS_METHOD_FOURIER)
    sep.set_potential_axes(*axes)
    tab = _abi.Plan(n, DX, 0.01, 1.0, _abi.WFS_C64, _abi.WFS_METHOD_FOURIER)
    tab.set_potential(V)
    a, b = psi.clone(), psi.clone()
    sep.step(a, None, 40)
    tab.step(b, None, 40)
    torch.cuda.synchronize()
    assert oracle.relative_l2(a.cpu().numpy(), b.cpu().numpy()) < 5e-6   # two roundings of the same V
    # a table set afterwards replaces the separable factors
    sep.set_potential(V)
    c = psi.clone()
    sep.step(c, None, 40)
    torch.cuda.synchronize()
    assert torch.equal(c, b)


def test_separable_rejects_bad_sizes(dev):
    plan = _abi.Plan((32, 32, 32), (0.5,) * 3, 0.01, 1.0, _abi.WFS_C64, _abi.WFS_METHOD_FOURIER)
    with pytest.raises(ValueError):
        plan.set_potential_axes(np.zeros(31), np.zeros(32), np.zeros(32))
