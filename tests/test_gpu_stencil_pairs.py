"""Power series: fused stage pairs against one launch per stage."""

import numpy as np
import pytest
import torch

import oracle
from wavefunctionsimpython_amd import _abi

pytestmark = pytest.mark.gpu

def _state(n, precision, dev, seed=0):
    rng = np.random.default_rng(seed)
    shape = tuple(n)
    psi = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * 0.1
    V = rng.uniform(-0.5, 0.5, shape) + 1j * rng.uniform(-0.05, 0.0, shape)
    dt = torch.complex128 if precision == _abi.WFS_C128 else torch.complex64
    return (torch.from_numpy(psi).to(dev, dt).contiguous(),
            torch.from_numpy(V.astype(np.complex64)).to(dev).contiguous())


@pytest.mark.parametrize("precision", [_abi.WFS_C64, _abi.WFS_C128])
@pytest.mark.parametrize("n,order,batch", [((64, 32, 96), 10, 1), ((32, 32, 32), 5, 1), ((16, 64, 32), 4, 2),
                                          ((8, 8, 8), 3, 1), ((96, 16, 8), 10, 1)])
def test_pairs_equal_single_stages(dev, precision, n, order, batch):
    shape = tuple(n) if batch == 1 else (batch,) + tuple(n)
    psi, V = _state(shape, precision, dev)
    dx = (0.4, 0.5, 0.3)
    single = _abi.Plan(n, dx, 0.01, 1.0, precision, _abi.WFS_METHOD_POWER_SERIES, order=order, batch=batch)
    single.set_variant("horner_stage", 0)
    a = psi.clone()
    single.step(a, V, 3)
    for variant in (1, 2, 3, 4, 5):
        paired = _abi.Plan(n, dx, 0.01, 1.0, precision, _abi.WFS_METHOD_POWER_SERIES, order=order, batch=batch)
        paired.set_variant("horner_stage", variant)
        b = psi.clone()
        paired.step(b, V, 3)
        torch.cuda.synchronize()
        assert torch.equal(a, b), f"variant {variant}"


def test_pairs_match_oracle_recurrence(dev):
    n = (32, 16, 32)
    psi, V = _state(n, _abi.WFS_C128, dev, seed=3)
    plan = _abi.Plan(n, (0.5, 0.5, 0.5), 0.01, 1.0, _abi.WFS_C128, _abi.WFS_METHOD_POWER_SERIES, order=10)
    ref = oracle.power_series_step(psi.cpu().numpy(), V.cpu().numpy(), 0.01, (0.5, 0.5, 0.5), 1.0, order=10,
                                   nsteps=2)
    plan.step(psi, V, 2)
    assert oracle.relative_l2(psi.cpu().numpy(), ref) < 1e-12
