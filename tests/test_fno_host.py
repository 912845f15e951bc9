"""Host logic of the Fourier neural operator, checked on the CPU.

The network's passes are strided GEMM calls plus elementwise kernels.  Here the
GEMM and the elementwise kernels are replaced by float64 torch restatements of
their documented semantics, so the stride bookkeeping of every forward and
backward contraction is checked against the FP64 autograd oracle without a GPU.
"""

import numpy as np
import pytest
import torch

from oracle import fno_oracle
from fno_emulation import EmulatedFNO
from wavefunctionsimpython_amd.fno import FNO3d


CASES = [  # n, width, modes, layers, cin, residual scale
    (8, 4, 2, 2, 3, 1.0),
    (8, 6, 4, 3, 3, 1.0),
    (6, 5, 6, 2, 2, 1.0),
    (8, 4, 4, 2, 5, 0.01),
]


@pytest.mark.parametrize("n,width,modes,layers,cin,rscale", CASES)
def test_fno_passes_match_fp64_autograd(n, width, modes, layers, cin, rscale):
    net = EmulatedFNO(n, width=width, modes=modes, layers=layers, in_channels=cin, seed=5,
                      device=torch.device("cpu"), splits=4, residual_scale=rscale)
    # larger output weights so the network term is not hidden by the residual
    net.params["Q"].mul_(10.0)
    rng = np.random.default_rng(11)
    P = n ** 3
    a = rng.standard_normal((cin, P))
    target = a[:2] + 0.3 * rng.standard_normal((2, P))
    params = {k: v.numpy().copy() for k, v in net.params.items()}
    loss_ref, out_ref, grads_ref = fno_oracle.loss_and_grads(params, a, target, n, modes, layers, rscale=rscale)

    at = torch.from_numpy(a.reshape(-1).copy())
    out = net.forward(at, keep=True).view(2, P)
    assert np.abs(out.numpy() - out_ref).max() < 1e-11 * np.abs(out_ref).max()
    tt = torch.from_numpy(target)
    tn = float((tt ** 2).sum())
    dout = (2.0 * (out - tt) / tn).reshape(-1).contiguous()
    net.zero_grad()
    net.backward(dout)
    for k, g in grads_ref.items():
        got = net.grads[k].numpy().reshape(g.shape)
        assert np.linalg.norm(got - g) <= 1e-10 * np.linalg.norm(g) + 1e-300, k


def test_fno_rejects_bad_modes():
    with pytest.raises(ValueError):
        FNO3d(8, modes=3, device=torch.device("cpu"))
    with pytest.raises(ValueError):
        FNO3d(8, modes=10, device=torch.device("cpu"))


def test_fno_inference_forward_equals_training_forward():
    """Without keep the layer outputs skip storing the pre-activations; the result is the same."""
    net = EmulatedFNO(8, width=4, modes=4, layers=3, in_channels=3, seed=2, device=torch.device("cpu"))
    a = torch.from_numpy(np.random.default_rng(1).standard_normal(3 * 8 ** 3))
    kept = net.forward(a, keep=True).clone()
    net._saved = None
    plain = net.forward(a, keep=False)
    assert torch.equal(kept, plain)
