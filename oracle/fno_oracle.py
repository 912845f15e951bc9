"""FP64 CPU oracle of the Fourier neural operator and its gradients.

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` as the checker of the GPU
network (``wavefunctionsimpython_amd.fno``); never shipped in or called by the
product path.  The reference has no learned component, so nothing here cites
it: this is the definition the GPU network must match.

The network (channel-major activations v[c] on an n^3 grid):

    v0      = P a + pb
    u_l     = W_l v_l + b_l + Re IFFT( scatter_m( R_l (.) gather_m(FFT(v_l)) ) )
    v_{l+1} = GELU_tanh(u_l) for l < L-1,   v_L = u_{L-1}
    out     = a[0:2] + r (Q v_L + qb)   (residual; r the residual scale)
              Q v_L + qb                (no residual)

gather_m keeps the m signed frequencies f in {0..m/2-1, -m/2..-1} per axis of
the full torch.fft spectrum (index f mod n), so the truncation and its table
construction on the GPU are checked against a plain FFT; R_l[q][i][o] is the
complex weight of mode q = (ix*m + iy)*m + iz from input channel i to output
channel o.  Gradients come from torch.autograd in float64.

Pins (tests/test_oracle_fno.py): m = n with R = identity reduces the spectral
path to the identity; a lone zero-frequency weight reduces it to the channel
mean; torch.autograd.gradcheck (central finite differences) on a 4^3 case.
"""

from __future__ import annotations

import math
from typing import Dict

import numpy as np
import torch


def mode_indices(n: int, m: int) -> np.ndarray:
    """Array indices (f mod n) of the kept signed frequencies, in kept order."""
    f = np.array([k if k < m // 2 else k - m for k in range(m)], dtype=np.int64)
    return np.mod(f, n)


def gelu_tanh(u: torch.Tensor) -> torch.Tensor:
    k = math.sqrt(2.0 / math.pi)
    return 0.5 * u * (1.0 + torch.tanh(k * (u + 0.044715 * u ** 3)))


def spectral(v: torch.Tensor, Rr: torch.Tensor, Ri: torch.Tensor, n: int, m: int) -> torch.Tensor:
    """Re IFFT of the mode-truncated, channel-mixed spectrum of v [C][n^3] -> [C][n^3]."""
    C = v.shape[0]
    Co = Rr.shape[2]
    idx = torch.from_numpy(mode_indices(n, m))
    V = torch.fft.fftn(v.reshape(C, n, n, n), dim=(1, 2, 3))
    Xs = V[:, idx][:, :, idx][:, :, :, idx].reshape(C, m ** 3)          # [C][q]
    R = torch.complex(Rr, Ri)                                           # [q][i][o]
    O = torch.einsum("qio,iq->oq", R, Xs)                               # [Co][q]
    full = torch.zeros(Co, n, n, n, dtype=V.dtype)
    ix = idx.view(m, 1, 1).expand(m, m, m).reshape(-1)
    iy = idx.view(1, m, 1).expand(m, m, m).reshape(-1)
    iz = idx.view(1, 1, m).expand(m, m, m).reshape(-1)
    full = full.index_put((torch.arange(Co).view(Co, 1), ix.view(1, -1), iy.view(1, -1), iz.view(1, -1)), O)
    return torch.fft.ifftn(full, dim=(1, 2, 3)).real.reshape(Co, n ** 3)


def forward(params: Dict[str, torch.Tensor], a: torch.Tensor, n: int, m: int, layers: int,
            residual: bool = True, rscale: float = 1.0) -> torch.Tensor:
    """FNO forward in float64: a [cin][n^3] -> out [cout][n^3]."""
    v = params["P"] @ a + params["pb"][:, None]
    for l in range(layers):
        u = params[f"W{l}"] @ v + params[f"b{l}"][:, None] + spectral(v, params[f"Rr{l}"], params[f"Ri{l}"], n, m)
        v = gelu_tanh(u) if l < layers - 1 else u
    out = params["Q"] @ v + params["qb"][:, None]
    if residual:
        out = a[:2] + rscale * out
    return out


def relative_loss(out: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
    """sum (out - target)^2 / sum target^2."""
    return ((out - target) ** 2).sum() / (target ** 2).sum()


def loss_and_grads(params_np: Dict[str, np.ndarray], a: np.ndarray, target: np.ndarray, n: int, m: int,
                   layers: int, residual: bool = True, rscale: float = 1.0):
    """(loss, out, {name: dloss/dparam}) in float64 for numpy inputs."""
    params = {k: torch.tensor(np.asarray(v, dtype=np.float64), requires_grad=True) for k, v in params_np.items()}
    at = torch.tensor(np.asarray(a, dtype=np.float64))
    tt = torch.tensor(np.asarray(target, dtype=np.float64))
    out = forward(params, at, n, m, layers, residual, rscale)
    loss = relative_loss(out, tt)
    names = list(params)
    grads = torch.autograd.grad(loss, [params[k] for k in names])
    return float(loss.detach()), out.detach().numpy(), {k: g.numpy() for k, g in zip(names, grads)}
