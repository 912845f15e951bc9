"""Float64 restatement of the FNO's device kernels, for checking its passes.

``EmulatedFNO`` runs the network's own host code (every strided GEMM call and
elementwise call of ``FNO3d``) with the device kernels replaced by float64
torch code of the documented semantics.  With ``bf16_operands`` every GEMM
operand is rounded to bf16 (nearest even) first, as the matrix-core kernel
does (as hi + lo parts for the operands the network splits), so the GPU
network can be compared with it closely.
"""

import numpy as np
import torch

from oracle import fno_oracle
from wavefunctionsimpython_amd.fno import FNO3d


def round_bf16(t: torch.Tensor) -> torch.Tensor:
    """float32 -> bf16 (round to nearest even) -> float64, elementwise."""
    a = np.ascontiguousarray(t.detach().cpu().numpy().astype(np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return torch.from_numpy(u.astype(np.uint32).view(np.float32).astype(np.float64)).to(t.device)


class EmulatedFNO(FNO3d):
    """FNO3d with the device kernels replaced by their float64 semantics."""

    dtype = torch.float64
    bf16_operands = False

    def _gemm(self, A, B, C, batch, M, N, K, sA, sB, sC, alpha=1.0, beta=0.0, splits=1, flags=0, bias=None,
              act=None):
        for t, st, dims in ((A, sA, (batch, M, K)), (B, sB, (batch, K, N)), (C, sC, (batch, M, N))):
            span = sum((d - 1) * s for d, s in zip(dims, st))
            assert t.is_contiguous() and min(st) >= 0 and span < t.numel()
        Av = torch.as_strided(A, (batch, M, K), sA)
        Bv = torch.as_strided(B, (batch, K, N), sB)
        Cv = torch.as_strided(C, (batch, M, N), sC)
        if self.bf16_operands:
            split = flags if self.split_activations else 0
            Ah, Bh = round_bf16(Av), round_bf16(Bv)
            prod = torch.bmm(Ah, Bh)
            if split & 1:   # A as hi + lo
                prod = prod + torch.bmm(round_bf16(Av.float().double() - Ah), Bh)
            if split & 2:
                prod = prod + torch.bmm(Ah, round_bf16(Bv.float().double() - Bh))
            r = alpha * prod
        else:
            r = alpha * torch.bmm(Av, Bv)
        if beta != 0.0:
            r = r + beta * Cv
        if bias is not None:
            r = r + bias[:M].view(1, M, 1)
        if act is not None:
            torch.as_strided(act, (batch, M, N), sC).copy_(fno_oracle.gelu_tanh(r))
        if not flags & 4:
            Cv.copy_(r)

    def _modes(self, Rr, Ri, X, O):
        """The per-mode complex mixing as the four GEMMs it replaces (same operand splits)."""
        C, m3 = self.C, self.m ** 3
        Xr, Xi, Or, Oi = X, X[m3:], O, O[m3:]
        sR, sX = (C * C, 1, C), (1, 2 * m3, 0)
        self._gemm(Rr, Xr, Or, m3, C, 1, C, sR, sX, sX, flags=2)
        self._gemm(Ri, Xi, Or, m3, C, 1, C, sR, sX, sX, alpha=-1.0, beta=1.0, flags=2)
        self._gemm(Ri, Xr, Oi, m3, C, 1, C, sR, sX, sX, flags=2)
        self._gemm(Rr, Xi, Oi, m3, C, 1, C, sR, sX, sX, beta=1.0, flags=2)

    def _layer(self, W, v, Z, bias, u, vn, T, cin, act):
        """The fused layer as the GEMM composition it replaces: u = W v + Z Gz + b,
        vn = GELU(u) or u, T = vn Fz (same operand splits)."""
        C, n, P, tm = self.C, self.n, self.P, 2 * self.m
        r = torch.empty(C * P, dtype=self.dtype, device=v.device)
        self._gemm(W, v, r, 1, C, P, cin, (0, cin, 1), (0, P, 1), (0, P, 1), flags=2)
        if Z is not None:
            S = torch.empty_like(r)
            self._gemm(Z, self.tab["Gz"], S, 1, C * n * n, n, tm, (0, tm, 1), (0, n, 1), (0, n, 1), flags=1)
            r = r + S
        r = (r.view(C, P) + bias[:C].view(C, 1)).reshape(-1)
        if u is not None:
            u.copy_(r)
        g = fno_oracle.gelu_tanh(r) if act else r
        if vn is not None:
            vn.copy_(g)
        if T is not None:
            self._gemm(vn, self.tab["Fz"], T, 1, C * n * n, tm, n, (0, n, 1), (0, tm, 1), (0, tm, 1), flags=1)

    def _gelu_backward(self, u, dv, du):
        ut = u.detach().clone().requires_grad_(True)
        (g,) = torch.autograd.grad(fno_oracle.gelu_tanh(ut), ut, dv)
        du.copy_(g)

    def _axpby(self, y, x, alpha, beta):
        y.copy_(alpha * x + beta * y if beta != 0.0 else alpha * x)
