"""Fourier neural operator on the bf16 matrix cores (no reference counterpart).

Architecture (channel-major FP32 activations v[c][p], p over the n^3 grid):

    v0      = P a + pb                                   (lift, a = [Re psi, Im psi, s*Re V])
    u_l     = W_l v_l + b_l + Re IDFT_m( R_l (.) DFT_m(v_l) )
    v_{l+1} = GELU(u_l)  (l < L-1),   v_L = u_{L-1}
    out     = Q v_L + qb + a[0:2]                         (residual on the state planes)

DFT_m is the 3-D DFT restricted to the m lowest frequencies per axis,
f in {0..m/2-1, -m/2..-1}; it is evaluated as three successive contractions
with small cos/sin tables (z, then y, then x), each one call of the strided
bf16 GEMM ``wfs_gemm_bf16``; IDFT_m (1/n per axis) likewise in reverse.  The
complex spectral weights R_l[q][i][o] mix channels per mode q.  Every weight
and activation enters the matrix cores as bf16; accumulation is FP32.  The
backward pass is the same contractions with transposed strides; weight
gradients use deterministic split-K reductions.  Adam updates FP32 master
weights.  Only elementwise work (bias, GELU, loss, Adam) runs outside the
GEMM, in the library's own kernels.
"""

from __future__ import annotations

import math
from typing import Dict, List

import numpy as np
import torch

from . import _abi

# operands carried as bf16 hi + lo parts: activations and gradients (weights and tables stay bf16)
SA, SB, NO_C = _abi.GEMM_SPLIT_A, _abi.GEMM_SPLIT_B, _abi.GEMM_NO_C
SAB = SA | SB


def mode_frequencies(m: int) -> np.ndarray:
    """Signed frequencies of the m kept modes: 0..m/2-1, -m/2..-1."""
    return np.array([k if k < m // 2 else k - m for k in range(m)], dtype=np.int64)


def dft_tables(n: int, m: int) -> Dict[str, np.ndarray]:
    """Forward/inverse truncated-DFT tables (FP64 on the host, stored FP32).

    Fz[z][kk]        : cos(th) for kk < m, -sin(th) for kk >= m          (z stage, e^{-i th})
    F2[k'][2y+part]  : k' < m: (cos, sin); k' >= m: (-sin, cos)          (y / x stages)
    G2[2x+part][k']  : part 0: (cos, -sin) / n; part 1: (sin, cos) / n  (inverse x / y)
    Gz[k'][z]        : part 0: cos / n; part 1: -sin / n                 (inverse z, real part)
    with th = 2 pi f z / n.
    """
    f = mode_frequencies(m)
    z = np.arange(n)
    th = 2.0 * np.pi * np.outer(z, f) / n          # [n][m]
    c, s = np.cos(th), np.sin(th)
    Fz = np.concatenate([c, -s], axis=1)            # [n][2m]
    F2 = np.zeros((2 * m, 2 * n))
    F2[:m, 0::2] = c.T
    F2[:m, 1::2] = s.T
    F2[m:, 0::2] = -s.T
    F2[m:, 1::2] = c.T
    G2 = np.zeros((2 * n, 2 * m))
    G2[0::2, :m] = c / n
    G2[0::2, m:] = -s / n
    G2[1::2, :m] = s / n
    G2[1::2, m:] = c / n
    Gz = np.concatenate([c.T / n, -s.T / n], axis=0)   # [2m][n]
    return {k: np.ascontiguousarray(t) for k, t in (("Fz", Fz), ("F2", F2), ("G2", G2), ("Gz", Gz))}


class FNO3d:
    """Weights, activations and the forward/backward passes of one FNO."""

    dtype = torch.float32

    def __init__(self, n: int, width: int = 32, modes: int = 12, layers: int = 4, in_channels: int = 3,
                 out_channels: int = 2, residual: bool = True, seed: int = 0, device=None, splits: int = 2048,
                 output_scale: float = 0.1, residual_scale: float = 1.0, split_activations: bool = True,
                 fused: bool = True):
        if modes % 2 or modes > n:
            raise ValueError("modes must be even and at most n")
        self.n, self.C, self.m, self.L = int(n), int(width), int(modes), int(layers)
        self.cin, self.cout = int(in_channels), int(out_channels)
        self.residual = residual and self.cout == 2 and self.cin >= 2
        self.P = self.n ** 3
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.splits = int(splits)
        # out = a[0:2] + residual_scale * (Q v + qb): a small scale lets the network learn a small
        # correction with O(1) weights and outputs
        self.rscale = float(residual_scale)
        # activations enter the matrix cores as bf16 hi + lo parts, so the network is a smooth
        # function of its input to about 16 bits (parareal differences G(U') - G(U) need that)
        self.split_activations = bool(split_activations)
        # layers as one fused kernel per grid line (wfs_fno_layer) where the sizes allow
        self.fused = (bool(fused) and self.split_activations and self.n % 32 == 0 and 32 <= self.n <= 128
                      and self.C <= 32 and self.cin <= 32 and 2 * self.m <= 32 and self.m % 2 == 0)
        rng = np.random.default_rng(seed)
        C, m3 = self.C, self.m ** 3

        def u(shape, bound):
            return rng.uniform(-bound, bound, size=shape).astype(np.float32)

        p: Dict[str, np.ndarray] = {
            "P": u((C, self.cin), 1.0 / math.sqrt(self.cin)), "pb": u((C,), 1.0 / math.sqrt(self.cin)),
            "Q": u((self.cout, C), 1.0 / math.sqrt(C)) * float(output_scale), "qb": np.zeros((self.cout,), np.float32),
        }
        for l in range(self.L):
            p[f"W{l}"] = u((C, C), 1.0 / math.sqrt(C))
            p[f"b{l}"] = u((C,), 1.0 / math.sqrt(C))
            scale = 1.0 / (C * C)
            p[f"Rr{l}"] = (scale * rng.random((m3, C, C))).astype(np.float32)
            p[f"Ri{l}"] = (scale * rng.random((m3, C, C))).astype(np.float32)
        self.params = {k: torch.from_numpy(v).to(self.device, self.dtype) for k, v in p.items()}
        self.grads = {k: torch.zeros_like(v) for k, v in self.params.items()}
        self.adam_m = {k: torch.zeros_like(v) for k, v in self.params.items()}
        self.adam_v = {k: torch.zeros_like(v) for k, v in self.params.items()}
        self.adam_step = 0
        t = dft_tables(self.n, self.m)
        self.tab = {k: torch.from_numpy(v).to(self.device, self.dtype).contiguous() for k, v in t.items()}
        self.ones = torch.ones(1, dtype=self.dtype, device=self.device)
        self._ws = None
        self._saved = None

    # ------------------------------------------------------------------ buffers
    def _workspace(self):
        if self._ws is None:
            C, n, m, P = self.C, self.n, self.m, self.P
            f = dict(dtype=self.dtype, device=self.device)
            self._ws = {
                "T": torch.empty(C * n * n * 2 * m, **f), "U": torch.empty(C * n * 2 * m * m, **f),
                "X": torch.empty(C * 2 * m ** 3, **f), "O": torch.empty(C * 2 * m ** 3, **f),
                "Y": torch.empty(C * 2 * n * m * m, **f), "Z": torch.empty(C * n * 2 * n * m, **f),
                "d1": torch.empty(C * P, **f), "d2": torch.empty(C * P, **f), "d3": torch.empty(C * P, **f),
                "out": torch.empty(self.cout * P, **f), "dq": torch.empty(self.cout * P, **f),
                "work": torch.empty(self.splits * C * max(C, self.cin), **f),
                "mse": torch.empty(8192, dtype=torch.float64, device=self.device),
                "mse_out": torch.empty(2, dtype=torch.float64, device=self.device),
            }
        return self._ws

    def _gemm(self, A, B, C, batch, M, N, K, sA, sB, sC, alpha=1.0, beta=0.0, splits=1, flags=0, bias=None,
              act=None):
        """C = alpha A B + beta C (+ bias[m]); act = GELU(C) when given; flags SA/SB split the
        activation operands (dropped when split_activations is off), NO_C skips storing C."""
        ws = self._workspace()
        if not self.split_activations:
            flags &= ~SAB
        _abi.gemm_bf16(A, B, C, batch, M, N, K, sA, sB, sC, alpha, beta, splits,
                       ws["work"] if splits > 1 else None, flags=flags, bias=bias, act=act)

    # ------------------------------------------------------------------ spectral convolution
    def _spectral(self, v: torch.Tensor, l: int, S: torch.Tensor) -> None:
        """S[c][p] = Re IDFT_m(R_l (.) DFT_m(v)) (written, not accumulated)."""
        ws = self._workspace()
        C, n, m = self.C, self.n, self.m
        tm = 2 * m
        tb = self.tab
        self._gemm(v, tb["Fz"], ws["T"], 1, C * n * n, tm, n, (0, n, 1), (0, tm, 1), (0, tm, 1), flags=SA)
        self._spectral_mid(l)
        self._gemm(ws["Z"], tb["Gz"], S, 1, C * n * n, n, tm, (0, tm, 1), (0, n, 1), (0, n, 1), flags=SA)

    def _spectral_mid(self, l: int) -> None:
        """Z = the spectral path of layer l between its z stages: T (the z-stage DFT of v, in the
        workspace) -> y and x DFTs -> R_l -> inverse x and y DFTs -> Z (workspace)."""
        ws = self._workspace()
        C, n, m = self.C, self.n, self.m
        m2, m3, tm = m * m, m ** 3, 2 * m
        tb = self.tab
        T, U, X, O, Y, Z = ws["T"], ws["U"], ws["X"], ws["O"], ws["Y"], ws["Z"]
        self._gemm(tb["F2"], T, U, C * n, tm, m, 2 * n, (0, 2 * n, 1), (2 * n * m, m, 1), (tm * m, m, 1), flags=SB)
        self._gemm(tb["F2"], U, X, C, tm, m2, 2 * n, (0, 2 * n, 1), (2 * n * m2, m2, 1), (tm * m2, m2, 1), flags=SB)
        Rr, Ri = self.params[f"Rr{l}"], self.params[f"Ri{l}"]
        Xr, Xi, Or, Oi = X, X[m3:], O, O[m3:]
        self._modes(Rr, Ri, X, O)
        self._gemm(tb["G2"], O, Y, C, 2 * n, m2, tm, (0, tm, 1), (tm * m2, m2, 1), (2 * n * m2, m2, 1), flags=SB)
        self._gemm(tb["G2"], Y, Z, C * n, 2 * n, m, tm, (0, tm, 1), (tm * m, m, 1), (2 * n * m, m, 1), flags=SB)

    def _modes(self, Rr, Ri, X, O) -> None:
        """O = R (.) X per mode, complex (Or = Rr Xr - Ri Xi, Oi = Ri Xr + Rr Xi); X, O [C][2][m^3]."""
        C, m3 = self.C, self.m ** 3
        if self.split_activations:
            _abi.fno_modes(Rr, Ri, X, O, m3, C)
            return
        Xr, Xi, Or, Oi = X, X[m3:], O, O[m3:]
        sR, sX = (C * C, 1, C), (1, 2 * m3, 0)
        self._gemm(Rr, Xr, Or, m3, C, 1, C, sR, sX, sX, flags=SB)
        self._gemm(Ri, Xi, Or, m3, C, 1, C, sR, sX, sX, alpha=-1.0, beta=1.0, flags=SB)
        self._gemm(Ri, Xr, Oi, m3, C, 1, C, sR, sX, sX, flags=SB)
        self._gemm(Rr, Xi, Oi, m3, C, 1, C, sR, sX, sX, beta=1.0, flags=SB)

    def _layer(self, W, v, Z, bias, u, vn, T, cin: int, act: bool) -> None:
        """u = W v + Z Gz + bias, vn = GELU(u) or u, T = vn Fz in one pass over the grid lines."""
        _abi.fno_layer(W, v, Z, bias, self.tab["Gz"], self.tab["Fz"], u, vn, T, cin, self.C, self.n, 2 * self.m,
                       act)

    def _spectral_backward(self, dS: torch.Tensor, l: int, v: torch.Tensor, dv: torch.Tensor) -> None:
        """dv += d/dv of the spectral path given dS; accumulates dR_l (needs X of this layer)."""
        ws = self._workspace()
        C, n, m = self.C, self.n, self.m
        m2, m3, tm = m * m, m ** 3, 2 * m
        tb = self.tab
        T, U, X, O, Y, Z = ws["T"], ws["U"], ws["X"], ws["O"], ws["Y"], ws["Z"]
        # recompute the forward modes X of this layer (cheaper than storing them per layer)
        self._gemm(v, tb["Fz"], T, 1, C * n * n, tm, n, (0, n, 1), (0, tm, 1), (0, tm, 1), flags=SA)
        self._gemm(tb["F2"], T, U, C * n, tm, m, 2 * n, (0, 2 * n, 1), (2 * n * m, m, 1), (tm * m, m, 1), flags=SB)
        self._gemm(tb["F2"], U, X, C, tm, m2, 2 * n, (0, 2 * n, 1), (2 * n * m2, m2, 1), (tm * m2, m2, 1), flags=SB)
        # inverse stages, transposed: dZ, dY, dO (reuse Z, Y, O)
        self._gemm(dS, tb["Gz"], Z, 1, C * n * n, tm, n, (0, n, 1), (0, 1, n), (0, tm, 1), flags=SA)
        self._gemm(tb["G2"], Z, Y, C * n, tm, m, 2 * n, (0, 1, tm), (2 * n * m, m, 1), (tm * m, m, 1), flags=SB)
        self._gemm(tb["G2"], Y, O, C, tm, m2, 2 * n, (0, 1, tm), (2 * n * m2, m2, 1), (tm * m2, m2, 1), flags=SB)
        Rr, Ri = self.params[f"Rr{l}"], self.params[f"Ri{l}"]
        gRr, gRi = self.grads[f"Rr{l}"], self.grads[f"Ri{l}"]
        Xr, Xi, dOr, dOi = X, X[m3:], O, O[m3:]
        sRt, sX = (C * C, C, 1), (1, 2 * m3, 0)
        # weight gradients (accumulated): dRr = Xr dOr^T + Xi dOi^T, dRi = Xr dOi^T - Xi dOr^T
        sA, sB, sC = (1, 2 * m3, 0), (1, 0, 2 * m3), (C * C, C, 1)
        self._gemm(Xr, dOr, gRr, m3, C, C, 1, sA, sB, sC, beta=1.0, flags=SAB)
        self._gemm(Xi, dOi, gRr, m3, C, C, 1, sA, sB, sC, beta=1.0, flags=SAB)
        self._gemm(Xr, dOi, gRi, m3, C, C, 1, sA, sB, sC, beta=1.0, flags=SAB)
        self._gemm(Xi, dOr, gRi, m3, C, C, 1, sA, sB, sC, alpha=-1.0, beta=1.0, flags=SAB)
        # mode gradients into X (overwrites the forward modes, no longer needed)
        dX = torch.empty_like(X)
        dXr, dXi = dX, dX[m3:]
        self._gemm(Rr, dOr, dXr, m3, C, 1, C, sRt, sX, sX, flags=SB)
        self._gemm(Ri, dOi, dXr, m3, C, 1, C, sRt, sX, sX, beta=1.0, flags=SB)
        self._gemm(Ri, dOr, dXi, m3, C, 1, C, sRt, sX, sX, alpha=-1.0, flags=SB)
        self._gemm(Rr, dOi, dXi, m3, C, 1, C, sRt, sX, sX, beta=1.0, flags=SB)
        # forward stages, transposed: dU, dT, then dv += dT Fz^T
        self._gemm(tb["F2"], dX, U, C, 2 * n, m2, tm, (0, 1, 2 * n), (tm * m2, m2, 1), (2 * n * m2, m2, 1), flags=SB)
        self._gemm(tb["F2"], U, T, C * n, 2 * n, m, tm, (0, 1, 2 * n), (tm * m, m, 1), (2 * n * m, m, 1), flags=SB)
        self._gemm(T, tb["Fz"], dv, 1, C * n * n, n, tm, (0, tm, 1), (0, 1, tm), (0, n, 1), beta=1.0, flags=SA)

    # ------------------------------------------------------------------ forward / backward
    # ------------------------------------------------------------------ elementwise kernels
    def _gelu_backward(self, u: torch.Tensor, dv: torch.Tensor, du: torch.Tensor) -> None:
        _abi.check(_abi.lib().wfs_fno_gelu_backward(_abi.ptr(u), _abi.ptr(dv), _abi.ptr(du), u.numel(),
                                                    _abi.stream_handle()), "wfs_fno_gelu_backward")

    def _axpby(self, y: torch.Tensor, x: torch.Tensor, alpha: float, beta: float) -> None:
        """y = alpha x + beta y."""
        _abi.check(_abi.lib().wfs_fno_axpby(_abi.ptr(y), _abi.ptr(x), alpha, beta, y.numel(), _abi.stream_handle()),
                   "wfs_fno_axpby")

    # ------------------------------------------------------------------ forward / backward
    def forward(self, a: torch.Tensor, keep: bool = False) -> torch.Tensor:
        """a: [cin][P] -> out [cout][P] (a view into a workspace buffer, valid until the next call)."""
        if a.dtype != self.dtype or a.numel() != self.cin * self.P or not a.is_contiguous():
            raise ValueError(f"input must be a contiguous {self.dtype} tensor of cin * n^3 elements")
        ws = self._workspace()
        C, P = self.C, self.P
        pr = self.params
        saved_v: List[torch.Tensor] = []
        saved_u: List[torch.Tensor] = []

        def act_buffer(avoid=None):
            # training keeps every activation for the backward pass; inference ping-pongs between
            # the workspace's gradient buffers (no allocation, so a forward can be graph-captured)
            if keep:
                return torch.empty(C * P, dtype=self.dtype, device=self.device)
            return ws["d2"] if avoid is ws["d1"] else ws["d1"]

        v = act_buffer()
        if self.fused:
            # lift and the first z-stage DFT, then per layer the middle of the spectral path and one
            # fused pass: u = W v + (last spectral stage) + b, v' = GELU(u), next z-stage DFT
            self._layer(pr["P"], a, None, pr["pb"], None, v, ws["T"], self.cin, act=False)
            for l in range(self.L):
                self._spectral_mid(l)
                last = l == self.L - 1
                u = (torch.empty(C * P, dtype=self.dtype, device=self.device) if keep
                     else ws["d3"] if last else None)
                vn = None if last else act_buffer(v)
                self._layer(pr[f"W{l}"], v, ws["Z"], pr[f"b{l}"], u, vn, None if last else ws["T"], C, act=True)
                if keep:
                    saved_v.append(v)
                    saved_u.append(u)
                v = u if last else vn
        else:
            self._gemm(pr["P"], a, v, 1, C, P, self.cin, (0, self.cin, 1), (0, P, 1), (0, P, 1), flags=SB,
                       bias=pr["pb"])
            for l in range(self.L):
                u = torch.empty(C * P, dtype=self.dtype, device=self.device) if keep else ws["d3"]
                self._spectral(v, l, u)
                last = l == self.L - 1
                vn = None if last else act_buffer(v)
                # u = W v + S + b and v' = GELU(u) in the GEMM's epilogue; u is stored only when
                # the backward pass needs it (or when it is the network's last activation)
                self._gemm(pr[f"W{l}"], v, u, 1, C, P, C, (0, C, 1), (0, P, 1), (0, P, 1), beta=1.0,
                           flags=SB | (NO_C if (not keep and not last) else 0), bias=pr[f"b{l}"], act=vn)
                if keep:
                    saved_v.append(v)
                    saved_u.append(u)
                v = u if last else vn
        out = ws["out"]
        self._gemm(pr["Q"], v, out, 1, self.cout, P, C, (0, C, 1), (0, P, 1), (0, P, 1), flags=SB,
                   bias=pr["qb"])
        if self.residual:
            # out = a[0:2] + rscale * out elementwise (the GEMM's inputs would round a to bf16)
            self._axpby(out, a[: 2 * P], 1.0, self.rscale)
        elif self.rscale != 1.0:
            self._axpby(out, out, self.rscale, 0.0)
        if keep:
            self._saved = (a, saved_v, saved_u, v)
        return out

    def backward(self, dout: torch.Tensor) -> None:
        """Accumulate parameter gradients for the last forward(keep=True)."""
        if self._saved is None:
            raise RuntimeError("forward(keep=True) must precede backward()")
        a, saved_v, saved_u, vL = self._saved
        ws = self._workspace()
        C, P = self.C, self.P
        pr, gr = self.params, self.grads
        # reductions over the grid (K = P): about 1024 points per split
        sp = max(1, min(self.splits, P // 1024))
        if self.rscale != 1.0:
            dq = ws["dq"]
            self._axpby(dq, dout, self.rscale, 0.0)
            dout = dq
        self._gemm(dout, vL, gr["Q"], 1, self.cout, C, P, (0, P, 1), (0, 1, P), (0, C, 1), beta=1.0, splits=sp, flags=SAB)
        self._gemm(dout, self.ones, gr["qb"], 1, self.cout, 1, P, (0, P, 1), (0, 0, 0), (0, 1, 0), beta=1.0,
                   splits=sp, flags=SA)
        dv = ws["d1"]
        self._gemm(pr["Q"], dout, dv, 1, C, P, self.cout, (0, 1, C), (0, P, 1), (0, P, 1), flags=SB)
        du, dvp = ws["d2"], ws["d3"]
        for l in reversed(range(self.L)):
            if l < self.L - 1:
                self._gelu_backward(saved_u[l], dv, du)
                g = du
            else:
                g = dv
            self._gemm(g, self.ones, gr[f"b{l}"], 1, C, 1, P, (0, P, 1), (0, 0, 0), (0, 1, 0), beta=1.0, splits=sp, flags=SA)
            self._gemm(g, saved_v[l], gr[f"W{l}"], 1, C, C, P, (0, P, 1), (0, 1, P), (0, C, 1), beta=1.0, splits=sp, flags=SAB)
            self._gemm(pr[f"W{l}"], g, dvp, 1, C, P, C, (0, 1, C), (0, P, 1), (0, P, 1), flags=SB)
            self._spectral_backward(g, l, saved_v[l], dvp)
            dv, dvp = dvp, dv
        self._gemm(dv, a, gr["P"], 1, C, self.cin, P, (0, P, 1), (0, 1, P), (0, self.cin, 1), beta=1.0, splits=sp, flags=SAB)
        self._gemm(dv, self.ones, gr["pb"], 1, C, 1, P, (0, P, 1), (0, 0, 0), (0, 1, 0), beta=1.0, splits=sp, flags=SA)
        self._saved = None

    def forward_cost(self) -> Dict[str, float]:
        """Matrix-core flops (split operands counted twice, as issued) and minimal HBM bytes of
        one inference forward pass (activations FP32, read or written once per GEMM)."""
        C, n, m, P, L = self.C, self.n, self.m, self.P, self.L
        tm = 2 * m
        per_layer = {
            "z": (C * n * n, tm, n, 2), "y": (C * n * tm, m, 2 * n, 2), "x": (C * tm, m * m, 2 * n, 2),
            "modes": (m ** 3 * C, 1, C, 4 * 2), "ix": (C * 2 * n, m * m, tm, 2), "iy": (C * n * 2 * n, m, tm, 2),
            "iz": (C * n * n, n, tm, 2), "w": (C, P, C, 2),
        }
        flops = 0.0
        for M, N, K, products in per_layer.values():   # products: MFMA products per GEMM (split = 2)
            flops += 2.0 * M * N * K * products
        flops *= L
        flops += 2.0 * 2 * (C * P * self.cin + self.cout * P * C)
        act = 4.0 * C * P
        if self.fused:
            # per layer one pass reads v and writes GELU(u) (u at the last layer); the lift reads a
            # and writes v0; the projection reads v_L and writes out; truncated spectra negligible
            bytes_ = L * 2 * act + 4.0 * P * (self.cin + C) + 4.0 * P * (C + self.cout)
        else:
            # z-stage reads v; inverse z writes S; W-GEMM reads v and S and writes GELU(u) (u too
            # at the last layer); the small truncated spectra are negligible
            bytes_ = L * 5 * act + 4.0 * P * (self.cin + C) + 4.0 * P * (C + self.cout)
        return {"flops": flops, "hbm_bytes": bytes_}

    def zero_grad(self) -> None:
        for g in self.grads.values():
            g.zero_()

    def adam(self, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> None:
        self.adam_step += 1
        lib = _abi.lib()
        s = _abi.stream_handle()
        for k, w in self.params.items():
            _abi.check(lib.wfs_fno_adam(_abi.ptr(w), _abi.ptr(self.grads[k]), _abi.ptr(self.adam_m[k]),
                                        _abi.ptr(self.adam_v[k]), w.numel(), lr, beta1, beta2, eps, self.adam_step,
                                        s), "wfs_fno_adam")

    def loss_and_grad(self, out: torch.Tensor, target: torch.Tensor, dout: torch.Tensor | None,
                      reference: torch.Tensor | None = None):
        """Relative squared error sum (out - target)^2 / D and its gradient into dout.

        D = sum target^2, or sum (reference - target)^2 when a reference is given (for a
        residual network whose reference is its base input: the loss is then 1 for a zero
        correction and measures what the correction gains, with gradients of order one).
        """
        ws = self._workspace()
        lib = _abi.lib()
        s = _abi.stream_handle()
        mse, res = ws["mse"], ws["mse_out"]
        _abi.check(lib.wfs_fno_mse(_abi.ptr(out), _abi.ptr(target), None, 0.0, out.numel(), _abi.ptr(mse),
                                   mse.numel(), _abi.ptr(res), s), "wfs_fno_mse")
        err, tn = res.cpu().numpy()
        if reference is not None:
            _abi.check(lib.wfs_fno_mse(_abi.ptr(reference), _abi.ptr(target), None, 0.0, out.numel(), _abi.ptr(mse),
                                       mse.numel(), _abi.ptr(res), s), "wfs_fno_mse")
            tn = float(res.cpu().numpy()[0])
        if dout is not None:
            _abi.check(lib.wfs_fno_mse(_abi.ptr(out), _abi.ptr(target), _abi.ptr(dout), 1.0 / tn, out.numel(),
                                       _abi.ptr(mse), mse.numel(), _abi.ptr(res), s), "wfs_fno_mse")
        return float(err / tn)

    def state_dict(self) -> Dict[str, np.ndarray]:
        return {k: v.cpu().numpy() for k, v in self.params.items()}

    def load_state_dict(self, d: Dict[str, np.ndarray]) -> None:
        for k, v in d.items():
            self.params[k].copy_(torch.from_numpy(np.asarray(v)).to(self.params[k].dtype))
