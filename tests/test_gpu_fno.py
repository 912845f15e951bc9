"""bf16 matrix-core GEMM and the Fourier neural operator on the GPU."""

import numpy as np
import pytest
import torch

from fno_emulation import EmulatedFNO, round_bf16
from oracle import fno_oracle
from wavefunctionsimpython_amd import _abi
from wavefunctionsimpython_amd.fno import FNO3d

pytestmark = pytest.mark.gpu


def _np_bf16(x):
    return round_bf16(torch.from_numpy(np.asarray(x, np.float32))).numpy()


GEMM_CASES = {
    # name: batch, M, N, K, A layout, B layout, C layout, splits, alpha, beta, flags
    "rowmajor_ragged": (1, 100, 70, 45, "mk", "kn", "mn", 1, 1.0, 0.0, 0),
    "tall_skinny_Atrans_batched": (3, 200, 20, 70, "km", "kn", "mn", 1, 1.0, 0.0, 0),
    "B_k_fast": (2, 64, 96, 130, "mk", "nk", "mn", 1, 1.0, 0.0, 0),
    "C_transposed": (1, 50, 40, 33, "mk", "kn", "nm", 1, -1.0, 0.0, 0),
    "splitk_alpha_beta": (1, 33, 40, 5000, "mk", "nk", "mn", 16, 0.5, 2.0, 0),
    "splitk_batched": (2, 16, 24, 3000, "mk", "nk", "mn", 8, 1.0, 1.0, 0),
    "split_A": (1, 100, 70, 45, "mk", "kn", "mn", 1, 1.0, 0.0, 1),
    "split_B_tall": (3, 200, 20, 70, "km", "kn", "mn", 1, 1.0, 0.0, 2),
    "split_both_splitk": (1, 33, 40, 5000, "mk", "nk", "mn", 16, 0.5, 2.0, 3),
    "vector_loads": (2, 200, 96, 256, "mk", "kn", "mn", 1, 1.0, 0.0, 3),
    "vector_loads_splitk": (1, 64, 128, 4096, "mk", "kn", "mn", 8, 1.0, 1.0, 0),
    "short_wide": (2, 20, 300, 40, "mk", "kn", "mn", 1, 1.0, 0.5, 2),
    "short_wide_vector": (1, 32, 1024, 32, "mk", "kn", "mn", 1, 1.0, 1.0, 2),
    "small_tile": (3, 24, 12, 256, "mk", "kn", "mn", 1, 1.0, 0.0, 2),
}


def _strides(layout, batch, rows, cols):
    # element strides (batch, row, col) of a packed [batch][...] operand
    if layout in ("mk", "kn", "mn"):
        return (rows * cols, cols, 1)
    return (rows * cols, 1, rows)


@pytest.mark.parametrize("name", list(GEMM_CASES))
def test_gemm_bf16_matches_numpy(dev, name):
    batch, M, N, K, la, lb, lc, splits, alpha, beta, flags = GEMM_CASES[name]
    rng = np.random.default_rng(abs(hash(name)) % 2 ** 32)
    sA, sB, sC = _strides(la, batch, M, K), _strides(lb, batch, K, N), _strides(lc, batch, M, N)
    A = rng.standard_normal(batch * M * K).astype(np.float32)
    B = rng.standard_normal(batch * K * N).astype(np.float32)
    C0 = rng.standard_normal(batch * M * N).astype(np.float32)
    view = np.lib.stride_tricks.as_strided
    Ah, Bh = _np_bf16(A), _np_bf16(B)
    Al, Bl = _np_bf16(A - Ah), _np_bf16(B - Bh)
    Av = view(Ah, (batch, M, K), [s * 8 for s in sA])
    Bv = view(Bh, (batch, K, N), [s * 8 for s in sB])
    Cv = view(C0.astype(np.float64), (batch, M, N), [s * 8 for s in sC])
    prod = np.matmul(Av, Bv)
    if flags & 1:   # A carried as hi + lo
        prod += np.matmul(view(Al, (batch, M, K), [s * 8 for s in sA]), Bv)
    if flags & 2:
        prod += np.matmul(Av, view(Bl, (batch, K, N), [s * 8 for s in sB]))
    ref = alpha * prod + beta * Cv
    scale = np.abs(alpha) * np.matmul(np.abs(Av), np.abs(Bv)) + np.abs(beta * Cv)
    # the split operands recover about 16 bits of the FP32 product
    exact = alpha * np.matmul(view(A.astype(np.float64), (batch, M, K), [s * 8 for s in sA]),
                              view(B.astype(np.float64), (batch, K, N), [s * 8 for s in sB])) + beta * Cv
    At, Bt, Ct = (torch.from_numpy(x).to(dev) for x in (A, B, C0))
    work = torch.empty(batch * splits * M * N, dtype=torch.float32, device=dev) if splits > 1 else None
    _abi.gemm_bf16(At, Bt, Ct, batch, M, N, K, sA, sB, sC, alpha, beta, splits, work, flags=flags)
    torch.cuda.synchronize()
    got = view(Ct.cpu().numpy().astype(np.float64), (batch, M, N), [s * 8 for s in sC])
    err = np.abs(got - ref) / scale
    err_exact = (np.abs(got - exact) / scale).max()
    print(f"{name}: max err / (|A||B|) = {err.max():.2e}; vs exact product {err_exact:.2e}")
    assert err.max() < 4e-6 * np.sqrt(K / splits) + 1e-6 * splits
    if flags == 3:
        assert err_exact < 1e-4
    elif flags:
        assert err_exact < 5e-3


def test_gemm_vector_paths_with_partial_quads(dev):
    """Row strides that are multiples of 4 with K and N that are not: the 16-byte load
    path must fall back to single elements in the last, partial quad."""
    rng = np.random.default_rng(7)
    M, N, K, lda, ldb = 70, 101, 70, 72, 104
    A = rng.standard_normal(M * lda).astype(np.float32)
    B = rng.standard_normal(K * ldb).astype(np.float32)
    C = np.zeros(M * N, np.float32)
    At, Bt, Ct = (torch.from_numpy(x).to(dev) for x in (A, B, C))
    _abi.gemm_bf16(At, Bt, Ct, 1, M, N, K, (0, lda, 1), (0, ldb, 1), (0, N, 1))
    torch.cuda.synchronize()
    Ab = _np_bf16(A).reshape(M, lda)[:, :K]
    Bb = _np_bf16(B).reshape(K, ldb)[:, :N]
    ref = Ab @ Bb
    scale = np.abs(Ab) @ np.abs(Bb)
    err = (np.abs(Ct.cpu().numpy().reshape(M, N) - ref) / scale).max()
    assert err < 4e-6 * np.sqrt(K)


def test_gemm_mode_mixing_pattern(dev):
    """Per-mode channel mixing: batch over modes, N = 1, operands interleaved in memory."""
    C, m3 = 8, 64
    rng = np.random.default_rng(4)
    R = rng.standard_normal((m3, C, C)).astype(np.float32)      # [q][i][o]
    X = rng.standard_normal((C, 2, m3)).astype(np.float32)      # [i][part][q]
    O = np.zeros((C, 2, m3), np.float32)
    Rt, Xt, Ot = (torch.from_numpy(x.reshape(-1)).to(dev) for x in (R, X, O))
    _abi.gemm_bf16(Rt, Xt, Ot, m3, C, 1, C, (C * C, 1, C), (1, 2 * m3, 0), (1, 2 * m3, 0))
    torch.cuda.synchronize()
    ref = np.einsum("qio,iq->oq", _np_bf16(R).reshape(m3, C, C), _np_bf16(X)[:, 0])
    got = Ot.cpu().numpy().reshape(C, 2, m3)
    assert np.abs(got[:, 0] - ref).max() < 1e-5 * np.abs(ref).max()
    assert np.all(got[:, 1] == 0)


def test_gemm_rejects_out_of_range_strides(dev):
    A = torch.zeros(10, device=dev)
    with pytest.raises(ValueError):
        _abi.gemm_bf16(A, A, A, 1, 4, 4, 4, (0, 4, 1), (0, 4, 1), (0, 4, 1))


FNO_CASES = [  # n, width, modes, layers
    (16, 8, 4, 3),
    (32, 8, 8, 2),
]


def _fno_pair(n, width, modes, layers, seed=5):
    gpu = FNO3d(n, width=width, modes=modes, layers=layers, seed=seed, splits=16)
    emu = EmulatedFNO(n, width=width, modes=modes, layers=layers, seed=seed, device=torch.device("cpu"),
                      splits=16)
    emu.bf16_operands = True
    gpu.params["Q"].mul_(10.0)    # so the network term is not hidden by the residual
    for k in gpu.params:
        emu.params[k].copy_(gpu.params[k].cpu().double())
    return gpu, emu


def _inputs(n, seed=11):
    rng = np.random.default_rng(seed)
    P = n ** 3
    a = rng.standard_normal((3, P))
    target = a[:2] + 0.3 * rng.standard_normal((2, P))
    return a, target


def _rel(x, y):
    return float(np.linalg.norm(np.asarray(x) - np.asarray(y)) / np.linalg.norm(np.asarray(y)))


@pytest.mark.parametrize("n,width,modes,layers", FNO_CASES)
def test_fno_forward_and_gradients(dev, n, width, modes, layers):
    gpu, emu = _fno_pair(n, width, modes, layers)
    a, target = _inputs(n)
    P = n ** 3
    params = {k: v.cpu().numpy().astype(np.float64) for k, v in gpu.params.items()}
    loss_ref, out_ref, grads_ref = fno_oracle.loss_and_grads(params, a, target, n, modes, layers)

    a32 = torch.from_numpy(a.astype(np.float32).reshape(-1)).to(dev)
    t32 = torch.from_numpy(target.astype(np.float32).reshape(-1)).to(dev)
    out = gpu.forward(a32, keep=True)
    dout = torch.empty_like(out)
    gpu.zero_grad()
    loss = gpu.loss_and_grad(out, t32, dout)
    out_np = out.cpu().numpy().reshape(2, P)
    gpu.backward(dout)
    torch.cuda.synchronize()

    a_e = torch.from_numpy(a.astype(np.float32).astype(np.float64).reshape(-1))
    out_e = emu.forward(a_e, keep=True).view(2, P).clone()
    tt = torch.from_numpy(target.astype(np.float32).astype(np.float64))
    emu.zero_grad()
    emu.backward((2.0 * (out_e - tt) / float((tt ** 2).sum())).reshape(-1).contiguous())

    net_ref = out_ref - a[:2]
    e_oracle = _rel(out_np - a[:2], net_ref)
    e_emu = _rel(out_np, out_e.numpy())
    print(f"forward: vs FP64 oracle {e_oracle:.2e}, vs bf16-operand emulation {e_emu:.2e}; "
          f"loss {loss:.6f} vs {loss_ref:.6f}")
    assert e_oracle < 2e-2
    assert e_emu < 1e-4
    assert abs(loss - loss_ref) < 3e-2 * loss_ref
    for k, g in grads_ref.items():
        got = gpu.grads[k].cpu().numpy().reshape(g.shape)
        ge = emu.grads[k].numpy().reshape(g.shape)
        # a spectral weight is one complex number: its real and imaginary gradients share one scale
        # (the imaginary part alone is ~1e3 smaller and dominated by cancellation)
        if k[:2] in ("Rr", "Ri"):
            other = ("Ri" if k[:2] == "Rr" else "Rr") + k[2:]
            norm = np.hypot(np.linalg.norm(g), np.linalg.norm(grads_ref[other]))
        else:
            norm = np.linalg.norm(g)
        eo, ee = np.linalg.norm(got - g) / norm, np.linalg.norm(got - ge) / norm
        print(f"  grad {k:4s}: vs FP64 autograd {eo:.2e}, vs emulation {ee:.2e}")
        assert eo < 3e-2, k
        assert ee < 2e-4, k
