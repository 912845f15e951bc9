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


@pytest.mark.parametrize("n,width,modes,layers", [(32, 8, 8, 2), (64, 16, 12, 3), (128, 32, 12, 4), (96, 32, 16, 2)])
def test_fused_layers_match_gemm_composition(dev, n, width, modes, layers):
    """The fused layer kernel (u = W v + Z Gz + b, GELU, next z stage per grid line) against the
    network's unfused GEMM path on the same weights: forward output, the saved activations of a
    training forward, and every gradient.  Both accumulate bf16 products in FP32 (in another
    order), so they agree to FP32 rounding, not bit for bit."""
    print(f"case: n={n} width={width} modes={modes} layers={layers}", flush=True)
    nets = [FNO3d(n, width=width, modes=modes, layers=layers, seed=3, splits=64, fused=f) for f in (True, False)]
    assert nets[0].fused and not nets[1].fused
    nets[0].params["Q"].mul_(10.0)
    nets[1].params["Q"].copy_(nets[0].params["Q"])
    P = n ** 3
    a = torch.from_numpy(np.random.default_rng(7).standard_normal(3 * P).astype(np.float32)).to(dev)
    outs = [net.forward(a).clone() for net in nets]
    e_inf = _rel(outs[0].cpu().numpy() - a[: 2 * P].cpu().numpy(), outs[1].cpu().numpy() - a[: 2 * P].cpu().numpy())
    target = a[: 2 * P] + 0.1
    grads = []
    for net in nets:
        out = net.forward(a, keep=True)
        saved = [t.clone() for t in net._saved[1] + net._saved[2]]
        dout = torch.empty_like(out)
        net.zero_grad()
        net.loss_and_grad(out, target, dout)
        net.backward(dout)
        grads.append(({k: g.clone() for k, g in net.grads.items()}, out.clone(), saved))
    torch.cuda.synchronize()
    e_keep = _rel(grads[0][1].cpu().numpy(), grads[1][1].cpu().numpy())
    e_saved = max(_rel(x.cpu().numpy(), y.cpu().numpy()) for x, y in zip(grads[0][2], grads[1][2]))
    g0 = {k: g.cpu().numpy() for k, g in grads[0][0].items()}
    g1 = {k: g.cpu().numpy() for k, g in grads[1][0].items()}

    def gnorm(k):
        # a spectral weight is one complex number: its two gradients share one scale
        if k[:2] in ("Rr", "Ri"):
            return np.hypot(np.linalg.norm(g1["Rr" + k[2:]]), np.linalg.norm(g1["Ri" + k[2:]]))
        return np.linalg.norm(g1[k])

    e_grad = max(np.linalg.norm(g0[k] - g1[k]) / gnorm(k) for k in g1 if gnorm(k) > 0)
    print(f"  fused vs GEMM path: forward (network term) {e_inf:.2e}, training forward {e_keep:.2e}, "
          f"saved activations {e_saved:.2e}, gradients {e_grad:.2e}")
    assert e_inf < 2e-5 and e_keep < 1e-6 and e_saved < 2e-5 and e_grad < 1e-3


def test_mode_mixing_kernel_matches_gemms(dev):
    """wfs_fno_modes against the four strided GEMMs it replaces (same bf16 operand rounding,
    FP32 accumulation in another order) and against float64 on the rounded operands."""
    C, m3 = 32, 1728
    rng = np.random.default_rng(8)
    Rr, Ri = (torch.from_numpy(rng.standard_normal(m3 * C * C).astype(np.float32) / C).to(dev) for _ in range(2))
    X = torch.from_numpy(rng.standard_normal(2 * C * m3).astype(np.float32)).to(dev)
    O = torch.empty_like(X)
    _abi.fno_modes(Rr, Ri, X, O, m3, C)
    O2 = torch.empty_like(X)
    Xr, Xi, Or, Oi = X, X[m3:], O2, O2[m3:]
    sR, sX = (C * C, 1, C), (1, 2 * m3, 0)
    _abi.gemm_bf16(Rr, Xr, Or, m3, C, 1, C, sR, sX, sX, flags=_abi.GEMM_SPLIT_B)
    _abi.gemm_bf16(Ri, Xi, Or, m3, C, 1, C, sR, sX, sX, alpha=-1.0, beta=1.0, flags=_abi.GEMM_SPLIT_B)
    _abi.gemm_bf16(Ri, Xr, Oi, m3, C, 1, C, sR, sX, sX, flags=_abi.GEMM_SPLIT_B)
    _abi.gemm_bf16(Rr, Xi, Oi, m3, C, 1, C, sR, sX, sX, beta=1.0, flags=_abi.GEMM_SPLIT_B)
    torch.cuda.synchronize()
    R = (_np_bf16(Rr.cpu().numpy()) + 1j * _np_bf16(Ri.cpu().numpy())).reshape(m3, C, C)
    xn = X.cpu().numpy().reshape(C, 2, m3)
    xs = [_np_bf16(p) + _np_bf16(p - _np_bf16(p)) for p in (xn[:, 0], xn[:, 1])]
    ref = np.einsum("qio,iq->oq", R, xs[0] + 1j * xs[1])
    got = O.cpu().numpy().reshape(C, 2, m3)
    gemm = O2.cpu().numpy().reshape(C, 2, m3)
    scale = np.abs(ref).max()
    e_ref = max(np.abs(got[:, 0] - ref.real).max(), np.abs(got[:, 1] - ref.imag).max()) / scale
    e_gemm = np.abs(got - gemm).max() / scale
    print(f"mode mixing: vs float64 on rounded operands {e_ref:.2e}, vs GEMM path {e_gemm:.2e}")
    assert e_ref < 1e-6 and e_gemm < 1e-6


def test_fused_layer_limits(dev):
    assert not FNO3d(16, width=8, modes=4, layers=1).fused          # n not a multiple of 32
    assert not FNO3d(32, width=40, modes=4, layers=1).fused         # width above 32
    assert not FNO3d(32, width=8, modes=18, layers=1).fused         # 2 modes above 32
    z = torch.zeros(64, device=dev)
    with pytest.raises(_abi.WfsError):
        rc = _abi.lib().wfs_fno_layer(_abi.ptr(z), _abi.ptr(z), None, _abi.ptr(z), _abi.ptr(z), _abi.ptr(z),
                                      None, _abi.ptr(z), None, 4, 4, 48, 8, 1, None)
        _abi.check(rc, "wfs_fno_layer")


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


def test_elementwise_kernels_match_torch(dev):
    """bias + GELU, GELU', axpby, relative MSE and Adam against float32/float64 torch."""
    lib, s = _abi.lib(), _abi.stream_handle()
    g = torch.Generator(device=dev).manual_seed(3)
    C, P = 5, 1000
    u = torch.randn(C * P, device=dev, generator=g)
    bias = torch.randn(C, device=dev, generator=g)
    v = torch.empty_like(u)
    ref_u = (u.view(C, P) + bias[:, None]).reshape(-1)
    _abi.check(lib.wfs_fno_bias_act(_abi.ptr(u), _abi.ptr(bias), _abi.ptr(v), C, P, 1, s), "bias_act")
    torch.cuda.synchronize()
    assert torch.allclose(u, ref_u, atol=0, rtol=0)
    assert torch.allclose(v, torch.nn.functional.gelu(ref_u, approximate="tanh"), atol=1e-6, rtol=1e-5)
    dv = torch.randn_like(u)
    du = torch.empty_like(u)
    _abi.check(lib.wfs_fno_gelu_backward(_abi.ptr(ref_u), _abi.ptr(dv), _abi.ptr(du), u.numel(), s), "gelu_bw")
    x = ref_u.double().requires_grad_(True)
    (gref,) = torch.autograd.grad(torch.nn.functional.gelu(x, approximate="tanh"), x, dv.double())
    torch.cuda.synchronize()
    assert torch.allclose(du.double(), gref, atol=1e-5, rtol=1e-5)
    y = torch.randn_like(u)
    y0 = y.clone()
    _abi.check(lib.wfs_fno_axpby(_abi.ptr(y), _abi.ptr(dv), 0.5, -2.0, u.numel(), s), "axpby")
    torch.cuda.synchronize()
    assert torch.allclose(y, 0.5 * dv - 2.0 * y0, atol=1e-6)
    pred, targ = torch.randn_like(u), torch.randn_like(u)
    dpred = torch.empty_like(u)
    work = torch.empty(8192, dtype=torch.float64, device=dev)
    out2 = torch.empty(2, dtype=torch.float64, device=dev)
    _abi.check(lib.wfs_fno_mse(_abi.ptr(pred), _abi.ptr(targ), _abi.ptr(dpred), 0.25, u.numel(), _abi.ptr(work),
                               8192, _abi.ptr(out2), s), "mse")
    torch.cuda.synchronize()
    d = pred.double() - targ.double()
    assert abs(float(out2[0]) - float((d * d).sum())) < 1e-9 * float((d * d).sum())
    assert abs(float(out2[1]) - float((targ.double() ** 2).sum())) < 1e-9 * float((targ.double() ** 2).sum())
    assert torch.allclose(dpred.double(), 2 * d * 0.25, atol=1e-6)
    w = torch.randn(100, device=dev, generator=g)
    grad = torch.randn_like(w)
    m, vv = torch.zeros_like(w), torch.zeros_like(w)
    opt_w = w.detach().clone().requires_grad_(True)
    opt = torch.optim.Adam([opt_w], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    for step in (1, 2):
        _abi.check(lib.wfs_fno_adam(_abi.ptr(w), _abi.ptr(grad), _abi.ptr(m), _abi.ptr(vv), w.numel(), 1e-2, 0.9,
                                    0.999, 1e-8, step, s), "adam")
        opt_w.grad = grad.clone()
        opt.step()
    torch.cuda.synchronize()
    assert torch.allclose(w, opt_w.detach(), atol=1e-6, rtol=1e-5)
