"""Pins of the FP64 FNO oracle against what the mathematics fixes."""

import numpy as np
import torch

from oracle import fno_oracle


def test_mode_indices_signed_order():
    assert list(fno_oracle.mode_indices(16, 4)) == [0, 1, 14, 15]
    assert list(fno_oracle.mode_indices(6, 6)) == [0, 1, 2, 3, 4, 5]


def test_spectral_all_modes_identity_weights_is_identity():
    n, C = 6, 3
    rng = np.random.default_rng(0)
    v = torch.from_numpy(rng.standard_normal((C, n ** 3)))
    Rr = torch.zeros(n ** 3, C, C, dtype=torch.float64)
    Rr[:, range(C), range(C)] = 1.0
    Ri = torch.zeros_like(Rr)
    out = fno_oracle.spectral(v, Rr, Ri, n, n)
    assert torch.allclose(out, v, atol=1e-13)


def test_spectral_zero_frequency_only_gives_channel_mean():
    n, C, m = 8, 2, 4
    rng = np.random.default_rng(1)
    v = torch.from_numpy(rng.standard_normal((C, n ** 3)))
    Rr = torch.zeros(m ** 3, C, C, dtype=torch.float64)
    Ri = torch.zeros_like(Rr)
    Rr[0, 0, 1] = 2.0      # mode (0,0,0): channel 0 -> channel 1, weight 2
    Ri[0, 1, 0] = 5.0      # imaginary weight at f=0 drops out of the real part
    out = fno_oracle.spectral(v, Rr, Ri, n, m)
    assert torch.allclose(out[1], 2.0 * v[0].mean() * torch.ones(n ** 3, dtype=torch.float64), atol=1e-13)
    assert torch.allclose(out[0], torch.zeros(n ** 3, dtype=torch.float64), atol=1e-13)


def test_spectral_single_mode_is_plane_wave():
    n, m = 8, 4
    x = np.arange(n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    th = 2 * np.pi * (1 * X - 2 * Y + 0 * Z) / n      # frequency (1, -2, 0)
    v = torch.from_numpy(np.cos(th).reshape(1, -1))
    Rr = torch.zeros(m ** 3, 1, 1, dtype=torch.float64)
    Ri = torch.zeros_like(Rr)
    q = (1 * m + 2) * m + 0                             # kept index of (1, -2, 0): f=-2 sits at slot 2
    Rr[q, 0, 0] = 1.0
    out = fno_oracle.spectral(v, Rr, Ri, n, m).numpy().reshape(-1)
    # only the (1,-2,0) half of the cosine survives: Re(e^{i th} / 2)
    assert np.allclose(out, 0.5 * np.cos(th).reshape(-1), atol=1e-13)


def test_gelu_matches_torch_tanh_approximation():
    u = torch.linspace(-6, 6, 101, dtype=torch.float64)
    assert torch.allclose(fno_oracle.gelu_tanh(u), torch.nn.functional.gelu(u, approximate="tanh"), atol=1e-15)


def test_forward_gradients_finite_differences():
    n, m, L, C, cin = 4, 2, 2, 3, 3
    rng = np.random.default_rng(2)
    shapes = {"P": (C, cin), "pb": (C,), "Q": (2, C), "qb": (2,)}
    for l in range(L):
        shapes.update({f"W{l}": (C, C), f"b{l}": (C,), f"Rr{l}": (m ** 3, C, C), f"Ri{l}": (m ** 3, C, C)})
    names = list(shapes)
    params = [torch.tensor(0.5 * rng.standard_normal(shapes[k]), requires_grad=True) for k in names]
    a = torch.from_numpy(rng.standard_normal((cin, n ** 3)))
    target = torch.from_numpy(rng.standard_normal((2, n ** 3)))

    def f(*ps):
        return fno_oracle.relative_loss(fno_oracle.forward(dict(zip(names, ps)), a, n, m, L), target)

    assert torch.autograd.gradcheck(f, params, eps=1e-6, atol=1e-7, rtol=1e-5)
