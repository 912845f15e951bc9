"""The two-stage factorisation of the y transform used by the slab scheme (CPU, numpy).

The kernels split a length-NY transform, NY = LY * B, y = B a + b, ky = ka + LY kb, into
  stage 1 (with the z work): Y[b][ka] = sum_a x[B a + b] W_LY^{a ka}, stored at row B ka + b
  stage 2 (with the x pass): X[ka + LY kb] = sum_b W_B^{b kb} W_NY^{b ka} Y[b][ka]
and run the inverse as the same stages backwards, unnormalised; for B = 4 the length-4 stage
is done by lane pairs (each lane two rows, one value swapped), for B = 2 by one swap.  This
checks the index algebra of those stages, written as the kernels order them, against
numpy's FFT for every (NY, LY, B) the kernels use.
"""

import numpy as np
import pytest

SHAPES = [(256, 64, 4), (192, 48, 4), (128, 64, 2)]


def _w(n, e):
    return np.exp(-2j * np.pi * np.asarray(e) / n)


def _stage1(x, ny, ly, b, sign):
    out = np.empty(ny, complex)
    for r in range(b):
        rows = b * np.arange(ly) + r
        out[rows] = np.fft.fft(x[rows]) if sign < 0 else np.fft.ifft(x[rows]) * ly
    return out


def _stage2_forward(st, ny, ly, b):
    X = np.empty(ny, complex)
    for ka in range(ly):
        Z = st[b * ka + np.arange(b)] * _w(ny, np.arange(b) * ka)
        if b == 2:
            # lane q keeps X[q]: X[0] = Z0 + Z1 (lane 0), X[1] = Z0 - Z1 (lane 1: recv - mine)
            out = {0: Z[0] + Z[1], 1: Z[0] - Z[1]}
        else:
            out = {}
            for q in (0, 1):
                E = {qq: (Z[qq] + Z[qq + 2], Z[qq] - Z[qq + 2]) for qq in (0, 1)}
                P, Q = E[0][q], E[1][q]
                Qt = Q * (-1j) ** q
                out[q], out[q + 2] = P + Qt, P - Qt
        for kb in range(b):
            X[ka + ly * kb] = out[kb]
    return X


def _stage2_inverse(X, ny, ly, b):
    st = np.empty(ny, complex)
    for ka in range(ly):
        Xk = X[ka + ly * np.arange(b)]
        if b == 2:
            o = {0: Xk[0] + Xk[1], 1: Xk[0] - Xk[1]}
        else:
            o = {}
            for q in (0, 1):
                F = {qq: (Xk[qq] + Xk[qq + 2], Xk[qq] - Xk[qq + 2]) for qq in (0, 1)}
                P, Q = F[0][q], F[1][q]
                Qt = Q * (1j) ** q
                o[q], o[q + 2] = P + Qt, P - Qt
        for r in range(b):
            st[b * ka + r] = o[r] * np.conj(_w(ny, r * ka))
    return st


@pytest.mark.parametrize("ny,ly,b", SHAPES)
def test_forward_stages_equal_the_dft(ny, ly, b):
    rng = np.random.default_rng(ny)
    x = rng.standard_normal(ny) + 1j * rng.standard_normal(ny)
    X = _stage2_forward(_stage1(x, ny, ly, b, -1), ny, ly, b)
    assert np.abs(X - np.fft.fft(x)).max() < 1e-12 * np.abs(x).sum()


@pytest.mark.parametrize("ny,ly,b", SHAPES)
def test_inverse_stages_equal_the_unnormalised_inverse_dft(ny, ly, b):
    rng = np.random.default_rng(ny + 1)
    X = rng.standard_normal(ny) + 1j * rng.standard_normal(ny)
    x = _stage1(_stage2_inverse(X, ny, ly, b), ny, ly, b, +1)
    assert np.abs(x - np.fft.ifft(X) * ny).max() < 1e-12 * np.abs(X).sum()


@pytest.mark.parametrize("ny,ly,b", SHAPES)
def test_round_trip_through_the_half_transformed_rows(ny, ly, b):
    """entry stage 1, then stage 2 and its inverse (the x launch), then the inverse stage 1:
    ny times the input, as the unnormalised transforms with 1/ny on the kinetic table give"""
    rng = np.random.default_rng(ny + 2)
    x = rng.standard_normal(ny) + 1j * rng.standard_normal(ny)
    st = _stage1(x, ny, ly, b, -1)
    back = _stage1(_stage2_inverse(_stage2_forward(st, ny, ly, b), ny, ly, b), ny, ly, b, +1)
    assert np.abs(back / ny - x).max() < 1e-12 * np.abs(x).max() * ny
