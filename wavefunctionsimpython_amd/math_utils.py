"""Coordinate conventions, analytic predictions and time integrals.

Host-side helpers of the path (sources/math_utils.py).  The grid convention is
r_d(i) = (i - N_d/2) * dx_d with integer N_d/2, so the grid spans
[-L/2, L/2 - dx].
"""

from __future__ import annotations

import math

import numpy as np

planck_constant = 2.0 * math.pi
reduced_planck_constant = 1.0
electron_rest_mass = 1.0


def transform_center_origin_to_corner_origin_system(pos, box_dimensions):
    """pos + L/2 -- sources/math_utils.py:41-42."""
    return np.asarray(pos) + np.asarray(box_dimensions) * 0.5


def transform_corner_origin_to_center_origin_system(pos, box_dimensions):
    """pos - L/2 -- sources/math_utils.py:45-46."""
    return np.asarray(pos) - np.asarray(box_dimensions) * 0.5


def classical_momentum(mass: float, velocity):
    """p = m v -- sources/math_utils.py:67-68."""
    return mass * np.asarray(velocity)


def get_de_broglie_wave_length_bohr_radii(momentum_h_bar_per_bohr_radius: float):
    """h / p (inf for p = 0) -- sources/math_utils.py:61-64."""
    if momentum_h_bar_per_bohr_radius == 0.0:
        return math.inf
    return planck_constant / momentum_h_bar_per_bohr_radius


def cut_bounding_box(arr, bottom, top):
    """arr[b0:t0, b1:t1, b2:t2] -- sources/math_utils.py:87-88."""
    return arr[int(bottom[0]):int(top[0]), int(bottom[1]):int(top[1]), int(bottom[2]):int(top[2])]


def normalize(v):
    v = np.asarray(v, dtype=np.float64)
    return v / math.sqrt(float(np.dot(v, v)))


def nearest_power_of_2(n: int) -> int:
    """Smallest power of two >= n -- sources/math_utils.py:134-146."""
    p = 1
    while p < n:
        p *= 2
    return p


def indefinite_simpson_integral(array, dt: float):
    """Running Simpson integral of a uniformly sampled series.

    sources/math_utils.py:199-234: entry 0 is 0, entry 1 the trapezoid of the
    first panel, then the odd and even chains each advance by Simpson panels.
    """
    a = np.asarray(array, dtype=np.float64)
    n = a.size
    if n == 0:
        return np.empty(0, dtype=np.float64)
    if n == 1:
        return np.zeros(1)
    if n == 2:
        return np.array([0.0, (a[0] + a[1]) / 2.0])
    panel = a[:-2] + 4.0 * a[1:-1] + a[2:]          # panel ending at index i+2
    out = np.zeros(n, dtype=np.float64)
    out[2::2] = np.cumsum(panel[0::2])                # even chain from 0
    out[1] = 1.5 * (a[0] + a[1])                      # trapezoid, pre-scaled by 3
    out[3::2] = out[1] + np.cumsum(panel[1::2])       # odd chain from 1
    return out * float(dt) / 3.0


def predict_free_space_standard_devation(delta_t: float, sigma0: float, mass: float, step_count: int):
    """sigma(t) = sqrt(a^2 + 4 t^2/(m^2 a^2)) / 2, a = 2 sigma0 -- sources/math_utils.py:237-251."""
    a = 2.0 * float(sigma0)
    t = np.arange(step_count, dtype=np.float64) * float(delta_t)
    return np.sqrt(a * a + 4.0 * t * t / (float(mass) ** 2 * a * a)) / 2.0


def predict_free_space_expected_location(delta_t: float, x0: float, velocity: float, step_count: int):
    """<x>(t) = x0 + v t -- sources/math_utils.py:254-267."""
    return float(x0) + np.arange(step_count, dtype=np.float64) * float(delta_t) * float(velocity)
