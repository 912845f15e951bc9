"""Probability current through a plane (sources/measurement.py:190-302)."""

from __future__ import annotations

from typing import Dict

import numpy as np
import torch

from . import math_utils
from .config_read_helper import try_read_param


class PlaneProbabilityCurrent:
    """Flux of j = Im(conj(psi) grad psi)/m through a sampled rectangle.

    Keys and defaults follow sources/measurement.py:207-218 (the reference
    passes the table path as the *default* of center/normal/enable_image,
    :210-215; here those defaults are the origin, +x and True).  The sample
    counts are resolution + 1 with an even count reduced by one (:219-223).
    """

    def __init__(self, config: Dict, sim_state):
        where = "measurement.plane_probability_currents"
        self.__name = try_read_param(config, "name", "Probability current", where)
        self.__center_bohr_radii_3 = np.array(try_read_param(config, "center_bohr_radii_3", [0.0, 0.0, 0.0], where),
                                              dtype=np.float64)
        self.__normal_vector_3 = math_utils.normalize(
            np.array(try_read_param(config, "normal_vector_3", [1.0, 0.0, 0.0], where), dtype=np.float64))
        self.__enable_image = bool(try_read_param(config, "enable_image", True, where))
        self.__size_bohr_radii_2 = np.array(try_read_param(config, "size_bohr_radii_2", [60.0, 60.0], where),
                                            dtype=np.float64)
        c = np.array(try_read_param(config, "resolution_2", [512, 512], where), dtype=np.int64) + 1
        self.__sample_count_2 = np.where(c % 2 == 0, c - 1, c)
        self.__delta_t = sim_state.get_delta_time_h_bar_per_hartree()
        self.__sample_interval_steps = 1
        self.__evolution = np.empty(shape=0, dtype=np.float64)
        self.__pending = None
        self.__density = None

    def get_name(self):
        return self.__name

    def set_name(self, n: str):
        self.__name = n

    def is_enable_image(self):
        return self.__enable_image

    def set_sample_interval_steps(self, k: int):
        """Steps between two samples (1 = every iteration, the reference's cadence)."""
        self.__sample_interval_steps = max(1, int(k))

    def calculate(self, sim_state):
        plan = sim_state.get_plan()
        psi = sim_state.get_wave_function()
        if self.__density is None:
            dt = torch.float64 if sim_state.is_double_precision() else torch.float32
            self.__density = torch.empty(tuple(int(v) for v in self.__sample_count_2), dtype=dt, device=psi.device)
        self.__pending = plan.plane_current(psi, self.__center_bohr_radii_3, self.__normal_vector_3,
                                            self.__size_bohr_radii_2, self.__sample_count_2, density=self.__density)

    def get_probability_current_density(self):
        """Last j.n map, [count_u][count_v]."""
        return self.__density

    def synchronize(self):
        if self.__pending is None:
            return
        self.__evolution = np.append(self.__evolution, float(self.__pending[0].item()))
        self.__pending = None

    def get_probability_current_evolution_with_name(self):
        return self.__evolution, self.__name

    def get_integrated_probability_current_evolution_with_name(self):
        """Running Simpson integral of the flux (sources/measurement.py:294-297)."""
        return (math_utils.indefinite_simpson_integral(self.__evolution,
                                                       self.__delta_t * self.__sample_interval_steps),
                self.__name)
