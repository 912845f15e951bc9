"""Initial wave packet (sources/wave_packet.py:39-151)."""

from __future__ import annotations

from typing import Dict

import numpy as np
import torch

from . import _abi
from . import math_utils
from .config_read_helper import try_read_param


class WavePacket:
    """Particle parameters shared by packet types (sources/wave_packet.py:39-61)."""

    def __init__(self):
        self._initial_position_bohr_radii_3 = np.array([0.0, 0.0, 0.0])
        self._initial_wp_velocity_bohr_radii_hartree_per_h_bar_3 = np.array([0.0, 0.0, 0.0])
        self._initial_momentum_h_per_bohr_radii_3 = np.array([0.0, 0.0, 0.0])
        self._particle_mass_electron_rest_mass = 1.0
        self._de_broglie_wave_length_bohr_radii = float("inf")

    def get_initial_wp_position_bohr_radii_3(self):
        return self._initial_position_bohr_radii_3

    def get_initial_wp_velocity_bohr_radii_hartree_per_h_bar_3(self):
        return self._initial_wp_velocity_bohr_radii_hartree_per_h_bar_3

    def get_initial_wp_momentum_h_per_bohr_radii_3(self):
        return self._initial_momentum_h_per_bohr_radii_3

    def get_particle_mass_electron_rest_mass(self):
        return self._particle_mass_electron_rest_mass

    def get_de_broglie_wave_length_bohr_radii(self):
        return self._de_broglie_wave_length_bohr_radii


class GaussianWavePacket(WavePacket):
    """Gaussian packet read from the ``wave_packet`` table (sources/wave_packet.py:63-151)."""

    def __init__(self, config: Dict):
        super().__init__()
        self.__path = try_read_param(config, "wave_packet.path", "")
        self._particle_mass_electron_rest_mass = float(
            try_read_param(config, "wave_packet.particle_mass_electron_rest_mass", 1.0))
        self._initial_position_bohr_radii_3 = np.array(
            try_read_param(config, "wave_packet.initial_position_bohr_radii_3", [0.0, 0.0, 0.0]), dtype=np.float64)
        self._initial_wp_velocity_bohr_radii_hartree_per_h_bar_3 = np.array(
            try_read_param(config, "wave_packet.initial_velocity_bohr_radii_hartree_per_h_bar_3", [0.0, 0.0, 0.0]),
            dtype=np.float64)
        self._initial_momentum_h_per_bohr_radii_3 = math_utils.classical_momentum(
            self._particle_mass_electron_rest_mass, self._initial_wp_velocity_bohr_radii_hartree_per_h_bar_3)
        p = float(np.dot(self._initial_momentum_h_per_bohr_radii_3, self._initial_momentum_h_per_bohr_radii_3)) ** 0.5
        self._de_broglie_wave_length_bohr_radii = math_utils.get_de_broglie_wave_length_bohr_radii(p)
        self.__initial_standard_deviation_bohr_radii_3 = np.array(
            try_read_param(config, "wave_packet.initial_standard_deviation_bohr_radii_3", [5.0, 5.0, 5.0]),
            dtype=np.float64)

    def get_initial_standard_deviation_bohr_radii(self):
        return self.__initial_standard_deviation_bohr_radii_3

    def init_wave_packet(self, delta_x_bohr_radii_3, shape, double_precision: bool = False,
                         device: torch.device | None = None) -> torch.Tensor:
        """Device tensor of the initial state (sources/wave_packet.py:97-151).

        With ``wave_packet.path`` set, the raw complex64/complex128 file is
        loaded and reshaped to the grid (:103-105); otherwise the Gaussian is
        generated on the GPU by ``wfs_init_gaussian``.
        """
        device = device or torch.device("cuda", torch.cuda.current_device())
        shape = tuple(int(s) for s in shape)
        if len(self.__path) > 0:
            arr = np.fromfile(self.__path, dtype=np.complex128 if double_precision else np.complex64).reshape(shape)
            return torch.from_numpy(arr).to(device)
        prec = _abi.WFS_C128 if double_precision else _abi.WFS_C64
        psi = torch.empty(shape, dtype=torch.complex128 if double_precision else torch.complex64, device=device)
        _abi.init_gaussian(shape, delta_x_bohr_radii_3, prec, self.__initial_standard_deviation_bohr_radii_3,
                           self._initial_position_bohr_radii_3, self._initial_momentum_h_per_bohr_radii_3, psi)
        return psi
