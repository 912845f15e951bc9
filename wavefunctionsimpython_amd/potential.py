"""Potential descriptions and their GPU rasterisation (sources/potential.py).

V is a complex64 device tensor on the grid; walls, the harmonic well and the
lattice add to its real part, the absorbing boundary to its imaginary part.
"""

from __future__ import annotations

from typing import Dict, List

import numpy as np
import torch

from . import _abi
from . import math_utils
from .config_read_helper import try_read_param

_RED = "\033[31m"
_RESET = "\033[0m"


class PreInitializedPotential:
    """Raw potential file (sources/potential.py:14-55)."""

    def __init__(self, config: Dict):
        self.__path = try_read_param(config, "potential.pre_initialized_potential.path", "")
        self.__enable = bool(try_read_param(config, "potential.pre_initialized_potential.enable", False))
        if len(self.__path) == 0:
            self.__enable = False
        self.__visible = try_read_param(config, "potential.pre_initialized_potential.visible", True)
        type_str = try_read_param(config, "potential.pre_initialized_potential.float_type", "float64")
        types = {"float64": np.float64, "float32": np.float32, "complex64": np.complex64, "complex128": np.complex128}
        if type_str in types:
            self.__float_type = types[type_str]
        else:
            self.__float_type = np.float64
            print(_RED + f"Unrecognised float type for the pre-initialized potential \"{self.__path}\"!\n"
                  "Available options are: float32, float64, complex64, complex128. Falling back to float64." + _RESET)

    def is_enabled(self):
        return self.__enable

    def is_visible(self):
        return self.__visible

    def add_potential(self, V: torch.Tensor) -> torch.Tensor:
        """V += file contents, real files widened to complex (sources/potential.py:47-55)."""
        arr = np.fromfile(self.__path, dtype=self.__float_type).reshape(tuple(V.shape))
        V += torch.from_numpy(arr.astype(np.complex64)).to(V.device)
        return V


class AbsorbingBoundaryCondition:
    """Complex absorbing potential outside the observation box (sources/potential.py:57-158)."""

    def __init__(self, config, bottom_corner, top_corner):
        self.boundary_bottom_corner_bohr_radii_3 = np.array(bottom_corner, dtype=np.float64)
        self.boundary_top_corner_bohr_radii_3 = np.array(top_corner, dtype=np.float64)
        self.outer_potential_in_positive_xyz_direction_hartree_3 = np.array(try_read_param(
            config, "absorbing_boundary_condition.outer_potential_in_positive_xyz_direction_hartree_3",
            [-100.0, -100.0, -100.0]), dtype=np.float64)
        self.outer_potential_in_negative_xyz_direction_hartree_3 = np.array(try_read_param(
            config, "absorbing_boundary_condition.outer_potential_in_negative_xyz_direction_hartree_3",
            [-100.0, -100.0, -100.0]), dtype=np.float64)
        self.start_offset = float(try_read_param(config, "absorbing_boundary_condition.start_offset", 1.0))
        self.slope_exponent_in_positive_xyz_direction_3 = np.array(try_read_param(
            config, "absorbing_boundary_condition.slope_exponent_in_positive_xyz_direction_3", [2.0, 2.0, 2.0]),
            dtype=np.float64)
        self.slope_exponent_in_negative_xyz_direction_3 = np.array(try_read_param(
            config, "absorbing_boundary_condition.slope_exponent_in_negative_xyz_direction_3", [2.0, 2.0, 2.0]),
            dtype=np.float64)
        self.enable = bool(try_read_param(config, "absorbing_boundary_condition.enable", True))

    def add_potential(self, V: torch.Tensor, delta_x_3, planes=None) -> torch.Tensor:
        """planes=(x_begin, x_end): only those x planes of V get the absorber."""
        if not self.enable:
            return V
        _abi.add_cap(tuple(V.shape), delta_x_3, self.boundary_bottom_corner_bohr_radii_3,
                     self.boundary_top_corner_bohr_radii_3, self.start_offset,
                     self.outer_potential_in_positive_xyz_direction_hartree_3,
                     self.outer_potential_in_negative_xyz_direction_hartree_3,
                     self.slope_exponent_in_positive_xyz_direction_3,
                     self.slope_exponent_in_negative_xyz_direction_3, V, planes=planes)
        return V


class PotentialWall:
    """A slab with optional slits (sources/potential.py:161-242)."""

    def __init__(self, wall_config: Dict):
        self.potential_hartree = float(try_read_param(wall_config, "potential_hartree", 20))
        self.center_bohr_radii_3 = np.array(try_read_param(wall_config, "center_bohr_radii_3", [0.0, 0.0, 0.0]),
                                            dtype=np.float64)
        self.normal_vector_3 = math_utils.normalize(
            np.array(try_read_param(wall_config, "normal_vector_3", [1.0, 0.0, 0.0]), dtype=np.float64))
        self.plateau_thickness_bohr_radii = float(try_read_param(wall_config, "plateau_thickness_bohr_radii", 5.0))
        self.slope_thickness_bohr_radii = float(try_read_param(wall_config, "slope_thickness_bohr_radii", 0.0))
        self.slope_exponent = float(try_read_param(wall_config, "slope_exponent", 1.0))
        self.velocity_bohr_radius_hartree_per_h_bar_3 = np.array(
            try_read_param(wall_config, "velocity_bohr_radius_hartree_per_h_bar_3", [0.0, 0.0, 0.0]),
            dtype=np.float64)
        self.angular_velocity_rad_hartree_per_h_bar_3 = np.array(
            try_read_param(wall_config, "angular_velocity_rad_hartree_per_h_bar_3", [0.0, 0.0, 0.0]),
            dtype=np.float64)
        self.potential_change_rate_hartree_sqr_per_h_bar = float(
            try_read_param(wall_config, "potential_change_rate_hartree_sqr_per_h_bar", 0.0))
        self.slit_count = int(try_read_param(wall_config, "slit_count", 0))
        if self.slit_count > 4:
            print(_RED + "Potential wall slit count is currently supported only up to 4 slits!" + _RESET)
            self.slit_count = 4
        elif self.slit_count < 0:
            print(_RED + "Negative potential wall slit count is invalid!" + _RESET)
            self.slit_count = 0
        self.slit_spacing_bohr_radii = float(try_read_param(wall_config, "slit_spacing_bohr_radii", 1.0))
        self.slit_width_bohr_radii = float(try_read_param(wall_config, "slit_width_bohr_radii", 0.5))
        self.slit_rotation_radian = float(try_read_param(wall_config, "slit_rotation_radian", 0.0))
        self.visible = try_read_param(wall_config, "visible", True)
        self.enable = bool(try_read_param(wall_config, "enable", True))

    def add_potential(self, V: torch.Tensor, delta_x_3, planes=None) -> torch.Tensor:
        """planes=(x_begin, x_end): only those x planes of V get the wall."""
        if not self.enable:
            return V
        _abi.add_wall(tuple(V.shape), delta_x_3, self.potential_hartree, self.center_bohr_radii_3,
                      self.normal_vector_3, self.plateau_thickness_bohr_radii, self.slope_thickness_bohr_radii,
                      self.slope_exponent, self.slit_count, self.slit_spacing_bohr_radii,
                      self.slit_width_bohr_radii, self.slit_rotation_radian, V, planes=planes)
        return V

    def x_planes(self, shape, delta_x_3, center=None):
        """The x planes [lo, hi) outside which the wall (at ``center``, default its own) adds
        nothing, or None when its normal is not along x (its slab then crosses every x plane).

        A voxel gets the wall where |n . (r - c)| <= plateau/2 + slope
        (sources/cuda_kernels/potential_wall.cu:37-38, :76-82); with n = (+-1, 0, 0) that is
        |r_x - c_x| <= plateau/2 + slope, r_x = dx (i - nx/2).  Two planes of margin on each side
        cover the kernel's float32 rounding of r and d."""
        nx = int(shape[0])
        if not self.enable:
            return (0, 0)
        n = self.normal_vector_3
        if n[1] != 0.0 or n[2] != 0.0 or n[0] == 0.0:
            return None
        c = float((self.center_bohr_radii_3 if center is None else center)[0])
        reach = (0.5 * self.plateau_thickness_bohr_radii + max(self.slope_thickness_bohr_radii, 0.0)) / abs(n[0])
        d = float(delta_x_3[0])
        lo = int(np.floor(nx / 2 + (c - reach) / d)) - 2
        hi = int(np.ceil(nx / 2 + (c + reach) / d)) + 3
        return (min(max(lo, 0), nx), min(max(hi, 0), nx))


class HarmonicPotential:
    """V += 1/2 m w^2 |r - c|^2 from the ``potential.harmonic`` table.

    The reference has this formula only as an uncalled helper
    (sources/potential.py:414-423); the table is an addition of this build.
    """

    def __init__(self, config: Dict, mass: float):
        self.enable = bool(try_read_param(config, "potential.harmonic.enable", False))
        self.angular_frequency = float(try_read_param(config, "potential.harmonic.angular_frequency", 0.0))
        self.center_bohr_radii_3 = np.array(
            try_read_param(config, "potential.harmonic.center_bohr_radii_3", [0.0, 0.0, 0.0]), dtype=np.float64)
        self.mass = float(mass)

    def add_potential(self, V: torch.Tensor, delta_x_3) -> torch.Tensor:
        if self.enable:
            _abi.add_harmonic(tuple(V.shape), delta_x_3, self.angular_frequency, self.mass, self.center_bohr_radii_3, V)
        return V


class OpticalLattice:
    """V += depth * sum_d cos^2(pi r_d / period) from ``potential.lattice`` (no reference counterpart)."""

    def __init__(self, config: Dict):
        self.enable = bool(try_read_param(config, "potential.lattice.enable", False))
        self.depth_hartree = float(try_read_param(config, "potential.lattice.depth_hartree", 0.0))
        self.period_bohr_radii = float(try_read_param(config, "potential.lattice.period_bohr_radii", 1.0))

    def add_potential(self, V: torch.Tensor, delta_x_3) -> torch.Tensor:
        if self.enable:
            _abi.add_lattice(tuple(V.shape), delta_x_3, self.depth_hartree, self.period_bohr_radii, V)
        return V


def generate_potential_from_walls_and_drain(V: torch.Tensor, V_vis: torch.Tensor, delta_x_3,
                                            absorbing_pontetial: AbsorbingBoundaryCondition,
                                            walls: List[PotentialWall], planes=None):
    """Dynamic-mode regeneration: V = 0, + walls, + absorbing boundary.

    sources/potential.py:508-522 (V_vis is returned unchanged there too).  planes=(x_begin,
    x_end) regenerates those x planes only, with the values the whole-grid regeneration gives
    them; the other planes keep theirs.
    """
    if planes is None:
        V.zero_()
    else:
        V[planes[0]:planes[1]].zero_()
    for w in walls:
        w.add_potential(V, delta_x_3, planes=planes)
    absorbing_pontetial.add_potential(V, delta_x_3, planes=planes)
    return V, V_vis
