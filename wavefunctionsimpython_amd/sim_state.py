"""Simulation state and the time step (sources/sim_state.py).

``SimState`` keeps the reference's configuration surface, geometry rules and
getters.  The state psi (complex64 or complex128) and the potential V
(complex64) are device tensors; every step runs in the HIP library through one
``wfs_step`` call, which merges consecutive split-step half kicks when several
steps are requested at once.
"""

from __future__ import annotations

import io
import os
from enum import Enum
from typing import Dict

import numpy as np
import torch

from . import _abi
from . import math_utils
from . import potential
from . import wave_packet
from .config_read_helper import try_read_param

_RED = "\033[31m"
_GREEN = "\033[32m"
_RESET = "\033[0m"


class SimulationMethod(Enum):
    """sources/sim_state.py:16-18."""
    FOURIER = 0
    POWER_SERIES = 1


class SimState:
    """Grid, time step, observation box, psi and V (sources/sim_state.py:21-762)."""

    def __init__(self, config: Dict, device: torch.device | None = None):
        self.__device = device or torch.device("cuda", torch.cuda.current_device() if torch.cuda.is_available() else 0)
        self.__cache_dir = try_read_param(config, "paths.cache_dir", "./cache")
        self.__output_dir = try_read_param(config, "paths.output_dir", "./output")
        self.__use_cache = True
        self.__enable_visual_output = True

        # method (sources/sim_state.py:70-77): fft/fourier -> FOURIER, fdtd/power_series -> POWER_SERIES,
        # anything else -> FOURIER
        method_name = str(try_read_param(config, "simulation.method", "fourier")).lower()
        if method_name in ("fdtd", "power_series"):
            self.__simulation_method = SimulationMethod.POWER_SERIES
        else:
            self.__simulation_method = SimulationMethod.FOURIER
        self.__order_of_approximation = int(try_read_param(config, "simulation.order_of_approximation", 10))
        self.__is_dynamic_potential_mode = bool(try_read_param(config, "simulation.is_dynamic_potential_mode", True))
        self.__enable_wave_function_saving = bool(
            try_read_param(config, "simulation.enable_wave_function_saving", False))
        self.__wave_function_save_iteration_interval = int(
            try_read_param(config, "simulation.wave_function_save_iteration_interval", 1))
        self.__double_precision = bool(try_read_param(config, "simulation.double_precision_calculation", False))

        self.__wave_packet = wave_packet.GaussianWavePacket(config)

        self.__simulated_volume_dimensions_bohr_radii_3 = np.array(
            try_read_param(config, "volume.simulated_volume_dimensions_bohr_radii_3", [100.0, 100.0, 100.0]),
            dtype=np.float64)
        self.__number_of_voxels_3 = np.array(
            try_read_param(config, "volume.number_of_voxels_3", [256, 256, 256]), dtype=np.int64)
        self.__delta_x_bohr_radii_3 = self.__simulated_volume_dimensions_bohr_radii_3 / self.__number_of_voxels_3
        dxm = float(np.max(self.__delta_x_bohr_radii_3))
        self.__upper_limit_on_delta_time_h_per_hartree = 4.0 / np.pi * (3.0 * dxm * dxm) / 3.0   # :95-100
        # required key: raises KeyError like sources/sim_state.py:101-103
        self.__delta_time_h_bar_per_hartree = float(config["simulation"]["delta_time_h_bar_per_hartree"])

        # observation box (sources/sim_state.py:105-161)
        L = self.__simulated_volume_dimensions_bohr_radii_3
        N = self.__number_of_voxels_3
        b = np.array(try_read_param(config, "volume.observation_box_bottom_corner_bohr_radii_3",
                                    [-25.0, -25.0, -25.0]), dtype=np.float64)
        t = np.array(try_read_param(config, "volume.observation_box_top_corner_bohr_radii_3",
                                    [25.0, 25.0, 25.0]), dtype=np.float64)
        for i in range(3):
            if b[i] > t[i]:
                b[i], t[i] = t[i], b[i]
        for i in range(3):
            if b[i] < -L[i] * 0.5:
                b[i] = -L[i] * 0.5
            if t[i] > L[i] * 0.5:
                t[i] = L[i] * 0.5
        self.__observation_box_bottom_corner_bohr_radii_3 = b
        self.__observation_box_top_corner_bohr_radii_3 = t
        self.__observation_box_bottom_corner_voxel_3 = np.array(
            math_utils.transform_center_origin_to_corner_origin_system(b, L) / self.__delta_x_bohr_radii_3,
            dtype=np.int32)
        self.__observation_box_top_corner_voxel_3 = np.array(
            math_utils.transform_center_origin_to_corner_origin_system(t, L) / self.__delta_x_bohr_radii_3,
            dtype=np.int32)
        for i in range(3):
            if self.__observation_box_top_corner_voxel_3[i] >= N[i]:
                self.__observation_box_top_corner_voxel_3[i] = N[i] - 1
            if self.__observation_box_bottom_corner_voxel_3[i] < 0:
                self.__observation_box_bottom_corner_voxel_3[i] = 0
        self.__observation_box_dimensions_bohr_radii_3 = t - b
        self.__observation_box_dimensions_voxel_3 = (self.__observation_box_top_corner_voxel_3
                                                     - self.__observation_box_bottom_corner_voxel_3)

        # potentials (sources/sim_state.py:167-183); lists are per instance
        self.__pre_initialized_potential = potential.PreInitializedPotential(config)
        if self.__pre_initialized_potential.is_enabled():
            self.__is_dynamic_potential_mode = False
        self.__absorbing_boundary_condition = potential.AbsorbingBoundaryCondition(
            config, bottom_corner=self.__observation_box_bottom_corner_bohr_radii_3,
            top_corner=self.__observation_box_top_corner_bohr_radii_3)
        self.__potential_walls = [potential.PotentialWall(w) for w in try_read_param(config, "potential.walls", [])]
        self.__harmonic = potential.HarmonicPotential(config, self.__wave_packet.get_particle_mass_electron_rest_mass())
        self.__lattice = potential.OpticalLattice(config)

        self.__wave_tensor: torch.Tensor | None = None
        self.__localised_potential_hartree: torch.Tensor | None = None
        self.__localised_potential_to_visualize_hartree: torch.Tensor | None = None
        self.__plan: _abi.Plan | None = None
        self.__potential_dirty = True   # the plan's phase table must be refreshed from V
        self.__dirty_planes = None      # else: only these x planes of it, (lo, hi)

    # ------------------------------------------------------------------ set-up
    def _shape(self):
        return tuple(int(v) for v in self.__number_of_voxels_3)

    def _build_potential(self):
        shape = self._shape()
        V = torch.zeros(shape, dtype=torch.complex64, device=self.__device)
        V_vis = torch.zeros(shape, dtype=torch.complex64, device=self.__device)
        if self.__pre_initialized_potential.is_enabled():
            self.__pre_initialized_potential.add_potential(V)
        self.__absorbing_boundary_condition.add_potential(V, self.__delta_x_bohr_radii_3)
        for wall in self.__potential_walls:
            wall.add_potential(V, self.__delta_x_bohr_radii_3)
            if wall.visible:
                wall.add_potential(V_vis, self.__delta_x_bohr_radii_3)
        self.__harmonic.add_potential(V, self.__delta_x_bohr_radii_3)
        self.__lattice.add_potential(V, self.__delta_x_bohr_radii_3)
        return V, V_vis

    def initialize_state(self):
        """Initial psi and V, plan creation, potential dump (sources/sim_state.py:185-300)."""
        print(self.get_sim_state_description_text(use_colors=True))
        shape = self._shape()
        full_init = True
        if self.__use_cache:
            try:
                arr = np.load(os.path.join(self.__cache_dir, "gaussian_wave_packet.npy"), allow_pickle=False)
                want = np.complex128 if self.__double_precision else np.complex64
                if arr.shape == shape and arr.dtype == want:
                    self.__wave_tensor = torch.from_numpy(arr).to(self.__device)
                    full_init = False
            except OSError:
                print("No cached gaussian_wave_packet.npy found.")
        if full_init:
            self.__wave_tensor = self.__wave_packet.init_wave_packet(
                self.__delta_x_bohr_radii_3, shape, self.__double_precision, self.__device)
            os.makedirs(self.__cache_dir, exist_ok=True)
            np.save(os.path.join(self.__cache_dir, "gaussian_wave_packet.npy"), self.__wave_tensor.cpu().numpy())

        full_init = True
        if self.__use_cache:
            try:
                v = np.load(os.path.join(self.__cache_dir, "localized_potential.npy"), allow_pickle=False)
                vv = np.load(os.path.join(self.__cache_dir, "localized_potential_to_visualize.npy"), allow_pickle=False)
                if v.shape == shape and v.dtype == np.complex64:
                    self.__localised_potential_hartree = torch.from_numpy(v).to(self.__device)
                    self.__localised_potential_to_visualize_hartree = torch.from_numpy(vv).to(self.__device)
                    full_init = False
            except OSError:
                print("No cached localized_potential.npy found.")
        if full_init:
            V, V_vis = self._build_potential()
            self.__localised_potential_hartree = V
            self.__localised_potential_to_visualize_hartree = V_vis
            os.makedirs(self.__cache_dir, exist_ok=True)
            np.save(os.path.join(self.__cache_dir, "localized_potential.npy"), V.cpu().numpy())
            np.save(os.path.join(self.__cache_dir, "localized_potential_to_visualize.npy"), V_vis.cpu().numpy())

        self.__plan = _abi.Plan(
            shape, self.__delta_x_bohr_radii_3, self.__delta_time_h_bar_per_hartree,
            self.get_particle_mass(), _abi.WFS_C128 if self.__double_precision else _abi.WFS_C64,
            _abi.WFS_METHOD_FOURIER if self.__simulation_method == SimulationMethod.FOURIER
            else _abi.WFS_METHOD_POWER_SERIES, self.__order_of_approximation)

        self.__potential_dirty = True
        # raw float32 dump of V (sources/sim_state.py:296-300)
        os.makedirs(os.path.join(self.__output_dir, "potential_raw"), exist_ok=True)
        self.__localised_potential_hartree.cpu().numpy().view(np.float32).tofile(
            os.path.join(self.__output_dir, "potential_raw/potential.bin"))

    # ------------------------------------------------------------------ getters
    def get_delta_time_h_bar_per_hartree(self):
        return self.__delta_time_h_bar_per_hartree

    def get_delta_x_bohr_radii_3(self):
        return self.__delta_x_bohr_radii_3

    def set_wave_function(self, wave_func):
        want = torch.complex128 if self.__double_precision else torch.complex64
        if not isinstance(wave_func, torch.Tensor):
            wave_func = torch.from_numpy(np.ascontiguousarray(wave_func))
        wave_func = wave_func.to(device=self.__device, dtype=want).contiguous()
        if tuple(wave_func.shape) != self._shape():
            raise ValueError(f"wave function shape {tuple(wave_func.shape)} != grid {self._shape()}")
        self.__wave_tensor = wave_func

    def get_wave_function(self) -> torch.Tensor:
        return self.__wave_tensor

    def get_potential(self) -> torch.Tensor:
        """Device complex64 potential used by the step."""
        return self.__localised_potential_hartree

    def set_potential(self, V: torch.Tensor):
        V = V.to(device=self.__device, dtype=torch.complex64).contiguous()
        if tuple(V.shape) != self._shape():
            raise ValueError("potential shape does not match the grid")
        self.__localised_potential_hartree = V
        self.__potential_dirty = True

    def get_plan(self) -> _abi.Plan:
        return self.__plan

    def is_wave_function_saving(self):
        return self.__enable_wave_function_saving

    def get_wave_function_save_interval(self):
        return self.__wave_function_save_iteration_interval

    def get_simulation_method(self):
        return self.__simulation_method

    def get_order_of_approximation(self):
        return self.__order_of_approximation

    def is_dynamic_potential_mode(self):
        return self.__is_dynamic_potential_mode

    def is_use_cache(self):
        return self.__use_cache

    def set_use_cache(self, uc: bool):
        self.__use_cache = uc

    def get_cache_dir(self):
        return self.__cache_dir

    def get_output_dir(self):
        return self.__output_dir

    def get_number_of_voxels_3(self):
        return self.__number_of_voxels_3

    def get_simulated_volume_dimensions_bohr_radii_3(self):
        return self.__simulated_volume_dimensions_bohr_radii_3

    def get_observation_box_bottom_corner_voxel_3(self):
        return self.__observation_box_bottom_corner_voxel_3

    def get_observation_box_top_corner_voxel_3(self):
        return self.__observation_box_top_corner_voxel_3

    def get_observation_box_bottom_corner_bohr_radii_3(self):
        return self.__observation_box_bottom_corner_bohr_radii_3

    def get_observation_box_top_corner_bohr_radii_3(self):
        return self.__observation_box_top_corner_bohr_radii_3

    def is_double_precision(self) -> bool:
        return self.__double_precision

    def get_device(self) -> torch.device:
        return self.__device

    def transform_physical_coordinate_to_voxel_3(self, pos_bohr_radii_3):
        """((pos + L/2) / dx) truncated to unsigned (sources/sim_state.py:354-358)."""
        return (math_utils.transform_center_origin_to_corner_origin_system(
            pos_bohr_radii_3, self.__simulated_volume_dimensions_bohr_radii_3)
            / self.__delta_x_bohr_radii_3).astype(np.uint64)

    def transform_voxel_to_physical_coordinate_3(self, voxel_3):
        """sources/sim_state.py:360-364."""
        return math_utils.transform_corner_origin_to_center_origin_system(
            np.asarray(voxel_3), self.__number_of_voxels_3) * self.__delta_x_bohr_radii_3

    def get_view_into_wave_function(self):
        """psi restricted to the observation box (sources/sim_state.py:366-383)."""
        return math_utils.cut_bounding_box(self.__wave_tensor, self.__observation_box_bottom_corner_voxel_3,
                                           self.__observation_box_top_corner_voxel_3)

    def get_view_into_potential(self):
        return math_utils.cut_bounding_box(self.__localised_potential_to_visualize_hartree.real,
                                           self.__observation_box_bottom_corner_voxel_3,
                                           self.__observation_box_top_corner_voxel_3)

    def get_view_into_complex_potential(self):
        return math_utils.cut_bounding_box(self.__localised_potential_to_visualize_hartree,
                                           self.__observation_box_bottom_corner_voxel_3,
                                           self.__observation_box_top_corner_voxel_3)

    def get_particle_mass(self):
        return self.__wave_packet.get_particle_mass_electron_rest_mass()

    def get_wave_packet(self):
        return self.__wave_packet

    def get_dxdydz(self):
        d = self.__delta_x_bohr_radii_3
        return d[0] * d[1] * d[2]

    # ------------------------------------------------------------------ dynamics
    def update_potential(self):
        """Dynamic-potential mode (sources/sim_state.py:466-509).

        Wall centres advect by v*dt and an x centre below zero is clamped to
        zero (the reference's ":472-473" test clamp); V is then rebuilt from
        walls and the absorbing boundary.  When no centre changed the rebuild
        would reproduce V bit for bit, so it is skipped.
        """
        if not self.__is_dynamic_potential_mode:
            return
        changed = False
        moved = []      # (wall, centre before the move)
        for w in self.__potential_walls:
            old = w.center_bohr_radii_3.copy()
            w.center_bohr_radii_3 = w.center_bohr_radii_3 + (w.velocity_bohr_radius_hartree_per_h_bar_3
                                                             * self.__delta_time_h_bar_per_hartree)
            if w.center_bohr_radii_3[0] < 0.0:
                w.center_bohr_radii_3[0] = 0.0
            if not np.array_equal(old, w.center_bohr_radii_3):
                changed = True
                moved.append((w, old))
        if not changed and self.__localised_potential_hartree is not None and getattr(self, "_regenerated", False):
            return
        planes = self._swept_planes(moved)
        self.__localised_potential_hartree, self.__localised_potential_to_visualize_hartree = \
            potential.generate_potential_from_walls_and_drain(
                V=self.__localised_potential_hartree, V_vis=self.__localised_potential_to_visualize_hartree,
                delta_x_3=self.__delta_x_bohr_radii_3, absorbing_pontetial=self.__absorbing_boundary_condition,
                walls=self.__potential_walls, planes=planes)
        if planes is None:
            self.__harmonic.add_potential(self.__localised_potential_hartree, self.__delta_x_bohr_radii_3)
            self.__lattice.add_potential(self.__localised_potential_hartree, self.__delta_x_bohr_radii_3)
            self.__potential_dirty = True
        elif not self.__potential_dirty and planes[0] < planes[1]:
            lo, hi = planes
            if self.__dirty_planes is not None:
                lo, hi = min(lo, self.__dirty_planes[0]), max(hi, self.__dirty_planes[1])
            self.__dirty_planes = (lo, hi)
        self._regenerated = True

    # Plane-wise regeneration (no reference counterpart: the reference rebuilds V whole): when
    # every wall that moved has its normal along x, only the x planes its slab covered before or
    # covers after the move change, and only those are rebuilt -- V and the plan's table come out
    # bit for bit as from the whole rebuild.  Used after a first whole rebuild, without the
    # harmonic well or the lattice (whole-grid terms), and while the planes are at most half of x.
    partial_regeneration = True

    def _swept_planes(self, moved):
        if not (self.partial_regeneration and getattr(self, "_regenerated", False)
                and self.__localised_potential_hartree is not None and moved):
            return None
        if self.__harmonic.enable or self.__lattice.enable:
            return None
        shape = tuple(self.__localised_potential_hartree.shape)
        lo, hi = shape[0], 0
        for w, old in moved:
            for c in (old, w.center_bohr_radii_3):
                span = w.x_planes(shape, self.__delta_x_bohr_radii_3, center=c)
                if span is None:
                    return None
                if span[0] < span[1]:
                    lo, hi = min(lo, span[0]), max(hi, span[1])
        if hi <= lo:
            return (0, 0)
        if 2 * (hi - lo) > shape[0]:
            return None
        return (lo, hi)

    def potential_is_static(self) -> bool:
        """True when update_potential() can no longer change V."""
        if not self.__is_dynamic_potential_mode:
            return True
        if not getattr(self, "_regenerated", False):
            return False
        for w in self.__potential_walls:
            if np.any(w.velocity_bohr_radius_hartree_per_h_bar_3 != 0.0):
                return False
        return True

    def potential_changes_every_step(self) -> bool:
        """True once V has been rebuilt in dynamic mode and a wall still moves: every further
        update_potential() changes V (an x centre clamped at 0 keeps moving in y and z, and an
        unchanged centre skips the rebuild)."""
        if not self.__is_dynamic_potential_mode or not getattr(self, "_regenerated", False):
            return False
        return any(np.any(w.velocity_bohr_radius_hartree_per_h_bar_3 != 0.0) for w in self.__potential_walls)

    def _refresh_potential(self):
        if self.__plan is None:
            raise RuntimeError("initialize_state() must be called before evolve_state()")
        if self.__potential_dirty:
            self.__plan.set_potential(self.__localised_potential_hartree)
            self.__potential_dirty = False
            self.__dirty_planes = None
        elif self.__dirty_planes is not None:
            lo, hi = self.__dirty_planes
            self.__plan.set_potential_planes(self.__localised_potential_hartree, lo, hi)
            self.__dirty_planes = None

    def evolve_state(self, steps: int = 1):
        """Advance psi by ``steps`` time steps in place (sources/sim_state.py:754-762)."""
        self._refresh_potential()
        self.__plan.step(self.__wave_tensor, None, int(steps))

    def evolve_state_observed(self, steps: int, observe_every: int, boxes, out: torch.Tensor | None = None,
                              update_between: bool = False):
        """``steps`` time steps with the box moments of the state before steps 0, observe_every, ...
        (< steps) -- the measure-then-evolve order of core_sim, sources/core_sim.py:19-24 -- in one
        wfs_step_observe call.  Returns the device float64 samples (n_samples, n_boxes, 7).

        update_between: update_potential() and the plan's table refresh run after every step
        inside the call (wfs_step_observe_between), as time_step's evolve-then-update order
        does between calls -- the dynamic mode with moving walls at the fused rate."""
        self._refresh_potential()
        between = self._update_between_steps if update_between else None
        return self.__plan.step_observe(self.__wave_tensor, None, int(steps), int(observe_every), boxes, out=out,
                                        between=between)

    # whether core_sim.run_iteration steps a potential that changes every step in stretches of
    # steps with the update between them (evolve_state_observed(update_between=True)) instead of
    # one evolve_state call per step
    dynamic_stretches = True

    def _update_between_steps(self, step: int):
        self.update_potential()
        self._refresh_potential()

    # ------------------------------------------------------------------ text
    def get_sim_state_description_text(self, use_colors=False):
        """Summary of the configured system (sources/sim_state.py:548-630, abridged)."""
        g = _GREEN if use_colors else ""
        r = _RESET if use_colors else ""
        red = _RED if use_colors else ""
        text = io.StringIO()
        text.write(g + "Simulated system state:\n" + r)
        text.write(f"Mass of the particle is {self.get_particle_mass()} electron rest mass.\n")
        text.write(f"Number of samples per axis is {tuple(int(v) for v in self.__number_of_voxels_3)}.\n")
        text.write(f"Space resolution is delta_r = {tuple(float(v) for v in self.__delta_x_bohr_radii_3)} Bohr radii.\n")
        text.write(f"The maximal viable time resolution < {self.__upper_limit_on_delta_time_h_per_hartree} h-bar / Hartree\n")
        text.write(f"Time resolution is delta = {self.__delta_time_h_bar_per_hartree} h-bar / Hartree.\n")
        if self.__delta_time_h_bar_per_hartree > 0.5 * self.__upper_limit_on_delta_time_h_per_hartree:
            text.write(red + "WARNING: delta_time exceeds theoretical limit!\n" + r)
        return text.getvalue()
