"""Observables and their result files (sources/measurement.py).

All box integrals of one measurement go to the GPU in a single
``wfs_observables`` call (Simpson weights, FP64 accumulation) and come back in
one device-to-host copy; the driver can also hand a whole stretch of
iterations to one ``wfs_step_observe`` call and append its samples with
``record_samples`` (one copy per stretch).  Visual outputs of the reference (vispy volume
rendering, matplotlib plots, mp4) are presentation and are not part of this
package; their configuration keys are accepted and ignored.
"""

from __future__ import annotations

import itertools
import os
from typing import Dict, List

import numpy as np
import torch

from . import math_utils
from .config_read_helper import try_read_param
from .iter_data import IterData
from .sim_state import SimState

_RED = "\033[31m"
_RESET = "\033[0m"


def _check_box(name, bottom, count, N):
    for a in range(3):
        if int(bottom[a]) < 0 or int(count[a]) < 1 or int(bottom[a]) + int(count[a]) > int(N[a]):
            raise ValueError(
                f"{name}: Simpson box [{int(bottom[a])}, {int(bottom[a]) + int(count[a])}) on axis {a} reaches "
                f"outside the grid of {int(N[a])} voxels (the reference would read out of bounds here)")


class VolumeProbability:
    """Probability inside a box (sources/measurement.py:65-188)."""

    def __init__(self, config: Dict, sim_state: SimState):
        self.__name = try_read_param(config, "name", "Volume probability", "measurement.volume_probabilities")
        b = np.array(try_read_param(config, "bottom_corner_bohr_radii_3", [-10.0, -10.0, -10.0],
                                    "measurement.volume_probabilities"), dtype=np.float64)
        t = np.array(try_read_param(config, "top_corner_bohr_radii_3", [10.0, 10.0, 10.0],
                                    "measurement.volume_probabilities"), dtype=np.float64)
        for i in range(3):
            if b[i] > t[i]:
                b[i], t[i] = t[i], b[i]
        self.__bottom_corner_bohr_radii_3 = b
        self.__top_corner_bohr_radii_3 = t
        self.__enable_image = bool(try_read_param(config, "enable_image", True, "measurement.volume_probabilities"))
        # voxel box: bottom, top + 1, odd sample count (sources/measurement.py:101-116)
        self.__bottom_voxel = sim_state.transform_physical_coordinate_to_voxel_3(b)
        top_voxel = sim_state.transform_physical_coordinate_to_voxel_3(t) + np.uint64(1)
        self.__sample_count = (top_voxel - self.__bottom_voxel).astype(np.int64)
        for i in range(3):
            if self.__sample_count[i] % 2 == 0:
                print(_RED + f"Volume probability \"{self.__name}\": truncating volume probability area along "
                      f"{i}. axis!" + _RESET)
                self.__sample_count[i] -= 1
        _check_box(f"Volume probability \"{self.__name}\"", self.__bottom_voxel, self.__sample_count,
                   sim_state.get_number_of_voxels_3())
        self.__probability_evolution = np.empty(shape=0, dtype=np.float64)
        self.__pending = None

    def get_name(self):
        return self.__name

    def set_name(self, n: str):
        self.__name = n

    def is_enable_image(self):
        return self.__enable_image

    def get_box(self):
        """(bottom voxel, sample count) of the Simpson box."""
        return tuple(int(v) for v in self.__bottom_voxel), tuple(int(v) for v in self.__sample_count)

    def calculate(self, sim_state: SimState):
        """Launch the box integral on the current state (sources/measurement.py:153-175)."""
        out = sim_state.get_plan().observables(sim_state.get_wave_function(), [self.get_box()])
        self.__pending = out[0, 0]

    def set_pending(self, value):
        self.__pending = value

    def synchronize(self):
        """Append the last calculated value (sources/measurement.py:183-187)."""
        if self.__pending is None:
            return
        v = float(self.__pending.item() if isinstance(self.__pending, torch.Tensor) else self.__pending)
        self.__probability_evolution = np.append(self.__probability_evolution, v)
        self.__pending = None

    def get_probability_evolution_with_name(self):
        return self.__probability_evolution, self.__name

    def clear(self):
        self.__probability_evolution = np.empty(shape=0, dtype=np.float64)


class ExpectedLocation:
    """<r> and sigma_r over the observation box (sources/measurement.py:305-407)."""

    def __init__(self, config: Dict, sim_state: SimState):
        self.__enable_image = bool(try_read_param(config, "measurement.expected_location.enable_image", False))
        self.__bottom_voxel = sim_state.get_observation_box_bottom_corner_voxel_3()
        # The reference adds one to the SimState's own top-voxel array and takes
        # one back when the count is even (sources/measurement.py:320-329); the
        # observation-box view of SimState sees the same array.
        top_voxel = sim_state.get_observation_box_top_corner_voxel_3()
        top_voxel += np.array([1, 1, 1], dtype=top_voxel.dtype)
        sample_count = (top_voxel - self.__bottom_voxel).astype(np.int64)
        for i in range(3):
            if sample_count[i] % 2 == 0:
                print(_RED + f"Expected location: truncating integration area along {i}. axis!" + _RESET)
                sample_count[i] -= 1
                top_voxel[i] -= 1
        self.__sample_count = sample_count
        _check_box("Expected location", self.__bottom_voxel, self.__sample_count, sim_state.get_number_of_voxels_3())
        self.__expected_location_evolution = np.zeros(shape=(0, 3), dtype=np.float64)
        self.__standard_deviation_evolution = np.zeros(shape=(0, 3), dtype=np.float64)
        self.__probability_evolution = np.empty(shape=0, dtype=np.float64)
        self.__pending = None

    def is_enable_image(self):
        return self.__enable_image

    def get_box(self):
        return tuple(int(v) for v in self.__bottom_voxel), tuple(int(v) for v in self.__sample_count)

    def calculate(self, sim_state: SimState):
        out = sim_state.get_plan().observables(sim_state.get_wave_function(), [self.get_box()])
        self.__pending = out[0]

    def set_pending(self, moments):
        self.__pending = moments

    def synchronize(self):
        """Append E[r] and sqrt(E[r^2] - E[r]^2) (sources/measurement.py:386-400)."""
        if self.__pending is None:
            return
        m = self.__pending.cpu().numpy() if isinstance(self.__pending, torch.Tensor) else np.asarray(self.__pending)
        e_r = m[1:4].reshape(1, 3)
        e_r2 = m[4:7].reshape(1, 3)
        self.__expected_location_evolution = np.concatenate((self.__expected_location_evolution, e_r), axis=0)
        with np.errstate(invalid="ignore"):
            sd = np.sqrt(e_r2 - np.power(e_r, 2))
        self.__standard_deviation_evolution = np.concatenate((self.__standard_deviation_evolution, sd), axis=0)
        self.__probability_evolution = np.append(self.__probability_evolution, m[0])
        self.__pending = None

    def get_expected_location_evolution(self):
        return self.__expected_location_evolution

    def get_standard_deviation_evolution(self):
        return self.__standard_deviation_evolution

    def get_norm_evolution(self):
        """Simpson norm over the observation box at each sample (not in the reference)."""
        return self.__probability_evolution


class MeasurementTools:
    """All observables of a run (sources/measurement.py:482-843)."""

    def __init__(self, config: Dict, sim_state: SimState):
        for key in ("measurement.volumetric.enable_animation", "measurement.volumetric.enable_image",
                    "measurement.per_axis_plot.enable_animation", "measurement.per_axis_plot.enable_image"):
            if try_read_param(config, key, False):
                print(_RED + f"{key}: visual output is not part of this package; ignored." + _RESET)
        self.__volume_probabilities: List[VolumeProbability] = []
        for conf in try_read_param(config, "measurement.volume_probabilities", []):
            new_v = VolumeProbability(conf, sim_state)
            self._resolve_naming_conflicts(self.__volume_probabilities, new_v)
            self.__volume_probabilities.append(new_v)
        self.__plane_probability_currents = []
        for conf in try_read_param(config, "measurement.plane_probability_currents", []):
            from .probability_current import PlaneProbabilityCurrent
            new_p = PlaneProbabilityCurrent(conf, sim_state)
            self._resolve_naming_conflicts(self.__plane_probability_currents, new_p)
            self.__plane_probability_currents.append(new_p)
        self.__expected_location = None
        if try_read_param(config, "measurement.expected_location.enable_image", False):
            self.__expected_location = ExpectedLocation(config, sim_state)
        self.__out = None
        self.__observe_every = 1

    @staticmethod
    def _resolve_naming_conflicts(items, new_item):
        """A name an item of the same kind already has gets the first free suffix _1, _2, ...,
        with the reference's two warnings (rename rule of sources/measurement.py:502-518)."""
        taken = {item.get_name() for item in items}
        name = new_item.get_name()
        if name not in taken:
            return
        print(_RED + f"Found multiple measurement tools of the same type with equal name: \"{name}\"" + _RESET)
        suffix = next(k for k in itertools.count(1) if f"{name}_{k}" not in taken)
        new_item.set_name(f"{name}_{suffix}")
        print(_RED + f"Renamed one to \"{new_item.get_name()}\"." + _RESET)

    def set_observe_every(self, k: int):
        """Samples are k steps apart (time integrals of the series use k * dt)."""
        self.__observe_every = max(1, int(k))
        for pc in self.__plane_probability_currents:
            pc.set_sample_interval_steps(self.__observe_every)

    def get_volume_probabilities(self):
        return self.__volume_probabilities

    def get_plane_probability_currents(self):
        return self.__plane_probability_currents

    def get_expected_location(self):
        return self.__expected_location

    def write_wave_function_to_file(self, sim_state: SimState, iter_data: IterData):
        """Observation-box dumps as .npy and raw interleaved floats (sources/measurement.py:626-642)."""
        if iter_data.i % sim_state.get_wave_function_save_interval() != 0:
            return
        out = sim_state.get_output_dir()
        os.makedirs(os.path.join(out, "wave_function_npy"), exist_ok=True)
        os.makedirs(os.path.join(out, "wave_function_raw"), exist_ok=True)
        wf = np.ascontiguousarray(sim_state.get_view_into_wave_function().cpu().numpy())
        np.save(os.path.join(out, f"wave_function_npy/wave_function_{iter_data.i:06d}.npy"), wf)
        type_str = "float64" if sim_state.is_double_precision() else "float32"
        wf.view(dtype=type_str).astype(type_str).tofile(
            os.path.join(out, f"wave_function_raw/wave_function_{iter_data.i:06d}.bin"))

    def observation_boxes(self):
        """(boxes, owners): the Simpson boxes of every enabled volume probability and of the
        expected location, in the order their rows come back from the library."""
        boxes = []
        owners = []
        for v in self.__volume_probabilities:
            if v.is_enable_image():
                boxes.append(v.get_box())
                owners.append(v)
        if self.__expected_location is not None:
            boxes.append(self.__expected_location.get_box())
            owners.append(self.__expected_location)
        return boxes, owners

    def needs_state_per_sample(self) -> bool:
        """True when a sample needs psi itself (enabled plane currents), not only box moments."""
        return any(pc.is_enable_image() for pc in self.__plane_probability_currents)

    @staticmethod
    def _hand_out(owners, rows):
        for owner, row in zip(owners, rows):
            if isinstance(owner, VolumeProbability):
                owner.set_pending(row[0])
            else:
                owner.set_pending(row)

    def record_samples(self, samples: torch.Tensor, owners) -> None:
        """Append the samples of a wfs_step_observe call -- (n_samples, n_boxes, 7) box moments,
        sample j taken before step j * observe_every -- to every series, in the order
        measure_and_render would have appended them one by one (one device-to-host copy)."""
        host = samples.cpu().numpy()
        for rows in host:
            self._hand_out(owners, rows)
            for owner in owners:
                owner.synchronize()

    def measure_and_render(self, sim_state: SimState, iter_data: IterData):
        """Save, then calculate all observables, then synchronise (sources/measurement.py:644-704)."""
        if sim_state.is_wave_function_saving():
            self.write_wave_function_to_file(sim_state, iter_data)
        boxes, owners = self.observation_boxes()
        if boxes:
            plan = sim_state.get_plan()
            if self.__out is None or self.__out.shape[0] != len(boxes):
                self.__out = torch.empty((len(boxes), 7), dtype=torch.float64,
                                         device=sim_state.get_wave_function().device)
            plan.observables(sim_state.get_wave_function(), boxes, out=self.__out)
            self._hand_out(owners, self.__out.cpu().numpy())
        for plane in self.__plane_probability_currents:
            if plane.is_enable_image():
                plane.calculate(sim_state)
        for v in self.__volume_probabilities:
            if v.is_enable_image():
                v.synchronize()
        for plane in self.__plane_probability_currents:
            if plane.is_enable_image():
                plane.synchronize()
        if self.__expected_location is not None:
            self.__expected_location.synchronize()

    def finish(self, sim_state: SimState):
        """Write the observable time series (sources/measurement.py:707-843, files only)."""
        type_str = "float64" if sim_state.is_double_precision() else "float32"
        out = sim_state.get_output_dir()
        f_npy = os.path.join(out, "measurable_npy")
        f_bin = os.path.join(out, "measurable_raw")
        os.makedirs(f_npy, exist_ok=True)
        os.makedirs(f_bin, exist_ok=True)

        def dump(name, arr):
            np.save(os.path.join(f_npy, name + ".npy"), arr)
            np.asarray(arr).astype(type_str).tofile(os.path.join(f_bin, name + ".bin"))

        if self.__expected_location is not None:
            dump("expected_location_evolution", self.__expected_location.get_expected_location_evolution())
            dump("standard_deviation_evolution", self.__expected_location.get_standard_deviation_evolution())
        for v in self.__volume_probabilities:
            if v.is_enable_image():
                evo, name = v.get_probability_evolution_with_name()
                dump(f"volume_probability_evolution_{name}", evo)
                dwell = math_utils.indefinite_simpson_integral(
                    evo, sim_state.get_delta_time_h_bar_per_hartree() * self.__observe_every)
                if dwell.size:
                    print(f"Dwell time in {name}: {dwell[-1]:.5f} ħ/Hartree")
        for pc in self.__plane_probability_currents:
            if pc.is_enable_image():
                evo, name = pc.get_probability_current_evolution_with_name()
                dump(f"probability_current_evolution_{name}", evo)
                ievo, name = pc.get_integrated_probability_current_evolution_with_name()
                np.save(os.path.join(f_npy, f"integrated_probability_current_{name}.npy"), ievo)
                np.asarray(ievo).astype(type_str).tofile(os.path.join(f_bin, f"integrated_probability_current_{name}.bin"))
                if ievo.size:
                    print(f"Transfer probability on {name}: {ievo[-1]:.5f}")
