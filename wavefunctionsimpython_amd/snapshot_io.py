"""Interrupt snapshot and resume (sources/snapshot_io.py:9-49).

The reference writes only the loop counters (its psi save is commented out,
:11-16) so a resume restarts from the initial state; here psi is saved too
(``wave_snapshot.npy``), which the reference's reader already looks for.
"""

from __future__ import annotations

import io
import os

import numpy as np

from .iter_data import IterData


def write_snapshot(sim_state, iter_data: IterData):
    print("\nCreating snapshot. Please wait for the application to exit by itself!")
    os.makedirs(sim_state.get_cache_dir(), exist_ok=True)
    try:
        np.save(os.path.join(sim_state.get_cache_dir(), "wave_snapshot.npy"),
                sim_state.get_wave_function().cpu().numpy())
    except OSError:
        print("\nFailed writing snapshot of the wave function.")
    with io.open(os.path.join(sim_state.get_cache_dir(), "data_snapshot.txt"), mode="w") as f:
        f.write(f"{iter_data.i};{iter_data.elapsed_system_time_s};{iter_data.average_iteration_system_time_s};")


def read_snapshot(sim_state, iter_data: IterData):
    path = os.path.join(sim_state.get_cache_dir(), "wave_snapshot.npy")
    if os.path.exists(path):
        try:
            sim_state.set_wave_function(np.load(path, allow_pickle=False))
        except (OSError, ValueError):
            print("No usable previous wave tensor found!")
    try:
        with io.open(os.path.join(sim_state.get_cache_dir(), "data_snapshot.txt"), mode="r") as f:
            parts = f.read().split(";")
            iter_data.i = int(parts[0])
            iter_data.elapsed_system_time_s = float(parts[1])
            iter_data.average_iteration_system_time_s = float(parts[2])
    except (IOError, ValueError, IndexError):
        print("Failed to read " + os.path.join(sim_state.get_cache_dir(), "data_snapshot.txt"))
    return sim_state, iter_data


def remove_snapshot(sim_state):
    for name in ("data_snapshot.txt", "wave_snapshot.npy"):
        p = os.path.join(sim_state.get_cache_dir(), name)
        if os.path.exists(p):
            os.remove(p)
