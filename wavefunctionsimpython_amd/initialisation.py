"""Non-interactive set-up of a run (sources/initialisation.py:16-124).

``load_config`` parses a TOML file (tomli); ``initialize`` builds the
SimState, the MeasurementTools and the IterData exactly as the reference does
after its interactive file pick: cache validity by comparing the config with
the one cached last time, state/potential initialisation, snapshot resume
(psi included), and a copy of the config in the cache and output folders
(JSON, since tomli only reads TOML).
"""

from __future__ import annotations

import json
import os
import sys
import time
from typing import Dict

import tomli

from . import snapshot_io
from .iter_data import IterData
from .measurement import MeasurementTools
from .sim_state import SimState

_RED = "\033[31m"
_BLUE = "\033[34m"
_RESET = "\033[0m"


def load_config(path: str) -> Dict:
    """TOML file -> dict; a decode error exits with status 1 (sources/initialisation.py:42-48)."""
    with open(path, mode="rb") as f:
        try:
            return tomli.load(f)
        except tomli.TOMLDecodeError:
            print(_RED + "Failed to decode TOML file. There are possibly errors in the file." + _RESET)
            sys.exit(1)


def initialize(config: Dict, use_cache: bool = True, device=None):
    """(SimState, MeasurementTools, IterData) for ``config``."""
    t0 = time.time()
    sim_state = SimState(config, device=device)
    cache_dir = sim_state.get_cache_dir()
    os.makedirs(cache_dir, exist_ok=True)
    cached = os.path.join(cache_dir, "cached_parameters.json")
    try:
        with open(cached) as f:
            if json.load(f) != json.loads(json.dumps(config)):
                print("Configuration file is different from the one used last time.\n"
                      "Falling back to full initialisation.")
                use_cache = False
    except (OSError, ValueError):
        use_cache = False
    sim_state.set_use_cache(use_cache)
    os.makedirs(sim_state.get_output_dir(), exist_ok=True)
    sim_state.initialize_state()
    measurement_tools = MeasurementTools(config, sim_state)
    iter_data = IterData(config)
    if (sim_state.is_use_cache()
            and os.path.exists(os.path.join(cache_dir, "data_snapshot.txt"))
            and os.path.exists(os.path.join(cache_dir, "wave_snapshot.npy"))):
        sim_state, iter_data = snapshot_io.read_snapshot(sim_state, iter_data)
        snapshot_io.remove_snapshot(sim_state)
        print(_BLUE + "Snapshot of an interrupted simulation loaded.\nResuming previous wave function.\n" + _RESET)
    with open(cached, "w") as f:
        json.dump(config, f)
    with open(os.path.join(sim_state.get_output_dir(), "config.json"), "w") as f:
        json.dump(config, f)
    print(f"Time spent with initialisation: {(time.time() - t0):.2f} s.\n")
    return sim_state, measurement_tools, iter_data
