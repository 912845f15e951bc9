"""Set-up of a run (sources/initialisation.py:16-124).

``initialize(use_cache)`` keeps the reference's signature: without a config it
picks the TOML file under ``config/`` (asking for an index when there are
several, as the reference does); a caller may pass the parsed ``config`` (and a
torch ``device``) instead.  It builds the SimState, the MeasurementTools and
the IterData: cache validity by comparing the config with the one cached last
time (``cached_parameters.toml``), state/potential initialisation, snapshot
resume (psi included), and a copy of the config in the output folder.
"""

from __future__ import annotations

import math
import os
import sys
import time
from typing import Dict, Optional

import tomli

from . import snapshot_io
from .iter_data import IterData
from .measurement import MeasurementTools
from .sim_state import SimState

_RED = "\033[31m"
_BLUE = "\033[34m"
_RESET = "\033[0m"

CONFIG_DIR = "config/"


def load_config(path: str) -> Dict:
    """TOML file -> dict; a decode error exits with status 1 (sources/initialisation.py:42-48)."""
    with open(path, mode="rb") as f:
        try:
            return tomli.load(f)
        except tomli.TOMLDecodeError:
            print(_RED + "Failed to decode TOML file. There are possibly errors in the file." + _RESET)
            sys.exit(1)


def _toml_value(v) -> str:
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, int):
        return str(v)
    if isinstance(v, float):
        if math.isnan(v):
            return "nan"
        if math.isinf(v):
            return "inf" if v > 0 else "-inf"
        return repr(v)
    if isinstance(v, str):
        return '"' + v.replace("\\", "\\\\").replace('"', '\\"') + '"'
    if isinstance(v, (list, tuple)):
        return "[" + ", ".join(_toml_value(x) for x in v) + "]"
    if isinstance(v, dict):
        return "{" + ", ".join(f"{_toml_key(k)} = {_toml_value(x)}" for k, x in v.items()) + "}"
    raise TypeError(f"cannot write {type(v).__name__} to TOML")


def _toml_key(k: str) -> str:
    k = str(k)
    if k and all(c.isalnum() or c in "_-" for c in k):
        return k
    return _toml_value(k)


def toml_dumps(config: Dict) -> str:
    """The config dict as TOML text (tables, arrays of tables, scalars and arrays): what
    tomli.loads reads back as an equal dict."""
    lines = []

    def table(d: Dict, path):
        simple = [(k, v) for k, v in d.items()
                  if not isinstance(v, dict) and not (isinstance(v, list) and v and all(isinstance(x, dict) for x in v))]
        for k, v in simple:
            lines.append(f"{_toml_key(k)} = {_toml_value(v)}")
        for k, v in d.items():
            sub = path + [_toml_key(k)]
            if isinstance(v, dict):
                lines.append("")
                lines.append("[" + ".".join(sub) + "]")
                table(v, sub)
            elif isinstance(v, list) and v and all(isinstance(x, dict) for x in v):
                for item in v:
                    lines.append("")
                    lines.append("[[" + ".".join(sub) + "]]")
                    table(item, sub)

    table(config, [])
    return "\n".join(lines).lstrip("\n") + "\n"


def select_config_file(config_dir: str = CONFIG_DIR) -> str:
    """The one TOML file under config_dir, or the one picked by index (sources/initialisation.py:16-40)."""
    if not os.path.exists(config_dir):
        print("No config file found under config folder.")
        print("Exiting application.")
        sys.exit(0)
    files = sorted(f for f in os.listdir(config_dir) if f.endswith(".toml"))
    if not files:
        print("No config file found under config folder.")
        print("Exiting application.")
        sys.exit(0)
    if len(files) == 1:
        return os.path.join(config_dir, files[0])
    answer = -1
    while answer not in range(len(files)):
        print("Available configuration files:")
        for idx, name in enumerate(files):
            print(f"{idx} {name}")
        print("Select one by entering its index number:", end=" ")
        try:
            answer = int(input())
        except ValueError:
            answer = -1
    return os.path.join(config_dir, files[answer])


def initialize(use_cache: bool = True, config: Optional[Dict] = None, device=None):
    """(SimState, MeasurementTools, IterData) of a run (sources/initialisation.py:16-124)."""
    t0 = time.time()
    if config is None:
        path = select_config_file()
        print(f"Opening {os.path.basename(path)}")
        config = load_config(path)
    sim_state = SimState(config, device=device)
    cache_dir = sim_state.get_cache_dir()
    os.makedirs(cache_dir, exist_ok=True)
    cached = os.path.join(cache_dir, "cached_parameters.toml")
    try:
        with open(cached, "rb") as f:
            if tomli.load(f) != tomli.loads(toml_dumps(config)):
                print("Configuration file is different from the one used last time.\n"
                      "Falling back to full initialisation.")
                use_cache = False
    except (OSError, tomli.TOMLDecodeError):
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
    text = toml_dumps(config)
    with open(cached, "w") as f:
        f.write(text)
    with open(os.path.join(sim_state.get_output_dir(), "config.toml"), "w") as f:
        f.write(text)
    print(f"Time spent with initialisation: {(time.time() - t0):.2f} s.\n")
    return sim_state, measurement_tools, iter_data
