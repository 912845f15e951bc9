"""Driver loop (sources/core_sim.py:19-65) with SIGINT snapshot handling.

``time_step`` keeps the reference order -- measure the current state, advance
one step, update the potential.  ``run_iteration`` additionally honours
``simulation.observe_every`` (an extension key, default 1 = reference
behaviour).  Whenever the potential cannot change and no sample needs psi
itself (no wave-function files, no plane currents), it hands a stretch of up
to ``STRETCH`` iterations to one ``wfs_step_observe`` call: the box moments of
every sample come back from the GPU in one copy per stretch, and the split-step
fuses each sample's exit, moments and re-entry into one launch.  The samples,
the final state and the series are those of the iteration-by-iteration loop.
"""

from __future__ import annotations

import signal
import sys
import time

from . import snapshot_io
from .iter_data import IterData
from .measurement import MeasurementTools
from .sim_state import SimState

# iterations handed to one wfs_step_observe call (SIGINT is answered between stretches)
STRETCH = 1000

_GREEN = "\033[32m"
_RED = "\033[31m"
_RESET = "\033[0m"


def register_signal_handler(iter_data: IterData):
    """SIGINT sets iter_data.is_quit (sources/signal_handling.py:8-14)."""
    def handler(sig, frame):
        if not iter_data.is_quit:
            print("\n" + _RED + "Simulation interrupted by the user." + _RESET)
        iter_data.is_quit = True
    return signal.signal(signal.SIGINT, handler)


def time_step(sim_state: SimState, measurement_tools: MeasurementTools, iter_data: IterData):
    """measure -> evolve -> update potential (sources/core_sim.py:19-24)."""
    measurement_tools.measure_and_render(sim_state, iter_data)
    sim_state.evolve_state()
    sim_state.update_potential()


def run_iteration(sim_state: SimState, measurement_tools: MeasurementTools, iter_data: IterData,
                  observe_every: int = 1, exit_on_quit: bool = True):
    """Main loop from iter_data.i to total_iteration_count (sources/core_sim.py:26-65).

    Sample j of every observable is the state after j * observe_every steps;
    the final state is not sampled (reference behaviour).  On SIGINT the
    snapshot (now including psi) and the result files are written and the
    process exits (or, with exit_on_quit=False, the loop returns).
    """
    print(_GREEN + "Simulating " + _RESET + "(Press <Ctrl-c> to quit.)")
    iter_data.is_quit = False
    previous = register_signal_handler(iter_data)
    observe_every = max(1, int(observe_every))
    measurement_tools.set_observe_every(observe_every)
    try:
        i = iter_data.i
        total = iter_data.total_iteration_count
        while i < total:
            t0 = time.time()
            iter_data.i = i
            stretch = (i % observe_every == 0 and not sim_state.is_wave_function_saving()
                       and not measurement_tools.needs_state_per_sample())
            static = sim_state.potential_is_static()
            dynamic = (stretch and not static and sim_state.dynamic_stretches
                       and sim_state.potential_changes_every_step()
                       and bool(measurement_tools.observation_boxes()[0]))
            if stretch and static:
                # a stretch of whole sampling intervals in one call
                span = min(max(STRETCH - STRETCH % observe_every, observe_every), total - i)
                boxes, owners = measurement_tools.observation_boxes()
                if boxes:
                    measurement_tools.record_samples(sim_state.evolve_state_observed(span, observe_every, boxes),
                                                     owners)
                else:
                    sim_state.evolve_state(span)
            elif dynamic:
                # a potential that changes every step: the stretch runs update_potential after
                # every step inside the call (time_step's evolve-then-update order)
                span = min(max(STRETCH - STRETCH % observe_every, observe_every), total - i)
                boxes, owners = measurement_tools.observation_boxes()
                measurement_tools.record_samples(
                    sim_state.evolve_state_observed(span, observe_every, boxes, update_between=True), owners)
            else:
                if i % observe_every == 0:
                    measurement_tools.measure_and_render(sim_state, iter_data)
                # steps until the next measured iteration
                span = min(observe_every - i % observe_every, total - i)
                if not (span > 1 and sim_state.potential_is_static() and not sim_state.is_wave_function_saving()):
                    span = 1
                sim_state.evolve_state(span)
            if not dynamic:
                sim_state.update_potential()   # (a dynamic stretch updated after every step)
            i += span
            iter_data.elapsed_system_time_s += time.time() - t0
            if iter_data.is_quit:
                iter_data.i = i
                snapshot_io.write_snapshot(sim_state, iter_data)
                measurement_tools.finish(sim_state)
                if exit_on_quit:
                    sys.exit(0)
                return sim_state, measurement_tools, iter_data
        iter_data.i = total
    finally:
        # a handler installed outside Python reads back as None: restore the default then
        signal.signal(signal.SIGINT, previous if previous is not None else signal.SIG_DFL)
    iter_data.total_simulated_time = sim_state.get_delta_time_h_bar_per_hartree() * iter_data.total_iteration_count
    iter_data.average_iteration_system_time_s = (iter_data.elapsed_system_time_s
                                                 / float(max(1, iter_data.total_iteration_count)))
    return sim_state, measurement_tools, iter_data
