"""MI355X-native propagation path of WaveFunctionSimPython.

Drop-in modules mirror the reference's ``sources`` package: ``sim_state``,
``core_sim``, ``measurement``, ``operators``, ``wave_packet``, ``potential``,
``iter_data``, ``config_read_helper``, ``math_utils``.  All arithmetic runs in
the hand-written gfx950 HIP library ``libwfsim.so`` through the ctypes
binding in ``_abi``; PyTorch only owns device memory and streams.
"""

__version__ = "0.1.0"
