"""Precomputed propagator tables (sources/operators.py:9-55).

The reference's two builders reference undefined names and are never called;
these implement their evident intent on the GPU: the kinetic half-step table
P_kin = exp(-i |k|^2 dt / (4m)) (the phase of kinetic_operator.cu:20-22) and
the potential table P_pot = cexp_i(-dt V) (potential_operator.cu:10-15).
"""

from __future__ import annotations

import numpy as np
import torch

from . import _abi


def init_kinetic_operator(delta_x_3, delta_time: float, mass: float, shape, double_precision: bool = False,
                          device: torch.device | None = None) -> torch.Tensor:
    """Device table exp(-i |k|^2 dt / (4 m)), k = 2 pi fftfreq(N, dx) per axis."""
    device = device or torch.device("cuda", torch.cuda.current_device())
    shape = tuple(int(s) for s in shape)
    prec = _abi.WFS_C128 if double_precision else _abi.WFS_C64
    out = torch.empty(shape, dtype=torch.complex128 if double_precision else torch.complex64, device=device)
    _abi.kinetic_operator(shape, np.asarray(delta_x_3, dtype=np.float64), float(delta_time), float(mass), prec, out)
    return out


def init_potential_operator(P_potential: torch.Tensor | None, V: torch.Tensor, delta_time: float) -> torch.Tensor:
    """Device table cexp_i(-dt V); written into ``P_potential`` when given (same shape as V)."""
    if P_potential is None:
        P_potential = torch.empty(tuple(V.shape), dtype=torch.complex64, device=V.device)
    prec = _abi.WFS_C128 if P_potential.dtype == torch.complex128 else _abi.WFS_C64
    _abi.potential_operator(V, float(delta_time), prec, P_potential)
    return P_potential
