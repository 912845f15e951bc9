"""Many independent packets advanced together (no reference counterpart).

The reference runs one SimState per process.  ``PacketBatch`` takes any
number of packet descriptions of possibly different grid sizes, groups them
into classes of identical (grid, voxel size, dt, mass), stores each class as
one stacked device tensor and advances every class with one batched plan: one
kernel launch per FFT pass covers every packet of the class.  The whole
multi-class run can be captured once into a hipGraph (one branch per class on
its own stream, joined at the end) and replayed with a single launch.

Each packet's state is bit-identical to the same packet run alone (the pass
kernels treat every column independently).  The harmonic + lattice potentials
are separable, so the plans hold per-axis phase factors instead of N^3 tables
(the potential pass reads 16 instead of 24 B per point); the N^3 potentials
are still built for packet_potential() and the observables.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Sequence, Tuple

import numpy as np
import torch

from . import _abi


@dataclass
class PacketSpec:
    """One packet: grid, Gaussian state and harmonic well."""
    n: Tuple[int, int, int]
    dx: Tuple[float, float, float]
    sigma: Tuple[float, float, float]
    r0: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    k0: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    omega: float = 0.0
    lattice_depth: float = 0.0
    lattice_period: float = 1.0
    mass: float = 1.0
    dt: float = 0.01


@dataclass
class _Class:
    key: tuple
    members: List[int]
    plan: _abi.Plan
    psi: torch.Tensor
    V: torch.Tensor
    stream: torch.cuda.Stream
    obs: torch.Tensor | None = None
    boxes: list = field(default_factory=list)


class PacketBatch:
    def __init__(self, specs: Sequence[PacketSpec], precision: int = _abi.WFS_C64, device=None,
                 autotune: bool = True):
        self.specs = list(specs)
        self.precision = precision
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        groups: Dict[tuple, List[int]] = {}
        for i, s in enumerate(self.specs):
            key = (tuple(int(v) for v in s.n), tuple(float(v) for v in s.dx), float(s.dt), float(s.mass))
            groups.setdefault(key, []).append(i)
        cdt = torch.complex128 if precision == _abi.WFS_C128 else torch.complex64
        self.classes: List[_Class] = []
        self._where: Dict[int, Tuple[int, int]] = {}
        for key, members in groups.items():
            n, dx, dt, mass = key
            B = len(members)
            psi = torch.empty((B,) + n, dtype=cdt, device=self.device)
            V = torch.zeros((B,) + n, dtype=torch.complex64, device=self.device)
            for j, i in enumerate(members):
                s = self.specs[i]
                _abi.init_gaussian(n, dx, precision, s.sigma, s.r0, s.k0, psi[j])
                if s.omega:
                    _abi.add_harmonic(n, dx, s.omega, s.mass, (0.0, 0.0, 0.0), V[j])
                if s.lattice_depth:
                    _abi.add_lattice(n, dx, s.lattice_depth, s.lattice_period, V[j])
                self._where[i] = (len(self.classes), j)
            plan = _abi.Plan(n, dx, dt, mass, precision, _abi.WFS_METHOD_FOURIER, batch=B)
            # harmonic + lattice is separable: per-axis phase factors instead of an N^3 table
            axes = [_abi.separable_axes(n, dx, self.specs[i].omega, self.specs[i].mass, (0.0, 0.0, 0.0),
                                        self.specs[i].lattice_depth, self.specs[i].lattice_period) for i in members]
            plan.set_potential_axes(*(np.stack([a[d] for a in axes]) for d in range(3)))
            if autotune:
                plan.autotune()         # after the potential: tunes the separable-potential path
            self.classes.append(_Class(key, members, plan, psi, V, torch.cuda.Stream(device=self.device)))
        torch.cuda.synchronize(self.device)

    # ------------------------------------------------------------------ access
    def packet_state(self, i: int) -> torch.Tensor:
        c, j = self._where[i]
        return self.classes[c].psi[j]

    def packet_potential(self, i: int) -> torch.Tensor:
        c, j = self._where[i]
        return self.classes[c].V[j]

    def total_points(self) -> int:
        return sum(int(np.prod(s.n)) for s in self.specs)

    def set_step_index(self, index: int) -> None:
        """Table copy of every class's next step (a captured graph replays its capture's copies)."""
        for cls in self.classes:
            cls.plan.set_step_index(index)

    # ------------------------------------------------------------------ stepping
    def _enqueue(self, cls: _Class, nsteps: int, observe_every: int, obs_offset: int) -> int:
        """Steps (and observables) of one class on its stream; returns observations written."""
        k = 0
        o = 0
        while k < nsteps:
            if cls.obs is not None and observe_every and k % observe_every == 0:
                cls.plan.observables(cls.psi, cls.boxes, out=cls.obs[obs_offset + o], stream=cls.stream)
                o += 1
            s = min(observe_every - k % observe_every, nsteps - k) if observe_every else nsteps - k
            cls.plan.step(cls.psi, None, s, stream=cls.stream)
            k += s
        return o

    def prepare_observables(self, n_obs: int, box_fraction: float = 0.9):
        """Per-class device buffer for n_obs samples of a centred odd box (fraction of each axis)."""
        for cls in self.classes:
            n = cls.key[0]
            bottom, count = [], []
            for a in range(3):
                c = max(1, int(n[a] * box_fraction))
                c -= (c + 1) % 2
                b = (n[a] - c) // 2
                bottom.append(b)
                count.append(c)
            cls.boxes = [(tuple(bottom), tuple(count))]
            cls.obs = torch.zeros((n_obs, len(cls.members), 1, 7), dtype=torch.float64, device=self.device)

    def run(self, nsteps: int, observe_every: int = 0) -> None:
        """Advance every packet by nsteps, each class on its own stream, eagerly."""
        main = torch.cuda.current_stream(self.device)
        start = torch.cuda.Event()
        start.record(main)
        ends = []
        for cls in self.classes:
            cls.stream.wait_event(start)
            self._enqueue(cls, nsteps, observe_every, 0)
            e = torch.cuda.Event()
            e.record(cls.stream)
            ends.append(e)
        for e in ends:
            main.wait_event(e)

    def capture(self, nsteps: int, observe_every: int = 0) -> _abi.Graph:
        """One hipGraph for nsteps of every class (fork/join across class streams)."""
        cap = torch.cuda.Stream(device=self.device)
        cap.wait_stream(torch.cuda.current_stream(self.device))
        with _abi.Graph(cap) as g:
            start = torch.cuda.Event()
            start.record(cap)
            ends = []
            for cls in self.classes:
                cls.stream.wait_event(start)
                self._enqueue(cls, nsteps, observe_every, 0)
                e = torch.cuda.Event()
                e.record(cls.stream)
                ends.append(e)
            for e in ends:
                cap.wait_event(e)
        g.capture_stream = cap
        return g

    def observations(self, i: int) -> np.ndarray:
        """Host copy of packet i's samples: (n_obs, 7) = P, E[r], E[r^2]."""
        c, j = self._where[i]
        return self.classes[c].obs[:, j, 0, :].cpu().numpy()


def random_packets(count: int = 48, seed: int = 0, sizes=(64, 96, 128, 192), dx: float = 0.5, dt: float = 0.01,
                   omega=(0.05, 0.5), sigma=(1.5, 4.0), k0=(-1.5, 1.5)) -> List[PacketSpec]:
    """The batched workload: grid sizes, harmonic frequencies, widths and momenta drawn from
    numpy.random.default_rng(seed); voxel size fixed (box = N * dx) and the packets centred."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(count):
        n = int(rng.choice(sizes))
        w = float(rng.uniform(*omega))
        s = float(rng.uniform(*sigma))
        k = tuple(float(v) for v in rng.uniform(*k0, size=3))
        specs.append(PacketSpec(n=(n, n, n), dx=(dx, dx, dx), sigma=(s, s, s), k0=k, omega=w, dt=dt))
    return specs
