"""Parareal in time on one GPU, with a learned coarse propagator (no reference counterpart).

The interval [0, T] is cut into N slices of ``steps_per_slice`` fine steps.
With a coarse propagator G and the fine split-step F, iteration k updates

    U_{n+1}^{k+1} = G(U_n^{k+1}) + F(U_n^k) - G(U_n^k)

(Lions, Maday and Turinici 2001; Gander and Vandewalle 2007).  All N fine
propagations of an iteration run as one batched plan (one launch per FFT
pass for every slice); the coarse sweep is sequential.  The correction sum is
formed in FP64 (``wfs_parareal_update``) and U is kept in complex128; the
fine and coarse operators see it rounded to complex64.

After k iterations the first k slices equal the serial fine solution exactly:
U_1^1 = G(U_0) + F(U_0) - G(U_0) is F(U_0) after rounding to complex64, and
so on by induction, because G is deterministic.  F and G must be the same maps
in every call: their plans restart the dithered table-copy schedule at copy 0
before each call (a plain run instead lets the copies run on from call to call).

Coarse propagators:
  IdentityCoarse    G(psi) = psi
  SplitStepCoarse   a few large split-step steps over the slice
  FNOCoarse         a trained Fourier neural operator, optionally as a
                    correction on top of a SplitStepCoarse (its residual
                    output is added to the base propagation)

Fallback: when the first iteration's relative change exceeds
``fallback_threshold`` (the coarse propagator is too poor to help) or the
iteration limit is reached, the remaining slices are finished by the serial
fine propagator from the last exact slice, so the result is always the fine
solution.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _abi
from .fno import FNO3d

C64 = torch.complex64


# ---------------------------------------------------------------------------- problem set-up
@dataclass
class LatticeProblem:
    """Harmonic well plus cos^2 optical lattice and a Gaussian packet on a cubic grid."""
    n: int
    dx: float
    omega: float
    depth: float
    period: float
    sigma: float
    r0: Tuple[float, float, float]
    k0: Tuple[float, float, float]
    dt: float = 0.005
    mass: float = 1.0

    @property
    def shape(self) -> Tuple[int, int, int]:
        return (self.n,) * 3

    def potential(self, device=None) -> torch.Tensor:
        V = torch.zeros(self.shape, dtype=C64, device=device or _cuda())
        _abi.add_harmonic(self.shape, (self.dx,) * 3, self.omega, self.mass, (0.0, 0.0, 0.0), V)
        if self.depth != 0.0:
            _abi.add_lattice(self.shape, (self.dx,) * 3, self.depth, self.period, V)
        return V

    def axes(self):
        """Per-axis parts (fx, fy, fz) of the potential, which is separable (FP64 host arrays)."""
        return _abi.separable_axes(self.shape, (self.dx,) * 3, self.omega, self.mass, (0.0, 0.0, 0.0), self.depth,
                                   self.period)

    def packet(self, device=None) -> torch.Tensor:
        psi = torch.empty(self.shape, dtype=C64, device=device or _cuda())
        _abi.init_gaussian(self.shape, (self.dx,) * 3, _abi.WFS_C64, (self.sigma,) * 3, self.r0, self.k0, psi)
        return psi


def random_lattice_problem(rng: np.random.Generator, n: int = 128, dx: float = 0.25, dt: float = 0.005,
                           omega=(0.05, 0.3), depth=(0.0, 2.0), period=(2.0, 6.0), sigma=(1.5, 3.0),
                           r0=3.0, k0=1.0) -> LatticeProblem:
    """omega, depth, period, sigma uniform in the given ranges; r0 and k0 components uniform in +-r0, +-k0."""
    return LatticeProblem(n=n, dx=dx, omega=float(rng.uniform(*omega)), depth=float(rng.uniform(*depth)),
                          period=float(rng.uniform(*period)), sigma=float(rng.uniform(*sigma)),
                          r0=tuple(float(v) for v in rng.uniform(-r0, r0, 3)),
                          k0=tuple(float(v) for v in rng.uniform(-k0, k0, 3)), dt=dt)


def _cuda():
    return torch.device("cuda", torch.cuda.current_device())


# ---------------------------------------------------------------------------- coarse propagators
class IdentityCoarse:
    name = "identity"

    def __call__(self, src: torch.Tensor, dst: torch.Tensor) -> None:
        dst.copy_(src)


class SplitStepCoarse:
    """``steps`` split-step steps of size slice_time / steps (one plan, potential cached)."""

    name = "splitstep"

    def __init__(self, n, dx, slice_time: float, mass: float, V: torch.Tensor, steps: int, batch: int = 1,
                 axes=None):
        self.steps = int(steps)
        self.plan = _abi.Plan(n, dx, slice_time / self.steps, mass, _abi.WFS_C64, _abi.WFS_METHOD_FOURIER,
                              batch=batch)
        if axes is not None:
            self.plan.set_potential_axes(*axes)
        else:
            self.plan.set_potential(V)

    def __call__(self, src: torch.Tensor, dst: torch.Tensor) -> None:
        if dst.data_ptr() != src.data_ptr():
            dst.copy_(src)
        self.plan.set_step_index(0)        # the same map on every call
        self.plan.step(dst, None, self.steps)


def pack_features(src: torch.Tensor, V: torch.Tensor, pscale: float, vscale: float, a: torch.Tensor,
                  base_out: Optional[torch.Tensor] = None) -> None:
    """Network input planes for state src: [Re psi, Im psi, V] scaled by (pscale, pscale, vscale),
    preceded by [Re B(psi), Im B(psi)] when the base output B(psi) is given."""
    lib, s = _abi.lib(), _abi.stream_handle()
    P = src.numel()
    if base_out is not None:
        _abi.check(lib.wfs_fno_pack(_abi.ptr(base_out), _abi.WFS_C64, None, float(pscale), 0.0, _abi.ptr(a), P, s),
                   "wfs_fno_pack")
        a = a[2 * P:]
    _abi.check(lib.wfs_fno_pack(_abi.ptr(src), _abi.WFS_C64, _abi.ptr(V), float(pscale), float(vscale),
                                _abi.ptr(a), P, s), "wfs_fno_pack")


def pack_target(psi: torch.Tensor, pscale: float, out: torch.Tensor) -> None:
    """Scaled [Re psi, Im psi] planes (the training target)."""
    _abi.check(_abi.lib().wfs_fno_pack(_abi.ptr(psi), _abi.WFS_C64, None, float(pscale), 0.0, _abi.ptr(out),
                                       psi.numel(), _abi.stream_handle()), "wfs_fno_pack")


class FNOCoarse:
    """G(psi) from a trained network.

    Without a base propagator the features are [Re psi, Im psi, V] and the
    network's residual output is a correction to psi; with a base B (a
    SplitStepCoarse) they are [Re B(psi), Im B(psi), Re psi, Im psi, V] and the
    output is a correction to B(psi).
    """

    name = "fno"

    def __init__(self, net: FNO3d, V: torch.Tensor, pscale: float, vscale: float,
                 base: Optional[SplitStepCoarse] = None):
        self.net, self.base = net, base
        self.pscale, self.vscale = float(pscale), float(vscale)
        self.V = V
        self.P = net.P
        want = 5 if base is not None else 3
        if net.cin != want or net.cout != 2 or not net.residual:
            raise ValueError(f"network must map {want} feature planes to 2 with a residual")
        if V.numel() != self.P:
            raise ValueError("potential and network grids differ")
        self.a = torch.empty(net.cin * self.P, dtype=torch.float32, device=V.device)
        self.tmp = torch.empty(V.shape, dtype=C64, device=V.device) if base is not None else None

    def __call__(self, src: torch.Tensor, dst: torch.Tensor) -> None:
        if self.base is not None:
            self.base(src, self.tmp)
        pack_features(src, self.V, self.pscale, self.vscale, self.a, self.tmp)
        out = self.net.forward(self.a)
        _abi.check(_abi.lib().wfs_fno_unpack(_abi.ptr(out), 1.0 / self.pscale, _abi.ptr(dst), self.P,
                                             _abi.stream_handle()), "wfs_fno_unpack")


class GraphedCoarse:
    """A coarse propagator replayed from one hipGraph: the call G(src) -> dst of ``coarse`` is
    captured once on fixed buffers (the first call runs it eagerly, which also allocates every
    workspace it needs, then captures), and each later call copies src in, replays the graph
    and copies the result out.  Same results as ``coarse``; the launch overhead of its many
    small kernels goes away."""

    def __init__(self, coarse, shape, device=None):
        self.coarse = coarse
        self.name = getattr(coarse, "name", "coarse") + "_graph"
        dev = device or _cuda()
        self.g_in = torch.empty(tuple(shape), dtype=C64, device=dev)
        self.g_out = torch.empty(tuple(shape), dtype=C64, device=dev)
        self.graph = None

    def _capture(self) -> None:
        self.coarse(self.g_in, self.g_out)            # eager: allocates workspaces, tables
        torch.cuda.synchronize()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            with _abi.Graph(side) as g:
                self.coarse(self.g_in, self.g_out)
        torch.cuda.synchronize()
        self.graph = g

    def __call__(self, src: torch.Tensor, dst: torch.Tensor) -> None:
        if self.graph is None:
            self.g_in.copy_(src)
            self._capture()
        self.g_in.copy_(src)
        self.graph.launch(torch.cuda.current_stream())
        dst.copy_(self.g_out)


class GridResampler:
    """Band-limited resampling between an n^3 and an nc^3 grid (complex64): ortho transform on
    one grid, the shared low frequencies copied (``wfs_spectral_resample``, zeros above the
    coarse band when going up), ortho inverse on the other.  Values of a function band-limited
    to the coarse grid are kept exactly (up to float rounding).  With a shift s (integer
    frequencies per axis) the coarse grid holds the band around fine frequency s instead of 0:
    ``down`` returns the function times exp(-2 pi i s.j / n) resampled, ``up_add`` multiplies back."""

    def __init__(self, n: int, nc: int, device=None):
        self.n, self.nc = int(n), int(nc)
        dev = device or _cuda()
        fourier = _abi.WFS_METHOD_FOURIER
        self.fine_plan = _abi.Plan((self.n,) * 3, (1.0,) * 3, 1.0, 1.0, _abi.WFS_C64, fourier)
        self.coarse_plan = _abi.Plan((self.nc,) * 3, (1.0,) * 3, 1.0, 1.0, _abi.WFS_C64, fourier)
        self.fine_buf = torch.empty((self.n,) * 3, dtype=C64, device=dev)
        self.coarse_buf = torch.empty((self.nc,) * 3, dtype=C64, device=dev)

    def down(self, x: torch.Tensor, out: torch.Tensor, shift=None) -> torch.Tensor:
        """out (nc^3) = x (n^3) resampled (the band around fine frequency shift); x is not changed."""
        self.fine_buf.copy_(x)
        self.fine_plan.fft3d(self.fine_buf, -1)
        _abi.spectral_resample(self.fine_buf, out, shift=shift)
        self.coarse_plan.fft3d(out, +1)
        return out

    def up_add(self, xc: torch.Tensor, dst: torch.Tensor, shift=None) -> None:
        """dst (n^3, contiguous complex64) += xc (nc^3) resampled, its band placed around fine
        frequency shift (xc is overwritten by its spectrum)."""
        if dst.dtype != C64 or not dst.is_contiguous() or dst.numel() != self.fine_buf.numel():
            raise ValueError("dst must be a contiguous complex64 tensor of the fine grid")
        self.coarse_plan.fft3d(xc, -1)
        back = None if shift is None else tuple(-int(v) for v in shift)
        _abi.spectral_resample(xc, self.fine_buf, shift=back)
        self.fine_plan.fft3d(self.fine_buf, +1)
        n2 = 2 * self.fine_buf.numel()
        _abi.check(_abi.lib().wfs_fno_axpby(_abi.ptr(dst), _abi.ptr(self.fine_buf), 1.0, 1.0, n2, _abi.stream_handle()),
                   "wfs_fno_axpby")


class CoarseGridFNOCoarse:
    """G(psi) = B(psi) + the network's correction computed on a coarser grid.

    The features [Re B(psi), Im B(psi), Re psi, Im psi, V] are resampled to the network's nc^3
    grid (``GridResampler``), the network predicts B(psi) + c there, and c, the correction, is
    resampled back and added to the fine B(psi).  The network is the same Fourier neural
    operator (4 layers, 12 modes, width 32 in the parareal workload) at nc^3 instead of n^3
    points: a call costs the base step, four transforms on each grid and one network pass at
    (nc/n)^3 of the points.  The correction is band-limited to the coarse grid.

    ``shift`` (integer frequencies per axis, default none) centres the coarse band on a carrier
    wave number, the packet's k0 L / (2 pi): the states enter the network as their envelopes
    (times exp(-i k.r) for that carrier), whose spectrum the network's low modes cover, and the
    correction is modulated back.  V enters unshifted (a pointwise factor commutes with the
    modulation).
    """

    name = "fno_coarse_grid"

    def __init__(self, net: FNO3d, V: torch.Tensor, pscale: float, vscale: float, base: SplitStepCoarse,
                 shift=None):
        if net.cin != 5 or net.cout != 2 or not net.residual:
            raise ValueError("network must map 5 feature planes to 2 with a residual")
        self.net, self.base = net, base
        self.pscale, self.vscale = float(pscale), float(vscale)
        self.shift = None if shift is None or not any(shift) else tuple(int(v) for v in shift)
        n = int(round(V.numel() ** (1.0 / 3.0)))
        if n ** 3 != V.numel() or net.n > n:
            raise ValueError("the potential's grid must be cubic and at least the network's")
        self.rs = GridResampler(n, net.n, V.device)
        cshape = (net.n,) * 3
        self.Vc = self.rs.down(V, torch.empty(cshape, dtype=C64, device=V.device))
        self.Pc = net.P
        self.a = torch.empty(5 * self.Pc, dtype=torch.float32, device=V.device)
        self.src_c = torch.empty(cshape, dtype=C64, device=V.device)
        self.base_c = torch.empty(cshape, dtype=C64, device=V.device)
        self.corr = torch.empty(2 * self.Pc, dtype=torch.float32, device=V.device)
        self.corr_c = torch.empty(cshape, dtype=C64, device=V.device)

    def features(self, src: torch.Tensor, base_out: torch.Tensor, Vc: torch.Tensor, a: torch.Tensor) -> None:
        """Network input at the coarse grid for state src with base output base_out (fine)."""
        self.rs.down(base_out, self.base_c, self.shift)
        self.rs.down(src, self.src_c, self.shift)
        pack_features(self.src_c, Vc, self.pscale, self.vscale, a, self.base_c)

    def target(self, psi: torch.Tensor, out: torch.Tensor) -> None:
        """Training target: psi (fine) resampled to the coarse grid, scaled planes."""
        self.rs.down(psi, self.src_c, self.shift)
        pack_target(self.src_c, self.pscale, out)

    def __call__(self, src: torch.Tensor, dst: torch.Tensor) -> None:
        self.base(src, dst)                                  # dst = B(src), fine grid
        self.features(src, dst, self.Vc, self.a)
        out = self.net.forward(self.a)                       # B(src) + c, coarse, scaled planes
        lib, s = _abi.lib(), _abi.stream_handle()
        P2 = 2 * self.Pc
        self.corr.copy_(out[:P2])
        _abi.check(lib.wfs_fno_axpby(_abi.ptr(self.corr), _abi.ptr(self.a), -1.0, 1.0, P2, s), "wfs_fno_axpby")
        _abi.check(lib.wfs_fno_unpack(_abi.ptr(self.corr), 1.0 / self.pscale, _abi.ptr(self.corr_c), self.Pc, s),
                   "wfs_fno_unpack")
        self.rs.up_add(self.corr_c, dst, self.shift)


# ---------------------------------------------------------------------------- parareal driver
@dataclass
class PararealResult:
    state: torch.Tensor                        # complex64 psi(T)
    iterations: int                            # parareal iterations done (fine sweeps)
    converged: bool
    fallback: bool                             # finished by the serial fine propagator
    history: List[float] = field(default_factory=list)   # max relative change per iteration
    timings: Dict[str, float] = field(default_factory=dict)


class Parareal:
    """N slices of ``steps_per_slice`` fine steps on one grid and potential."""

    def __init__(self, n, dx, dt: float, mass: float, V: torch.Tensor, steps_per_slice: int, slices: int,
                 autotune: bool = False, axes=None):
        self.n = tuple(int(v) for v in n)
        self.dx = tuple(float(v) for v in dx)
        self.dt, self.mass = float(dt), float(mass)
        self.steps = int(steps_per_slice)
        self.N = int(slices)
        if self.N < 1 or self.steps < 1:
            raise ValueError("slices and steps_per_slice must be >= 1")
        dev = V.device
        self.V = V
        self.fine = _abi.Plan(self.n, self.dx, self.dt, self.mass, _abi.WFS_C64, _abi.WFS_METHOD_FOURIER,
                              batch=self.N)
        self.serial = _abi.Plan(self.n, self.dx, self.dt, self.mass, _abi.WFS_C64, _abi.WFS_METHOD_FOURIER)
        if axes is not None:
            # separable potential: per-axis phase factors, one copy per slice
            self.fine.set_potential_axes(*(np.tile(np.asarray(a, dtype=np.float64), (self.N, 1)) for a in axes))
            self.serial.set_potential_axes(*axes)
        else:
            self.fine.set_potential(V.expand((self.N,) + self.n).contiguous())
            self.serial.set_potential(V)
        if autotune:
            self.fine.autotune()
            self.serial.autotune()
        self.P = int(np.prod(self.n))
        shape = (self.N,) + self.n
        self.Fin = torch.empty(shape, dtype=C64, device=dev)         # fine inputs, then outputs
        self.Gold = torch.empty(shape, dtype=C64, device=dev)        # G(U_n^k)
        self.Gnew = torch.empty(self.n, dtype=C64, device=dev)
        self.cur = torch.empty(self.n, dtype=C64, device=dev)
        self.U = torch.empty((self.N + 1,) + self.n, dtype=torch.complex128, device=dev)
        self.partials = torch.empty(8192, dtype=torch.float64, device=dev)
        self.norms = torch.empty((self.N, 2), dtype=torch.float64, device=dev)

    @property
    def slice_time(self) -> float:
        return self.steps * self.dt

    # -- serial fine reference ---------------------------------------------------------------
    def serial_fine(self, psi0: torch.Tensor, keep: bool = False):
        """psi(T) by the plain fine propagator; with keep, the state after every slice as well."""
        psi = psi0.clone()
        states = [psi.clone()] if keep else None
        for _ in range(self.N):
            self.serial.set_step_index(0)
            self.serial.step(psi, None, self.steps)
            if keep:
                states.append(psi.clone())
        return (psi, states) if keep else psi

    # -- parareal ------------------------------------------------------------------------------
    def _narrow(self, src: torch.Tensor, dst: torch.Tensor) -> None:
        _abi.check(_abi.lib().wfs_narrow_c128(_abi.ptr(src), _abi.ptr(dst), self.P, _abi.stream_handle()),
                   "wfs_narrow_c128")

    def _update(self, gnew, fold, gold, U, out2) -> None:
        _abi.check(_abi.lib().wfs_parareal_update(_abi.ptr(gnew), _abi.ptr(fold), _abi.ptr(gold), _abi.ptr(U),
                                                  self.P, _abi.ptr(self.partials), self.partials.numel(),
                                                  _abi.ptr(out2), _abi.stream_handle()), "wfs_parareal_update")

    def run(self, psi0: torch.Tensor, coarse, max_iterations: Optional[int] = None, tol: float = 1e-5,
            fallback_threshold: float = 1.0, min_iterations: int = 1, record=None) -> PararealResult:
        """Iterate until the max relative change of any slice end-state is below tol.

        ``record(k, U)`` (optional) is called after every iteration with the
        complex128 slice states U[0..N].
        """
        _abi.check_state(psi0, self.n, _abi.WFS_C64)
        N = self.N
        K = N if max_iterations is None else min(int(max_iterations), N)
        ev = {k: torch.cuda.Event(enable_timing=True) for k in ("t0", "t1")}
        t_coarse = t_fine = t_corr = 0.0
        wall0 = time.perf_counter()
        torch.cuda.synchronize()

        def tick():
            torch.cuda.synchronize()
            return time.perf_counter()

        t = tick()
        self.U[0].copy_(psi0)
        self.cur.copy_(psi0)
        for n in range(N):                                   # initial coarse sweep
            self.Fin[n].copy_(self.cur)
            coarse(self.cur, self.Gold[n])
            self.U[n + 1].copy_(self.Gold[n])
            self.cur.copy_(self.Gold[n])
        t2 = tick()
        t_coarse += t2 - t
        history: List[float] = []
        converged = fallback = False
        k = 0
        while k < K:
            t = tick()
            self.fine.set_step_index(0)                      # F is the same map in every iteration
            self.fine.step(self.Fin, None, self.steps)       # Fin[n] = F(U_n^k), all slices at once
            t2 = tick()
            t_fine += t2 - t
            k += 1
            self.cur.copy_(psi0)
            for n in range(N):                               # sequential correction sweep
                coarse(self.cur, self.Gnew)
                self._update(self.Gnew, self.Fin[n], self.Gold[n], self.U[n + 1], self.norms[n])
                self.Gold[n].copy_(self.Gnew)
                self.Fin[n].copy_(self.cur)                  # next fine input of slice n: U_n^{k}
                self._narrow(self.U[n + 1], self.cur)
            t3 = tick()
            t_corr += t3 - t2                                # coarse + FP64 correction (interleaved)
            nr = self.norms.cpu().numpy()
            change = float(np.sqrt(np.max(nr[:, 0] / np.maximum(nr[:, 1], 1e-300))))
            history.append(change)
            if record is not None:
                record(k, self.U)
            if k >= min_iterations and change < tol:
                converged = True
                break
            if k == 1 and change > fallback_threshold:
                fallback = True
                break
        state = torch.empty(self.n, dtype=C64, device=psi0.device)
        if converged:
            self._narrow(self.U[N], state)
        else:
            # slices 0..k are exact: finish serially from U_k
            fallback = True
            t = tick()
            self._narrow(self.U[k], state)
            for _ in range(k, N):
                self.serial.set_step_index(0)
                self.serial.step(state, None, self.steps)
            t_fine += tick() - t
        total = time.perf_counter() - wall0
        return PararealResult(state=state, iterations=k, converged=converged, fallback=fallback, history=history,
                              timings={"coarse_initial_s": t_coarse, "fine_s": t_fine,
                                       "coarse_and_correction_s": t_corr, "total_s": total})


# ---------------------------------------------------------------------------- training data
class TrajectoryPairs:
    """Pairs (psi(t), psi(t + slice)) generated by the fine solver.

    ``count`` trajectories (each its own problem) advance in lockstep in one
    batched plan; every call of ``advance`` moves all of them by one slice and
    yields ``count`` new pairs.  The coarse base propagation of the inputs, if
    a base is used, runs batched as well.
    """

    def __init__(self, problems: Sequence[LatticeProblem], steps_per_slice: int, base_steps: int = 0,
                 autotune: bool = False):
        p0 = problems[0]
        for p in problems:
            if (p.n, p.dx, p.dt, p.mass) != (p0.n, p0.dx, p0.dt, p0.mass):
                raise ValueError("all trajectories must share grid, dt and mass")
        self.problems = list(problems)
        self.B = len(problems)
        self.n = p0.shape
        self.steps = int(steps_per_slice)
        dev = _cuda()
        self.V = torch.stack([p.potential(dev) for p in problems]).contiguous()
        self.psi = torch.stack([p.packet(dev) for p in problems]).contiguous()
        self.prev = torch.empty_like(self.psi)
        dx = (p0.dx,) * 3
        axes = [p.axes() for p in problems]
        stacked = tuple(np.stack([a[d] for a in axes]) for d in range(3))
        self.fine = _abi.Plan(self.n, dx, p0.dt, p0.mass, _abi.WFS_C64, _abi.WFS_METHOD_FOURIER, batch=self.B)
        self.fine.set_potential_axes(*stacked)
        if autotune:
            self.fine.autotune()
        self.base = None
        self.base_out = None
        if base_steps > 0:
            self.base = SplitStepCoarse(self.n, dx, self.steps * p0.dt, p0.mass, self.V, base_steps, batch=self.B,
                                        axes=stacked)
            self.base_out = torch.empty_like(self.psi)

    def advance(self) -> None:
        """prev = psi(t); psi = psi(t + slice); base_out = B(prev) when a base is used."""
        self.prev.copy_(self.psi)
        self.fine.step(self.psi, None, self.steps)
        if self.base is not None:
            self.base(self.prev, self.base_out)


def _mix_planes(dst: torch.Tensor, x: torch.Tensor, y: torch.Tensor, a: complex, b: complex, P: int,
                complex_planes: int) -> None:
    """dst = a x + b y over ``complex_planes`` (Re, Im) plane pairs; later planes copied from x."""
    def axpby(out, src, alpha, beta):
        _abi.check(_abi.lib().wfs_fno_axpby(_abi.ptr(out), _abi.ptr(src), float(alpha), float(beta), P,
                                            _abi.stream_handle()), "wfs_fno_axpby")
    for c in range(complex_planes):
        re, im = slice(2 * c * P, (2 * c + 1) * P), slice((2 * c + 1) * P, (2 * c + 2) * P)
        axpby(dst[re], x[re], a.real, 0.0)
        axpby(dst[re], x[im], -a.imag, 1.0)
        axpby(dst[re], y[re], b.real, 1.0)
        axpby(dst[re], y[im], -b.imag, 1.0)
        axpby(dst[im], x[im], a.real, 0.0)
        axpby(dst[im], x[re], a.imag, 1.0)
        axpby(dst[im], y[im], b.real, 1.0)
        axpby(dst[im], y[re], b.imag, 1.0)
    rest = slice(2 * complex_planes * P, x.numel())
    if rest.start < x.numel():
        axpby(dst[rest], x[rest], 1.0, 0.0)


def train_coarse(net: FNO3d, pairs: TrajectoryPairs, slices: int, pscale: float, vscale: float,
                 lr: float = 1e-3, lr_final: float = 1e-4, epochs: int = 1, seed: int = 0, mix: float = 0.0,
                 log=None, featurizer: Optional["CoarseGridFNOCoarse"] = None) -> List[float]:
    """Train ``net`` on ``slices`` advances of every trajectory (slices * B pairs).

    epochs == 1: online, each pair is used once right after the solver made it
    (no pair storage, for large grids); epochs > 1: the pair set is stored as
    feature planes and visited ``epochs`` times in shuffled order.  Adam with
    cosine learning-rate decay from lr to lr_final.  Returns every step's loss.
    With a ``featurizer`` (a CoarseGridFNOCoarse of this network) the features and
    targets are the pairs resampled to the network's coarser grid.
    """
    if (pairs.base is None) != (net.cin == 3):
        raise ValueError("network input width does not match the pair features")
    if featurizer is not None and pairs.base is None:
        raise ValueError("the coarse-grid features need the base propagation")
    dev = pairs.psi.device
    coarse_V = None
    if featurizer is not None:
        rs = featurizer.rs
        coarse_V = [rs.down(pairs.V[b], torch.empty((net.n,) * 3, dtype=C64, device=dev)) for b in range(pairs.B)]
    B, P = pairs.B, net.P
    total = slices * B * epochs
    rng = np.random.default_rng(seed)
    dout = torch.empty(2 * P, dtype=torch.float32, device=dev)
    losses: List[float] = []

    def step(feat, targ):
        it = len(losses)
        lr_t = lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * it / max(total - 1, 1)))
        net.zero_grad()
        out = net.forward(feat, keep=True)
        losses.append(net.loss_and_grad(out, targ, dout, reference=feat[: 2 * P]))
        net.backward(dout)
        net.adam(lr=lr_t)
        if log is not None and (it % 50 == 0 or it == total - 1):
            log(it, losses[-1])

    def make(b, feat, targ):
        base_out = pairs.base_out[b] if pairs.base is not None else None
        if featurizer is not None:
            featurizer.features(pairs.prev[b], base_out, coarse_V[b], feat)
            featurizer.target(pairs.psi[b], targ)
            return
        pack_features(pairs.prev[b], pairs.V[b], pscale, vscale, feat, base_out)
        pack_target(pairs.psi[b], pscale, targ)

    if epochs == 1:
        feat = torch.empty(net.cin * P, dtype=torch.float32, device=dev)
        targ = torch.empty(2 * P, dtype=torch.float32, device=dev)
        for _ in range(slices):
            pairs.advance()
            for b in rng.permutation(B):
                make(int(b), feat, targ)
                step(feat, targ)
        return losses
    feats = torch.empty((slices * B, net.cin * P), dtype=torch.float32, device=dev)
    targs = torch.empty((slices * B, 2 * P), dtype=torch.float32, device=dev)
    for s_ in range(slices):
        pairs.advance()
        for b in range(B):
            make(b, feats[s_ * B + b], targs[s_ * B + b])
    mf = torch.empty(net.cin * P, dtype=torch.float32, device=dev) if mix > 0 else None
    mt = torch.empty(2 * P, dtype=torch.float32, device=dev) if mix > 0 else None
    nplanes = (net.cin - 1) // 2          # complex feature planes before the potential plane
    for e in range(epochs):
        for j in rng.permutation(slices * B):
            j = int(j)
            if mix > 0 and slices > 1 and rng.random() < mix:
                b, s1 = j % B, j // B
                s2 = int(rng.choice([t for t in range(slices) if t != s1]))
                ph = rng.uniform(0, 2 * np.pi, 2)
                w = rng.uniform(0.3, 1.0, 2)
                a1, a2 = w[0] * np.exp(1j * ph[0]), w[1] * np.exp(1j * ph[1]) * rng.choice([-1.0, 1.0])
                j2 = s2 * B + b
                _mix_planes(mf, feats[j], feats[j2], a1, a2, P, nplanes)
                _mix_planes(mt, targs[j], targs[j2], a1, a2, P, 1)
                step(mf, mt)
            else:
                step(feats[j], targs[j])
    return losses
