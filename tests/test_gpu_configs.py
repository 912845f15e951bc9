"""The benchmark workloads at their full sizes, against the complex128 path.

The complex128 path is pinned to the CPU FP64 oracle at 1e-12 on every kernel
(test_gpu_kernels, test_gpu_slab, test_gpu_dropin), so at sizes the CPU oracle
cannot reach it is the FP64 reference.  Every result states its relative L2
error and its phase error arg<psi_ref, psi> against that reference; the bounds
below are the stated complex64 horizons of these workloads.

  256^3 double slit + CAP     split-step, 1000 steps observed every 10; power
                              series p = 10
  48 packets, one hipGraph    {64, 96, 128, 192}^3 harmonic packets
                              (default_rng(0)), 64 steps
  128^3 lattice parareal      32 slices x 50 steps, width-32 / 12-mode FNO
                              correction of a one-step split-step
  512^3 lattice + CAP         split-step 100 steps, power series 20 steps

The fine step they all rest on is sources/sim_state.py:641-697 (split-step)
and :721-752 (power series).
"""

import math

import numpy as np
import pytest
import torch

from wavefunctionsimpython_amd import _abi
from wavefunctionsimpython_amd import parareal as pr
from wavefunctionsimpython_amd.batch import PacketBatch, random_packets
from wavefunctionsimpython_amd.fno import FNO3d

pytestmark = pytest.mark.gpu

C64, C128 = _abi.WFS_C64, _abi.WFS_C128
FOURIER, POWER = _abi.WFS_METHOD_FOURIER, _abi.WFS_METHOD_POWER_SERIES


def errors(psi, ref):
    """(relative L2, phase error arg<ref, psi>) in FP64."""
    a = psi.to(torch.complex128).reshape(-1)
    b = ref.to(torch.complex128).reshape(-1)
    rel = float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
    phase = float(torch.angle(torch.vdot(b, a)))
    return rel, phase


def double_slit_256(dev):
    """256^3, box 80 Bohr: 5 Ha wall 2 Bohr thick with two 1.5 Bohr slits 6 Bohr apart, cubic CAP
    outside +-32 Bohr (-100 Ha), packet sigma 4 at x = -16 with k0 = (2, 0, 0)."""
    n, L = 256, 80.0
    shape, dx = (n,) * 3, (L / n,) * 3
    psi = torch.empty(shape, dtype=torch.complex64, device=dev)
    _abi.init_gaussian(shape, dx, C64, (4.0,) * 3, (-16.0, 0.0, 0.0), (2.0, 0.0, 0.0), psi)
    V = torch.zeros(shape, dtype=torch.complex64, device=dev)
    _abi.add_cap(shape, dx, (-32.0,) * 3, (32.0,) * 3, 0.0, (-100.0,) * 3, (-100.0,) * 3, (3.0,) * 3, (3.0,) * 3, V)
    _abi.add_wall(shape, dx, 5.0, (0.0, 0.0, 0.0), (1.0, 0.0, 0.0), 2.0, 0.0, 1.0, 2, 6.0, 1.5, 0.0, V)
    b, t = int((-32.0 + L / 2) / dx[0]), int((32.0 + L / 2) / dx[0])
    cnt = t + 1 - b
    cnt -= (cnt + 1) % 2
    boxes = [((b, b, b), (cnt, cnt, cnt)), ((b, b, b), (63, cnt, cnt)), ((n // 2, b, b), (101, cnt, cnt))]
    return shape, dx, psi, V, boxes


def test_double_slit_256_split_step_1000_steps_observed(dev):
    """1000 steps of dt 0.005 observed every 10 steps (100 samples): complex64 (the slab scheme,
    one fused observing launch per sample) against the complex128 run."""
    shape, dx, psi0, V, boxes = double_slit_256(dev)
    p32 = _abi.Plan(shape, dx, 0.005, 1.0, C64, FOURIER)
    p64 = _abi.Plan(shape, dx, 0.005, 1.0, C128, FOURIER)
    assert p32.variants()["yz_slab"] != 0
    a = psi0.clone()
    b = psi0.to(torch.complex128)
    obs32 = p32.step_observe(a, V, 1000, 10, boxes)
    obs64 = p64.step_observe(b, V, 1000, 10, boxes)
    torch.cuda.synchronize()
    rel, phase = errors(a, b)
    o32, o64 = obs32.cpu().numpy(), obs64.cpu().numpy()
    norm_err = float(np.max(np.abs(o32[:, 0, 0] - o64[:, 0, 0])))
    x_err = float(np.max(np.abs(o32[:, 0, 1] / o32[:, 0, 0] - o64[:, 0, 1] / o64[:, 0, 0])))
    print(f"256^3 double slit, 1000 split steps: complex64 vs complex128 rel L2 {rel:.2e}, phase {phase:+.1e} rad; "
          f"norm {o64[0, 0, 0]:.6f} -> {o64[-1, 0, 0]:.6f}, max |dP| {norm_err:.1e}, max |d<x>| {x_err:.1e} Bohr")
    # stated complex64 horizon of this workload: 1.3e-5 relative L2 after 1000 steps (measured)
    assert rel < 2.5e-5 and abs(phase) < 1e-6
    assert norm_err < 2.5e-5 and x_err < 1e-4


def test_double_slit_256_power_series(dev):
    """Power series p = 10 at 256^3 (20 steps) against its complex128 run, and after 20 steps
    within its method error of the split-step (7-point Laplacian vs spectral)."""
    shape, dx, psi0, V, _ = double_slit_256(dev)
    p32 = _abi.Plan(shape, dx, 0.005, 1.0, C64, POWER, 10)
    p64 = _abi.Plan(shape, dx, 0.005, 1.0, C128, POWER, 10)
    a = psi0.clone()
    b = psi0.to(torch.complex128)
    p32.step(a, V, 20)
    p64.step(b, V, 20)
    torch.cuda.synchronize()
    rel, phase = errors(a, b)
    fs = psi0.to(torch.complex128)
    _abi.Plan(shape, dx, 0.005, 1.0, C128, FOURIER).step(fs, V, 20)
    method, _ = errors(b, fs)
    print(f"256^3 power series p=10, 20 steps: complex64 vs complex128 rel L2 {rel:.2e}, phase {phase:+.1e}; "
          f"complex128 power series vs split-step {method:.2e}")
    assert rel < 5e-6 and abs(phase) < 5e-6
    assert method < 5e-2


def test_batched_48_packets_graph(dev):
    """48 packets on {64, 96, 128, 192}^3 (default_rng(0)), autotuned, 64 steps in one graph:
    every packet against its complex128 run; one packet per grid bit for bit equal to the same
    packet run alone with the same kernel variants; <x> of every well-contained packet follows
    Ehrenfest, <x>(t) = (k0 / m w) sin(w t) for a packet starting at the origin."""
    steps = 64
    specs = random_packets(48, seed=0)
    batch = PacketBatch(specs)
    g = batch.capture(steps)
    ref = PacketBatch(specs, precision=C128, autotune=False)
    for i in range(len(specs)):
        ref.packet_state(i).copy_(batch.packet_state(i).to(torch.complex128))
    torch.cuda.synchronize()
    g.launch()
    ref.run(steps)
    torch.cuda.synchronize()
    worst = 0.0
    for i in range(len(specs)):
        rel, phase = errors(batch.packet_state(i), ref.packet_state(i))
        worst = max(worst, rel)
        assert rel < 1e-5 and abs(phase) < 1e-5, f"packet {i} {specs[i].n}: {rel:.2e} {phase:+.1e}"
    # bit identity with a lone run, one packet per grid size
    first = {}
    for i, s in enumerate(specs):
        first.setdefault(s.n[0], i)
    for n, i in sorted(first.items()):
        solo = PacketBatch([specs[i]], autotune=False)
        cls = batch.classes[batch._where[i][0]]
        for k, v in cls.plan.variants().items():
            solo.classes[0].plan.set_variant(k, v)
        solo.run(steps)
        torch.cuda.synchronize()
        assert torch.equal(solo.packet_state(0), batch.packet_state(i)), f"packet {i} ({n}^3)"
    # Ehrenfest (the packet's Gaussian tail at the box edge below 1e-8 of its peak), moments over
    # the centred box of 0.9 of each axis
    batch.prepare_observables(1)
    moments = {}
    for cls in batch.classes:
        out = cls.plan.observables(cls.psi, cls.boxes).cpu().numpy().reshape(len(cls.members), -1, 7)
        for j, i in enumerate(cls.members):
            moments[i] = out[j, 0]
    t = steps * specs[0].dt
    checked = 0
    for i, s in enumerate(specs):
        if 0.5 * s.n[0] * s.dx[0] < 6.0 * s.sigma[0]:
            continue
        o = moments[i]
        x = o[1] / o[0]
        want = s.k0[0] / (s.mass * s.omega) * math.sin(s.omega * t)
        assert abs(x - want) < 2e-4, f"packet {i}: <x> {x:.6f} vs {want:.6f}"
        checked += 1
    print(f"48 packets x {steps} steps: worst complex64 vs complex128 rel L2 {worst:.2e}; "
          f"Ehrenfest checked on {checked} packets; bit identity on grids {sorted(first)}")
    assert checked >= 10


def test_parareal_128_width32_fno(dev):
    """128^3 harmonic + cos^2 lattice (default_rng(2)), 32 slices x 50 steps, with the width-32,
    12-mode, 4-layer FNO correction of a one-step split-step as the coarse propagator (trained
    briefly on solver pairs): converges to within 1e-4 of the complex128 run."""
    n, dx, S, NS = 128, 0.25, 50, 32
    rng = np.random.default_rng(1)
    probs = [pr.random_lattice_problem(rng, n=n, dx=dx, k0=0.5, r0=3.0) for _ in range(8)]
    pscale, vscale = math.sqrt((n * dx) ** 3), 0.2
    net = FNO3d(n, width=32, modes=12, layers=4, in_channels=5, output_scale=0.0, residual_scale=1e-3)
    pairs = pr.TrajectoryPairs(probs, S, base_steps=1)
    losses = pr.train_coarse(net, pairs, 4, pscale, vscale, lr=1e-3, lr_final=1e-4, epochs=1)
    del pairs
    test = pr.random_lattice_problem(np.random.default_rng(2), n=n, dx=dx, k0=0.5, r0=3.0)
    V, psi0 = test.potential(), test.packet()
    par = pr.Parareal(test.shape, (dx,) * 3, test.dt, 1.0, V, S, NS, axes=test.axes())
    base = pr.SplitStepCoarse(test.shape, (dx,) * 3, S * test.dt, 1.0, V, 1, axes=test.axes())
    fno = pr.FNOCoarse(net, V, pscale, vscale, base=base)
    res = par.run(psi0, fno, tol=1e-4)
    p64 = _abi.Plan(test.shape, (dx,) * 3, test.dt, 1.0, C128, FOURIER)
    p64.set_potential_axes(*test.axes())
    ref = psi0.to(torch.complex128)
    p64.step(ref, None, S * NS)
    torch.cuda.synchronize()
    rel, phase = errors(res.state, ref)
    print(f"128^3 parareal, 32 x 50 steps, width-32 FNO ({len(losses)} training pairs): {res.iterations} iterations, "
          f"converged {res.converged}, fallback {res.fallback}; vs complex128 rel L2 {rel:.2e}, phase {phase:+.1e}")
    assert res.converged and not res.fallback
    assert rel < 1e-4 and abs(phase) < 1e-4


@pytest.fixture(scope="module")
def lattice_512(dev):
    n, box = 512, 128.0
    shape, dx = (n,) * 3, (box / n,) * 3
    V = torch.zeros(shape, dtype=torch.complex64, device=dev)
    _abi.add_lattice(shape, dx, 3.0, 4.0, V)
    _abi.add_cap(shape, dx, (-48.0,) * 3, (48.0,) * 3, 0.0, (-100.0,) * 3, (-100.0,) * 3, (3.0,) * 3, (3.0,) * 3, V)
    psi0 = torch.empty(shape, dtype=torch.complex64, device=dev)
    _abi.init_gaussian(shape, dx, C64, (5.0,) * 3, (0.0, 0.0, 0.0), (1.0, 1.0, 0.0), psi0)
    yield shape, dx, V, psi0
    del V, psi0
    torch.cuda.empty_cache()


@pytest.mark.parametrize("method,steps,tol", [(FOURIER, 100, 1e-5), (POWER, 20, 2e-6)])
def test_lattice_512(dev, lattice_512, method, steps, tol):
    """512^3 complex64 (cos^2 lattice 3 Ha, period 4 Bohr, + CAP 16 Bohr; sigma 5, k0 (1, 1, 0);
    dt 0.002) against the complex128 run of the same method."""
    shape, dx, V, psi0 = lattice_512
    a = psi0.clone()
    p32 = _abi.Plan(shape, dx, 0.002, 1.0, C64, method, 10)
    p32.step(a, V, steps)
    del p32
    b = psi0.to(torch.complex128)
    p64 = _abi.Plan(shape, dx, 0.002, 1.0, C128, method, 10)
    p64.step(b, V, steps)
    torch.cuda.synchronize()
    del p64
    rel, phase = errors(a, b)
    name = "split-step" if method == FOURIER else "power series p=10"
    print(f"512^3 {name}, {steps} steps: complex64 vs complex128 rel L2 {rel:.2e}, phase {phase:+.1e}")
    del a, b
    torch.cuda.empty_cache()
    assert rel < tol and abs(phase) < tol
