"""Batched packets (stacked plans, one hipGraph over all size classes)."""

import numpy as np
import pytest
import torch

import oracle
from wavefunctionsimpython_amd import _abi
from wavefunctionsimpython_amd.batch import PacketBatch, PacketSpec, random_packets

pytestmark = pytest.mark.gpu


def _specs():
    rng = np.random.default_rng(3)
    specs = []
    for i in range(7):
        n = (32, 32, 32) if i % 2 else (64, 32, 16)
        s = float(rng.uniform(1.2, 2.0))
        specs.append(PacketSpec(n=n, dx=(0.5, 0.5, 0.5), sigma=(s, s, s),
                                k0=tuple(rng.uniform(-1, 1, 3)), omega=float(rng.uniform(0.1, 0.5)), dt=0.01))
    return specs


def test_batched_packets_bit_identical_to_single_runs(dev):
    batch = PacketBatch(_specs(), autotune=False)
    init = [batch.packet_state(i).clone() for i in range(len(batch.specs))]
    batch.run(13)
    torch.cuda.synchronize()
    for i, s in enumerate(batch.specs):
        p = _abi.Plan(s.n, s.dx, s.dt, s.mass, _abi.WFS_C64, _abi.WFS_METHOD_FOURIER)
        p.set_potential_axes(*_abi.separable_axes(s.n, s.dx, s.omega, s.mass, (0.0, 0.0, 0.0), s.lattice_depth,
                                                  s.lattice_period))
        psi = init[i].clone()
        p.step(psi, None, 13)
        assert torch.equal(psi, batch.packet_state(i)), f"packet {i}"


def test_batched_state_matches_oracle(dev):
    batch = PacketBatch(_specs()[:3], autotune=False)
    i = 2
    psi0 = batch.packet_state(i).cpu().numpy()
    V = batch.packet_potential(i).cpu().numpy()
    batch.run(20)
    torch.cuda.synchronize()
    ref = oracle.split_step(psi0, V, (0.5,) * 3, 0.01, 1.0, nsteps=20)
    assert oracle.relative_l2(batch.packet_state(i).cpu().numpy(), ref) < 1e-5


def test_graph_replays_match_eager(dev):
    eager = PacketBatch(_specs(), autotune=False)
    graphed = PacketBatch(_specs(), autotune=False)
    g = graphed.capture(10)
    assert g.node_count() > 0
    for replay in range(3):
        eager.set_step_index(0)       # a replay uses the table copies of its capture
        eager.run(10)
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        # eager work on the default stream between replays
        junk = torch.ones(1000, device=dev).sum().item()
        assert junk == 1000.0
        for i in range(len(eager.specs)):
            assert torch.equal(eager.packet_state(i), graphed.packet_state(i)), f"replay {replay} packet {i}"


def test_batched_observables_and_ehrenfest(dev):
    specs = [PacketSpec(n=(64, 64, 64), dx=(0.5, 0.5, 0.5), sigma=(1.5, 1.5, 1.5), r0=(1.0, 0.0, 0.0),
                        k0=(0.6, 0.0, -0.3), omega=w, dt=0.01) for w in (0.2, 0.35, 0.5)]
    batch = PacketBatch(specs, autotune=False)
    batch.prepare_observables(n_obs=5)
    b, c = batch.classes[0].boxes[0]
    psi0 = batch.packet_state(1).cpu().numpy()
    g = batch.capture(200, observe_every=40)
    g.launch()
    torch.cuda.synchronize()
    for i, s in enumerate(specs):
        obs = batch.observations(i)
        t = np.arange(5) * 40 * 0.01
        x = 1.0 * np.cos(s.omega * t) + 0.6 / s.omega * np.sin(s.omega * t)
        z = -0.3 / s.omega * np.sin(s.omega * t)
        assert np.allclose(obs[:, 1] / obs[:, 0], x, atol=2e-3)
        assert np.allclose(obs[:, 3] / obs[:, 0], z, atol=2e-3)
    er, _ = oracle.expected_location(psi0, (0.5,) * 3, b, c)
    assert np.allclose(batch.observations(1)[0, 1:4], er, rtol=1e-10, atol=1e-12)


def test_random_workload_shapes():
    specs = random_packets(48, seed=0)
    assert len(specs) == 48
    assert {s.n[0] for s in specs} <= {64, 96, 128, 192}
    assert all(0.05 <= s.omega <= 0.5 and 1.5 <= s.sigma[0] <= 4 for s in specs)
