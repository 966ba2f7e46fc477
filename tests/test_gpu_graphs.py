"""Captured graphs of the native step and its reverse pass: replayed many times, with eager work and host
synchronisations between the replays, every replay matches the same work queued directly (PICT_STEP_GRAPHS=0):
the forward step bitwise (its kernels are deterministic), the reverse pass to FP64 round-off (its scatters use
float atomics, whose summation order varies from run to run with or without graphs)."""
import pytest
import torch

import scenarios
from pict_amd import PISOtorch_simulation as S
from pict_amd import checkpoint as CK

pytestmark = pytest.mark.gpu


def _sim(d, dtype):
    tol = 1e-10 if dtype == torch.float64 else 1e-6
    return S.Simulation(d, time_step=0.05, corrector_steps=2, non_orthogonal=False, advection_tol=tol,
                        pressure_tol=tol, differentiable=False, advect_passive_scalar=False)


def _forward_states(name, dtype, steps, graphs, monkeypatch):
    monkeypatch.setenv("PICT_STEP_GRAPHS", "1" if graphs else "0")
    d = scenarios.ALL[name](dtype=dtype, device="cuda")
    sim = _sim(d, dtype)
    assert sim._native_step_eligible()
    out = []
    for i in range(steps):
        assert sim._PISO_split_step(1)
        # eager work on the same stream between replays
        noise = torch.ones(4096, device="cuda").cumsum(0)
        torch.cuda.synchronize()
        assert float(noise[-1]) == 4096.0
        out.append([b.velocity.detach().clone() for b in d.blocks] + [b.pressure.detach().clone() for b in d.blocks])
    return out


@pytest.mark.parametrize("name", ["periodic_two_block", "cavity", "box3d"])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_step_graph_replays_match_direct(name, dtype, monkeypatch):
    print("case", name, dtype, flush=True)
    steps = 6
    a = _forward_states(name, dtype, steps, True, monkeypatch)
    b = _forward_states(name, dtype, steps, False, monkeypatch)
    for i, (x, y) in enumerate(zip(a, b)):
        for t, u in zip(x, y):
            assert torch.equal(t, u), "step %d differs" % i


def _rollout_grads(name, steps, graphs, keep_tape, monkeypatch):
    monkeypatch.setenv("PICT_STEP_GRAPHS", "1" if graphs else "0")
    d = scenarios.ALL[name](dtype=torch.float64, device="cuda")
    g = torch.Generator().manual_seed(3)
    srcs = []
    for b in d.blocks:
        s = ((torch.rand(b.velocity.shape, generator=g, dtype=torch.float64) * 2 - 1) * 0.1).cuda()
        s.requires_grad_(True)
        b.setVelocitySource(s)
        srcs.append(s)
    d.UpdateDomainData()
    sim = _sim(d, torch.float64)
    target = [((torch.rand(b.velocity.shape, generator=g, dtype=torch.float64) * 2 - 1) * 0.3).cuda()
              for b in d.blocks]
    out = CK.rollout(sim, steps, params=srcs, checkpoint=True, keep_tape=keep_tape)
    loss = sum(((o - t) ** 2).sum() for o, t in zip(out, target))
    loss.backward()
    torch.cuda.synchronize()
    return [s.grad.detach().clone() for s in srcs] + [o.detach().clone() for o in out]


@pytest.mark.parametrize("name", ["periodic_two_block", "cavity_static_fields"])
@pytest.mark.parametrize("keep_tape", [False, True])
def test_reverse_pass_graph_replays_match_direct(name, keep_tape, monkeypatch):
    print("case", name, keep_tape, flush=True)
    a = _rollout_grads(name, 5, True, keep_tape, monkeypatch)
    b = _rollout_grads(name, 5, False, keep_tape, monkeypatch)
    for x, y in zip(a, b):
        err = float((x - y).abs().max()) / max(float(y.abs().max()), 1e-300)
        print("  graph vs direct: %.3e" % err, flush=True)
        assert err < 1e-10
