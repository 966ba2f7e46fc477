"""Deferred solver checks (Simulation.defer_solver_checks, pict_piso_step_async / pict_piso_step_vjp_async /
pict_step_wait): native steps and checkpointed reverse passes queued without the host wait for their solver
results.  The results are those of the waiting calls, flush_solver_checks raises what the waiting calls
would have raised (and collects every queued step), and more than 16 queued steps are collected on the way."""
import pytest
import torch

import scenarios
from pict_amd import PISOtorch_diff as PD
from pict_amd import PISOtorch_simulation as S
from pict_amd import checkpoint as CK

pytestmark = pytest.mark.gpu


def _rollout(defer, steps=4, max_iter=5000):
    dom = scenarios.periodic_two_block(dtype=torch.float32, device="cuda", nx=64, ny=32, seed=5)
    sim = S.Simulation(dom, time_step=0.02, corrector_steps=2, non_orthogonal=False, advection_tol=1e-6,
                       pressure_tol=1e-6, advect_passive_scalar=False)
    sim.linear_solve_max_iterations = max_iter
    sim.defer_solver_checks = defer
    srcs = [torch.zeros_like(b.velocity).requires_grad_(True) for b in dom.blocks]
    for b, s in zip(dom.blocks, srcs):
        b.setVelocitySource(s)
    dom.UpdateDomainData()
    out = CK.rollout(sim, steps, params=srcs, checkpoint=True)
    loss = sum((o ** 2).mean() for o in out)
    loss.backward()
    sim.flush_solver_checks()
    return [o.detach() for o in out], [s.grad for s in srcs], sim


def test_deferred_rollout_matches_waiting_calls():
    """Outputs bitwise; gradients to 1e-6 relative: the reverse pass accumulates some gradients with atomic adds
    (as the reference's backward kernels do), so two runs of the same waiting rollout differ at 1e-7."""
    print("case deferred rollout", flush=True)
    out_w, g_w, _ = _rollout(False)
    out_d, g_d, sim = _rollout(True)
    assert not sim._pending_checks
    for a, b in zip(out_w, out_d):
        assert torch.equal(a, b)
    for a, b in zip(g_w, g_d):
        err = float((a - b).abs().max() / b.abs().max())
        print(" gradient difference %.2e" % err, flush=True)
        assert err < 1e-6


def test_deferred_errors_are_raised_at_the_flush():
    print("case deferred solver failure", flush=True)
    with pytest.raises(PD.LinsolveError):
        _rollout(False, steps=2, max_iter=2)
    with pytest.raises(PD.LinsolveError):
        _rollout(True, steps=2, max_iter=2)


def test_more_than_sixteen_queued_steps():
    print("case 20 deferred forward steps", flush=True)
    res = []
    for defer in (False, True):
        dom = scenarios.periodic_two_block(dtype=torch.float32, device="cuda", nx=32, ny=16, seed=6)
        sim = S.Simulation(dom, time_step=0.02, corrector_steps=2, non_orthogonal=False, advection_tol=1e-6,
                           pressure_tol=1e-6, advect_passive_scalar=False)
        sim.defer_solver_checks = defer
        assert sim._PISO_split_step(20)
        assert len(sim._pending_checks) <= 16
        sim.flush_solver_checks()
        assert sim.total_step == 20
        res.append([b.velocity.clone() for b in dom.blocks])
    for a, b in zip(*res):
        assert torch.equal(a, b)
