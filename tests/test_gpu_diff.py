"""Differentiable PISO steps (PISOtorch_diff backend of the simulation driver) on the GPU: gradients of a loss
after a short rollout with respect to the initial velocity, a per-cell velocity source and a static viscosity,
against central finite differences of the forward step (FP64, solver tolerance 1e-12, eps 1e-6).  The bar is the
1e-4 relative agreement of directional derivatives."""
import pytest
import torch

import scenarios
from pict_amd import PISOtorch as P
from pict_amd import PISOtorch_simulation as S

pytestmark = pytest.mark.gpu

DT = 0.05


def _sim(domain, differentiable):
    return S.Simulation(domain, time_step=DT, corrector_steps=2, non_orthogonal=False, advection_tol=1e-12,
                        pressure_tol=1e-12, differentiable=differentiable, advect_passive_scalar=False)


def _setup(name, src_seed=9):
    d = scenarios.ALL[name](dtype=torch.float64, device="cuda")
    g = torch.Generator().manual_seed(src_seed)
    for b in d.blocks:
        if b.velocitySource is None:
            b.setVelocitySource(((torch.rand(b.velocity.shape, generator=g, dtype=torch.float64) * 2 - 1) * 0.1)
                                .cuda().contiguous())
    d.UpdateDomainData()
    return d


def _state(d):
    return [b.velocity.detach().clone() for b in d.blocks], [b.velocitySource.detach().clone() for b in d.blocks]


def _loss(d, target):
    return sum(((b.velocity - t) ** 2).sum() for b, t in zip(d.blocks, target))


def _forward_loss(name, v0, s0, target, steps, visc=None):
    d = _setup(name)
    for b, v, s in zip(d.blocks, v0, s0):
        b.setVelocity(v.clone())
        b.setVelocitySource(s.clone())
    if visc is not None:
        d.setViscosity(visc.clone())
    d.UpdateDomainData()
    _sim(d, False)._PISO_split_step(steps)
    return float(_loss(d, target))


@pytest.mark.parametrize("name", ["periodic_two_block", "cavity", "rotated_connection", "box3d"])
def test_rollout_gradients_match_finite_differences(name):
    print("case", name, flush=True)
    steps = 2
    d = _setup(name)
    v0, s0 = _state(d)
    g = torch.Generator().manual_seed(21)
    target = [((torch.rand(v.shape, generator=g, dtype=torch.float64) * 2 - 1) * 0.3).cuda() for v in v0]
    # differentiable run
    vs = [v.clone().requires_grad_(True) for v in v0]
    ss = [s.clone().requires_grad_(True) for s in s0]
    visc = d.viscosity.clone().requires_grad_(True)
    for b, v, s in zip(d.blocks, vs, ss):
        b.setVelocity(v)
        b.setVelocitySource(s)
    d.setViscosity(visc)
    d.UpdateDomainData()
    assert _sim(d, True)._PISO_split_step(steps)
    L = _loss(d, target)
    L.backward()
    grads_v = [v.grad.detach() for v in vs]
    grads_s = [s.grad.detach() for s in ss]
    assert all(torch.isfinite(x).all() for x in grads_v + grads_s)
    eps = 1e-6
    checks = []
    for kind in ("velocity", "source", "viscosity"):
        if kind == "viscosity":
            if all(b.viscosity is not None for b in d.blocks):
                continue
            an = float(visc.grad[0])
            lp = _forward_loss(name, v0, s0, target, steps, visc=visc.detach() + eps)
            lm = _forward_loss(name, v0, s0, target, steps, visc=visc.detach() - eps)
        else:
            base = v0 if kind == "velocity" else s0
            dirs = [((torch.rand(v.shape, generator=g, dtype=torch.float64) * 2 - 1)).cuda() for v in base]
            grads = grads_v if kind == "velocity" else grads_s
            an = sum(float((gr * dv).sum()) for gr, dv in zip(grads, dirs))
            plus = [x + eps * dv for x, dv in zip(base, dirs)]
            minus = [x - eps * dv for x, dv in zip(base, dirs)]
            if kind == "velocity":
                lp = _forward_loss(name, plus, s0, target, steps)
                lm = _forward_loss(name, minus, s0, target, steps)
            else:
                lp = _forward_loss(name, v0, plus, target, steps)
                lm = _forward_loss(name, v0, minus, target, steps)
        fd = (lp - lm) / (2 * eps)
        rel = abs(fd - an) / max(abs(fd), 1e-12)
        print("  %-9s fd %.10e  adjoint %.10e  rel %.2e" % (kind, fd, an, rel), flush=True)
        checks.append((kind, rel))
    assert checks and all(rel <= 1e-4 for _, rel in checks), checks


@pytest.mark.parametrize("name", ["periodic_two_block", "rotated_connection", "box3d", "cavity_static_fields"])
@pytest.mark.parametrize("keep_tape", [False, True])
def test_checkpointed_rollout_matches_full_graph(name, keep_tape):
    """Per-step checkpointing -- on these orthogonal steps the native path: forward as one library call, reverse
    pass (pict_piso_step_vjp) from a recomputed or kept tape -- gives the gradients of the whole PISOtorch_diff
    autograd graph, w.r.t. the input velocities, the velocity sources and the domain viscosity, and leaves the
    domain at the rollout's final state."""
    from pict_amd import checkpoint as CK
    print("case", name, keep_tape, flush=True)
    steps = 3
    grads = {}
    finals = {}
    for ck in (False, True):
        d = _setup(name)
        v0, s0 = _state(d)
        g = torch.Generator().manual_seed(5)
        target = [((torch.rand(v.shape, generator=g, dtype=torch.float64) * 2 - 1) * 0.3).cuda() for v in v0]
        srcs = [s.clone().requires_grad_(True) for s in s0]
        vin = [v.clone().requires_grad_(True) for v in v0]
        visc = d.viscosity.clone().requires_grad_(True)
        for b, s, v in zip(d.blocks, srcs, vin):
            b.setVelocitySource(s)
            b.setVelocity(v)
        d.setViscosity(visc)
        d.UpdateDomainData()
        sim = _sim(d, False)
        if ck:
            assert sim._native_step_eligible()
        out = CK.rollout(sim, steps, params=srcs + [visc], checkpoint=ck, keep_tape=keep_tape)
        L = sum(((o - t) ** 2).sum() for o, t in zip(out, target))
        L.backward()
        grads[ck] = [x.grad.detach().clone() for x in srcs + vin] + \
            ([visc.grad.detach().clone()] if visc.grad is not None else [])
        finals[ck] = [b.velocity.detach().clone() for b in d.blocks]
        assert sim.total_step == steps
    assert len(grads[True]) == len(grads[False])
    for a, b in zip(grads[True], grads[False]):
        err = float((a - b).abs().max()) / max(float(b.abs().max()), 1e-300)
        print("  checkpointed vs full graph: %.3e" % err, flush=True)
        assert err < 1e-10
    for a, b in zip(finals[True], finals[False]):
        assert torch.equal(a, b)


@pytest.mark.parametrize("name", ["periodic_two_block_coords", "rotated_connection", "box3d"])
def test_step_gradient_matches_oracle_adjoint(name):
    """Gradient of one differentiable GPU step (PISOtorch_diff graph: stage VJP kernels, adjoint solves, SDDMM)
    against the oracle's reverse pass of the same step (oracle.piso_step_vjp), FP64, w.r.t. the input velocities,
    the velocity sources and the domain viscosity."""
    from oracle import oracle as O
    print("case", name, flush=True)
    d = _setup(name)
    c = O.cpu_twin(d)
    v0, s0 = _state(d)
    g = torch.Generator().manual_seed(31)
    gout = [((torch.rand(v.shape, generator=g, dtype=torch.float64) * 2 - 1)).cuda() for v in v0]
    vs = [v.clone().requires_grad_(True) for v in v0]
    ss = [s.clone().requires_grad_(True) for s in s0]
    visc = d.viscosity.clone().requires_grad_(True)
    for b, v, s in zip(d.blocks, vs, ss):
        b.setVelocity(v)
        b.setVelocitySource(s)
    d.setViscosity(visc)
    d.UpdateDomainData()
    assert _sim(d, True)._PISO_split_step(1)
    L = sum((b.velocity * go).sum() for b, go in zip(d.blocks, gout))
    L.backward()
    gu, gs, gv = O.piso_step_vjp(c, DT, [x.cpu() for x in gout], corrector_steps=2)
    for a, b in zip(vs, gu):
        e = float((a.grad.cpu() - b).abs().max() / b.abs().max())
        print("  velocity grad rel err %.3e" % e, flush=True)
        assert e < 1e-8
    for a, b in zip(ss, gs):
        e = float((a.grad.cpu() - b).abs().max() / b.abs().max())
        print("  source grad rel err %.3e" % e, flush=True)
        assert e < 1e-8
    if not all(b.viscosity is not None for b in d.blocks):
        e = abs(float(visc.grad[0]) - gv) / abs(gv)
        print("  viscosity grad rel err %.3e" % e, flush=True)
        assert e < 1e-8
