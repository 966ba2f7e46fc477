"""Per-step checkpointing of differentiable PISO rollouts.

A rollout of K steps through the PISOtorch_diff backend keeps every stage's saved tensors alive until backward
(matrices, right-hand sides, solver results of all K steps).  `checkpointed_step` instead runs each step's forward
with the plain PISOtorch backend (no autograd bookkeeping), keeps only the step's input velocities, and re-runs that
one step through PISOtorch_diff when its gradient is needed, so that memory is one step's graph regardless of K.
The recomputed step is bit-identical to the forward one: every kernel, including the Krylov reductions, is
deterministic.
"""
import torch


class _PISOStep(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sim, n_blocks, *tensors):
        """tensors = block velocities (n_blocks) followed by the parameter tensors already set on the domain."""
        ctx.sim = sim
        ctx.n_blocks = n_blocks
        ctx.save_for_backward(*tensors)
        domain = sim.domain
        was_diff = sim.differentiable
        reuse = (sim.advect_reuse_result, sim.pressure_reuse_result)
        with torch.no_grad():
            for b, v in zip(domain.getBlocks(), tensors[:n_blocks]):
                b.setVelocity(v.detach().clone())
                # the plain backend writes results into the block tensors in place: never into a tensor that an
                # earlier step returned
                b.setPressure(b.pressure.detach().clone())
            domain.UpdateDomainData()
            # the differentiable recomputation starts every solve from zero; so does this forward
            sim.differentiable = False
            sim.advect_reuse_result = sim.pressure_reuse_result = False
            try:
                ok = sim._PISO_split_step(1)
            finally:
                sim.differentiable = was_diff
                sim.advect_reuse_result, sim.pressure_reuse_result = reuse
        if not ok:
            raise RuntimeError("PISO step failed (linear solve did not converge)")
        outs = [b.velocity.detach().clone() for b in domain.getBlocks()]
        outs += [b.pressure.detach().clone() for b in domain.getBlocks()]
        return tuple(outs)

    @staticmethod
    def backward(ctx, *grads):
        sim, n = ctx.sim, ctx.n_blocks
        saved = ctx.saved_tensors
        domain = sim.domain
        vel_in = [t.detach().clone().requires_grad_(True) for t in saved[:n]]
        params = list(saved[n:])
        leaves = [p.detach().requires_grad_(p.requires_grad) for p in params]
        swaps = _swap_params(domain, params, leaves)
        counters = (sim.total_step, sim.total_time)
        # the recomputation runs on the live domain: its block state is put back afterwards
        state = [(b.velocity, b.pressure) for b in domain.getBlocks()]
        was_diff = sim.differentiable
        try:
            with torch.enable_grad():
                for b, v in zip(domain.getBlocks(), vel_in):
                    b.setVelocity(v)
                domain.UpdateDomainData()
                sim.differentiable = True
                if not sim._PISO_split_step(1):
                    raise RuntimeError("PISO step recomputation failed")
                outs = [b.velocity for b in domain.getBlocks()] + [b.pressure for b in domain.getBlocks()]
                pairs = [(o, g) for o, g in zip(outs, grads) if g is not None and o.requires_grad]
                wrt = vel_in + [x for x in leaves if x.requires_grad]
                got = torch.autograd.grad([o for o, _ in pairs], wrt, [g for _, g in pairs], allow_unused=True)
        finally:
            sim.differentiable = was_diff
            sim.total_step, sim.total_time = counters
            for b, (v, p) in zip(domain.getBlocks(), state):
                b.setVelocity(v)
                b.setPressure(p)
            _restore_params(domain, swaps)
        got = list(got)
        g_vel = got[:n]
        g_par = []
        it = iter(got[n:])
        for x in leaves:
            g_par.append(next(it) if x.requires_grad else None)
        return (None, None, *g_vel, *g_par)


def _param_slots(domain):
    """Where a parameter tensor may live on the domain: block velocity sources and viscosities, the domain
    viscosity and fixed-boundary velocities."""
    for b in domain.getBlocks():
        yield (lambda b=b: b.velocitySource, b.setVelocitySource)
        yield (lambda b=b: b.viscosity, b.setViscosity)
        for _, fb in b.getFixedBoundaries():
            yield (lambda fb=fb: fb.velocity, fb.setVelocity)
    yield (lambda: domain.viscosity, domain.setViscosity)


def _swap_params(domain, params, leaves):
    swaps = []
    for get, set_ in _param_slots(domain):
        cur = get()
        for p, leaf in zip(params, leaves):
            if cur is not None and cur is p:
                set_(leaf)
                swaps.append((set_, p))
    return swaps


def _restore_params(domain, swaps):
    for set_, p in swaps:
        set_(p)
    domain.UpdateDomainData()


def checkpointed_step(sim, params=()):
    """One PISO step of `sim` whose autograd graph is recomputed in backward.  `params` are tensors already set on
    the domain (e.g. a learnable block.velocitySource) whose gradients are wanted.  The block velocities on entry
    are the step's differentiable input; on return the blocks hold the step's output velocities and pressures,
    which carry the gradient path."""
    domain = sim.domain
    blocks = domain.getBlocks()
    outs = _PISOStep.apply(sim, len(blocks), *[b.velocity for b in blocks], *params)
    n = len(blocks)
    for b, v, p in zip(blocks, outs[:n], outs[n:]):
        b.setVelocity(v)
        b.setPressure(p)
    domain.UpdateDomainData()
    return outs[:n]


def rollout(sim, steps, params=(), checkpoint=True):
    """`steps` differentiable PISO steps; with checkpoint=False the whole rollout graph is kept (PISOtorch_diff)."""
    if checkpoint:
        out = None
        for _ in range(steps):
            out = checkpointed_step(sim, params)
        return out
    was_diff = sim.differentiable
    sim.differentiable = True
    try:
        if not sim._PISO_split_step(steps):
            raise RuntimeError("PISO rollout failed (linear solve did not converge)")
    finally:
        sim.differentiable = was_diff
    return [b.velocity for b in sim.domain.getBlocks()]
