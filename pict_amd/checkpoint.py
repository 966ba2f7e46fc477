"""Per-step checkpointing of differentiable PISO rollouts.

A rollout of K steps through the PISOtorch_diff backend keeps every stage's saved tensors alive until backward
(matrices, right-hand sides, solver results of all K steps).  `checkpointed_step` instead keeps only each step's
input velocities and recomputes the step when its gradient is needed, so that memory is one step's intermediates
regardless of K.

Two implementations of the same contract:
- native: when the step is what pict_piso_step runs (orthogonal, no hooks, ...), the forward is one library call;
  the backward re-runs it recording its intermediates on a tape (pict_step_tape) and runs the reverse pass as one
  more library call (pict_piso_step_vjp).  With keep_tape=True the tape of the forward is kept instead (no
  recomputation; memory of one tape per step).
- general: the forward runs the plain PISOtorch backend, the backward re-runs the step through PISOtorch_diff and
  lets autograd differentiate it.
Every kernel is deterministic, so a recomputed step is bit-identical to its forward.
"""
import ctypes
import weakref

import torch

from . import PISOtorch, _abi


# ----------------------------------------------------------------------------
# parameters a checkpointed step may differentiate: where they live on the domain and where their gradient goes
# ----------------------------------------------------------------------------

def _param_slots(domain):
    """(getter, setter, gradient getter) of every tensor a parameter may be: block velocity sources and
    viscosities, fixed-boundary velocities and the domain viscosity."""
    for b in domain.getBlocks():
        yield (lambda b=b: b.velocitySource, b.setVelocitySource, lambda b=b: b.velocitySourceGrad)
        yield (lambda b=b: b.viscosity, b.setViscosity, lambda b=b: b.viscosityGrad)
        for _, fb in b.getFixedBoundaries():
            yield (lambda fb=fb: fb.velocity, fb.setVelocity, lambda fb=fb: fb.velocityGrad)
    yield (lambda: domain.viscosity, domain.setViscosity, lambda: domain.viscosityGrad)


def _same(a, b):
    return a is b or (a is not None and b is not None and a.data_ptr() == b.data_ptr() and a.shape == b.shape
                      and a.device == b.device)


def _slots_of(domain, params):
    """Index into _param_slots(domain) of the slot each parameter occupies now (None: not a domain parameter).
    Recorded when a step runs: a rollout may give every step other parameter tensors (a velocity source computed
    from the velocities, say), and the step's backward must see the ones it ran with."""
    slots = [get() for get, _, _ in _param_slots(domain)]
    out = []
    for p in params:
        out.append(next((i for i, cur in enumerate(slots) if _same(cur, p)), None))
    return out


def _install_params(domain, tensors, slot_idx):
    """Put each tensor into its recorded slot; returns what to put back (_restore_params)."""
    slots = list(_param_slots(domain))
    swaps = []
    for t, i in zip(tensors, slot_idx):
        if i is None:
            continue
        get, set_, _ = slots[i]
        cur = get()
        if cur is not t:
            set_(t)
            swaps.append((set_, cur))
    if swaps:
        domain.UpdateDomainData()
    return swaps


def _restore_params(domain, swaps):
    for set_, p in reversed(swaps):
        set_(p)
    domain.UpdateDomainData()


def _param_grads(domain, params, slot_idx):
    """The accumulated gradient of each parameter, from its recorded slot (None where it is not a domain
    parameter)."""
    slots = list(_param_slots(domain))
    out = []
    for p, i in zip(params, slot_idx):
        g = None
        if i is not None:
            gt = slots[i][2]()
            if gt is not None:
                g = gt.detach().clone().to(p.device)
        out.append(g)
    return out


# ----------------------------------------------------------------------------
# native tape and reverse pass
# ----------------------------------------------------------------------------

class _Tape:
    """Device buffers of one pict_step_tape."""

    def __init__(self, domain, K):
        n, nnz, D = domain.getTotalSize(), domain.C.getSize(), domain.getSpatialDims()
        opts = dict(dtype=domain.getDtype(), device=domain.getDevice())
        self.tensors = [torch.empty(nnz, **opts), torch.empty(n, **opts), torch.empty(nnz, **opts),
                        torch.empty(D * n, **opts), torch.empty(max(K, 1) * D * n, **opts),
                        torch.empty(max(K, 1) * n, **opts), torch.empty(max(K, 1) * D * n, **opts)]
        self.desc = _abi.StepTape(*[t.data_ptr() for t in self.tensors])


def _grad_tensors(domain):
    """Every gradient tensor the reverse pass reads or writes (None where one is missing)."""
    ts = [domain.A_grad, domain.C_grad.value if domain.C_grad is not None else None,
          domain.P_grad.value if domain.P_grad is not None else None, domain.velocityRHS_grad,
          domain.velocityResult_grad, domain.pressureRHS_grad, domain.pressureRHSdiv_grad,
          domain.pressureResult_grad, domain.viscosity_grad]
    for b in domain.getBlocks():
        ts += [b._velocity_grad, b._pressure_grad]
    return ts + _param_grad_tensors(domain)


def _param_grad_tensors(domain):
    """The parameter gradients the reverse pass accumulates into (block sources and viscosities, fixed-boundary
    velocities, the domain viscosity)."""
    ts = [domain.viscosity_grad]
    for b in domain.getBlocks():
        ts += [b.velocitySource_grad, b._viscosity_grad]
        ts += [fb._velocity_grad for _, fb in b.getFixedBoundaries()]
    return [t for t in ts if t is not None]


def _grad_key(domain):
    ts = _grad_tensors(domain)
    if any(t is None for t in ts[:9 + 2 * len(domain.getBlocks())]):
        return None
    return tuple(t.data_ptr() for t in ts)


_GRAD_KEYS = weakref.WeakKeyDictionary()  # domain -> addresses of the gradient tensors created here


_RECOMPUTE_TAPES = weakref.WeakKeyDictionary()  # domain -> tape reused by recomputing backward passes


def _recompute_tape(domain, K):
    """The tape a recomputing backward pass records into: filled and consumed within one backward call, so one per
    domain is reused."""
    t = _RECOMPUTE_TAPES.get(domain)
    key = (K, domain.getTotalSize(), domain.C.getSize(), domain.getSpatialDims(), domain.getDtype(), domain.getDevice())
    if t is None or t[0] != key:
        t = (key, _Tape(domain, K))
        _RECOMPUTE_TAPES[domain] = t
    return t[1]


def _ensure_grad_tensors(domain):
    """Gradient tensors the reverse pass writes.  They are created once per domain and reused: the reverse pass
    overwrites every intermediate gradient before it reads it, so only the parameter gradients, which its kernels
    accumulate into, are zeroed again."""
    key = _grad_key(domain)
    if key is not None and _GRAD_KEYS.get(domain) == key:
        for t in _param_grad_tensors(domain):
            t.zero_()
        return
    _create_grad_tensors(domain)
    _GRAD_KEYS[domain] = _grad_key(domain)


def _create_grad_tensors(domain):
    domain.CreateAGrad()
    domain.CGrad.CreateValue()
    domain.PGrad.CreateValue()
    domain.CreateVelocityRHSGrad()
    domain.CreateVelocityResultGrad()
    domain.CreatePressureRHSGrad()
    domain.CreatePressureRHSdivGrad()
    domain.CreatePressureResultGrad()
    domain.CreateVelocityGradOnBlocks()
    domain.CreatePressureGradOnBlocks()
    domain.CreateVelocitySourceGradOnBlocks()
    domain.CreateViscosityGrad()
    domain.CreateVelocityGradOnBoundaries()


_GRAD_VECTORS = weakref.WeakKeyDictionary()  # domain -> the global gradient vector of its reverse passes


def _to_global(tensors, D, domain=None):
    """Per-block [1,D,...] tensors -> global component-major vector (the domain's persistent one if given: its
    address is part of what a captured reverse pass replays)."""
    parts = [t[0, c].reshape(-1) for c in range(D) for t in tensors]
    if domain is None:
        return torch.cat(parts).contiguous()
    n = D * domain.getTotalSize()
    g = _GRAD_VECTORS.get(domain)
    if g is None or g.numel() != n or g.dtype != tensors[0].dtype or g.device != tensors[0].device:
        g = torch.empty(n, dtype=tensors[0].dtype, device=tensors[0].device)
        _GRAD_VECTORS[domain] = g
    torch.cat(parts, out=g)
    return g


def _from_global(v, like, D):
    out = [torch.empty_like(t) for t in like]
    n = sum(t[0, 0].numel() for t in like)
    for c in range(D):
        off = c * n
        for o in out:
            k = o[0, c].numel()
            o[0, c].copy_(v[off:off + k].view_as(o[0, c]))
            off += k
    return out


class _Work:
    """Private block velocity and pressure tensors a domain steps in place during native checkpointed rollouts.
    Installed once, their addresses stay in the domain's descriptor, so a step needs no repacking; step inputs
    and outputs are copied in and out.  `last` holds the outputs (autograd tensors) of the last native step whose
    values the buffers hold."""

    def __init__(self, domain):
        blocks = domain.getBlocks()
        self.vel = [torch.empty_like(b.velocity, requires_grad=False) for b in blocks]
        self.pres = [torch.empty_like(b.pressure, requires_grad=False) for b in blocks]
        self.last = None

    def matches(self, domain):
        return all(w.shape == b.velocity.shape and w.dtype == b.velocity.dtype and w.device == b.velocity.device
                   for w, b in zip(self.vel, domain.getBlocks())) and len(self.vel) == len(domain.getBlocks())

    def installed(self, domain):
        return all(b.velocity is v and b.pressure is p for b, v, p in zip(domain.getBlocks(), self.vel, self.pres))

    def install(self, domain):
        """Make the domain step in these buffers (their values are set by the caller)."""
        for b, v, p in zip(domain.getBlocks(), self.vel, self.pres):
            b.setVelocity(v)
            b.setPressure(p)
        domain.UpdateDomainData()
        self.last = None


_WORK = weakref.WeakKeyDictionary()  # domain -> _Work


def _work(domain):
    w = _WORK.get(domain)
    if w is None or not w.matches(domain):
        w = _Work(domain)
        _WORK[domain] = w
    return w


def _same_storage(a, b):
    return a is b or (a.data_ptr() == b.data_ptr() and a.shape == b.shape and a.dtype == b.dtype)


def _native_forward(sim, tape):
    """One native step (recording `tape`)."""
    return sim._native_steps(1, sim._get_time_step_torch(), tape=tape.desc if tape is not None else None)


def _native_backward(sim, tape, vel_grads):
    """Reverse pass of the step recorded on `tape`: gradient w.r.t. the input block velocities; parameter
    gradients are left on the domain's gradient tensors."""
    domain = sim.domain
    D = domain.getSpatialDims()
    _ensure_grad_tensors(domain)
    domain.UpdateDomainData()
    g = _to_global(vel_grads, D, domain)
    o = sim._native_opts(sim._get_time_step_torch())
    K = o.corrector_steps
    infos = (_abi.SolveInfo * (D + K))()
    lib = _abi.load_library()
    timing = PISOtorch.solve_log is not None
    if sim.defer_solver_checks and not timing:
        sim._defer_check(lib.pict_piso_step_vjp_async, o, True, domain._atlas, ctypes.byref(o),
                         ctypes.byref(tape.desc), ctypes.c_void_p(g.data_ptr()))
        return g
    if timing:
        _abi.check(lib.pict_set_step_timing(1))
    try:
        _abi.check(lib.pict_piso_step_vjp(domain._atlas, ctypes.byref(o), ctypes.byref(tape.desc),
                                          ctypes.c_void_p(g.data_ptr()), infos,
                                          PISOtorch._stream(domain.getDevice())), "PISO step reverse pass")
        kms = None
        if timing:
            ms = (ctypes.c_double * (1 + K))()
            lib.pict_last_step_solve_ms(ms, 1 + K)
            kms = list(ms)
    finally:
        if timing:
            _abi.check(lib.pict_set_step_timing(0))
    sim._native_record(infos, o, kms, adjoint=True)
    return g


class _PISOStep(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sim, n_blocks, keep_tape, *tensors):
        """tensors = block velocities (n_blocks) followed by the parameter tensors already set on the domain."""
        ctx.set_materialize_grads(False)
        ctx.sim = sim
        ctx.n_blocks = n_blocks
        ctx.save_for_backward(*tensors)
        ctx.slot_idx = _slots_of(sim.domain, tensors[n_blocks:])
        domain = sim.domain
        was_diff = sim.differentiable
        reuse = (sim.advect_reuse_result, sim.pressure_reuse_result)
        vin = tensors[:n_blocks]
        with torch.no_grad():
            # the differentiable recomputation starts every solve from zero; so does this forward
            sim.differentiable = False
            sim.advect_reuse_result = sim.pressure_reuse_result = False
            try:
                ctx.native = sim._native_step_eligible()
                ctx.tape = None
                if ctx.native:
                    # step in the domain's private buffers: the native step does not read the pressure, and the
                    # velocity is copied in unless the buffers already hold this input (the previous step's output)
                    w = _work(domain)
                    held = w.installed(domain) and w.last is not None and all(
                        _same_storage(v, l) for v, l in zip(vin, w.last))
                    if not w.installed(domain):
                        w.install(domain)
                    if not held:
                        for wv, v in zip(w.vel, vin):
                            wv.copy_(v)
                    ctx.tape = _Tape(domain, sim.corrector_steps) if keep_tape else None
                    ok = _native_forward(sim, ctx.tape)
                    outs = [t.clone() for t in w.vel + w.pres]
                else:
                    for b, v in zip(domain.getBlocks(), vin):
                        b.setVelocity(v.detach().clone())
                        # the forward writes results into the block tensors in place: never into a tensor that an
                        # earlier step returned
                        b.setPressure(b.pressure.detach().clone())
                    domain.UpdateDomainData()
                    ok = sim._PISO_split_step(1)
                    outs = [b.velocity.detach().clone() for b in domain.getBlocks()]
                    outs += [b.pressure.detach().clone() for b in domain.getBlocks()]
            finally:
                sim.differentiable = was_diff
                sim.advect_reuse_result, sim.pressure_reuse_result = reuse
        if not ok:
            raise RuntimeError("PISO step failed (linear solve did not converge)")
        sim._checkpoint_native = ctx.native
        return tuple(outs)

    @staticmethod
    def backward(ctx, *grads):
        n = ctx.n_blocks
        pressure_grads = [g for g in grads[n:] if g is not None and bool((g != 0).any())]
        if ctx.native and not pressure_grads:
            return (None, None, None, *_PISOStep._backward_native(ctx, grads[:n]))
        return (None, None, None, *_PISOStep._backward_autograd(ctx, grads))

    @staticmethod
    def _backward_native(ctx, vel_grads):
        sim, n = ctx.sim, ctx.n_blocks
        domain = sim.domain
        saved = ctx.saved_tensors
        params = list(saved[n:])
        vel_grads = [torch.zeros_like(v) if g is None else g.contiguous() for g, v in zip(vel_grads, saved[:n])]
        counters = (sim.total_step, sim.total_time)
        tape = ctx.tape
        # the recomputation and the reverse pass overwrite the block velocities and pressures: they run in the
        # private buffers, whose values (or the domain's own tensors) are put back afterwards
        w = _work(domain)
        state = None if w.installed(domain) else [(b.velocity, b.pressure) for b in domain.getBlocks()]
        last = w.last
        swaps = []
        try:
            with torch.no_grad():
                # the parameters this step ran with (a later step may have set others)
                swaps = _install_params(domain, [p.detach() for p in params], ctx.slot_idx)
                snapshot = [t.clone() for t in w.vel + w.pres] if state is None else None
                if state is not None:
                    w.install(domain)
                if tape is None:
                    # recompute the step from its input, recording the tape
                    tape = _recompute_tape(domain, sim.corrector_steps)
                    for wv, v in zip(w.vel, saved[:n]):
                        wv.copy_(v)
                    reuse = (sim.advect_reuse_result, sim.pressure_reuse_result)
                    sim.advect_reuse_result = sim.pressure_reuse_result = False
                    try:
                        _native_forward(sim, tape)
                    finally:
                        sim.advect_reuse_result, sim.pressure_reuse_result = reuse
                g = _native_backward(sim, tape, vel_grads)
                g_vel = _from_global(g, list(saved[:n]), domain.getSpatialDims())
                g_par = [gp if p.requires_grad else None
                         for gp, p in zip(_param_grads(domain, params, ctx.slot_idx), params)]
        finally:
            sim.total_step, sim.total_time = counters
            with torch.no_grad():
                if swaps:
                    _restore_params(domain, swaps)
                if state is None:
                    for t, v in zip(w.vel + w.pres, snapshot):
                        t.copy_(v)
                    w.last = last
                else:
                    for b, (v, p) in zip(domain.getBlocks(), state):
                        b.setVelocity(v)
                        b.setPressure(p)
                    domain.UpdateDomainData()
        ctx.tape = None
        return (*g_vel, *g_par)

    @staticmethod
    def _backward_autograd(ctx, grads):
        sim, n = ctx.sim, ctx.n_blocks
        saved = ctx.saved_tensors
        domain = sim.domain
        vel_in = [t.detach().clone().requires_grad_(True) for t in saved[:n]]
        params = list(saved[n:])
        leaves = [p.detach().requires_grad_(p.requires_grad) for p in params]
        swaps = _install_params(domain, leaves, ctx.slot_idx)
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
        return (*g_vel, *g_par)


def checkpointed_step(sim, params=(), keep_tape=False):
    """One PISO step of `sim` whose gradient is computed by recomputing the step in backward (keep_tape=True on
    the native path: by keeping its tape).  `params` are tensors already set on the domain (e.g. a learnable
    block.velocitySource) whose gradients are wanted.  The block velocities on entry are the step's
    differentiable input (after a native checkpointed step: that step's returned outputs).  Returns the output
    velocities, which carry the gradient path.  On return the blocks hold the step's output velocities and
    pressures: as those tensors themselves on the stage-by-stage path, as their values in private buffers the
    domain keeps stepping in on the native path."""
    domain = sim.domain
    blocks = domain.getBlocks()
    n = len(blocks)
    w = _WORK.get(domain)
    if w is not None and w.last is not None and w.installed(domain):
        vin = w.last  # the domain holds the values of the previous native step's outputs
    else:
        vin = [b.velocity for b in blocks]
    outs = _PISOStep.apply(sim, n, keep_tape, *vin, *params)
    if sim._checkpoint_native:
        # the domain keeps stepping in its private buffers, which hold these outputs' values
        _WORK[domain].last = list(outs[:n])
    else:
        for b, v, p in zip(blocks, outs[:n], outs[n:]):
            b.setVelocity(v)
            b.setPressure(p)
        domain.UpdateDomainData()
    return outs[:n]


def rollout(sim, steps, params=(), checkpoint=True, keep_tape=False):
    """`steps` differentiable PISO steps; with checkpoint=False the whole rollout graph is kept (PISOtorch_diff)."""
    if checkpoint:
        out = None
        for _ in range(steps):
            out = checkpointed_step(sim, params, keep_tape)
        return out
    was_diff = sim.differentiable
    sim.differentiable = True
    try:
        if not sim._PISO_split_step(steps):
            raise RuntimeError("PISO rollout failed (linear solve did not converge)")
    finally:
        sim.differentiable = was_diff
    return [b.velocity for b in sim.domain.getBlocks()]


def rollout_with_sources(sim, steps, source_fn, keep_tape=False):
    """A checkpointed rollout whose block velocity sources are recomputed before every step from the current
    velocities: sources = source_fn(velocities) (per block, differentiable: e.g. a static forcing plus a CNN
    sub-grid model).  Each step's gradient reaches its sources, and through them whatever source_fn depends on
    (its parameters and the velocities it read).  Returns the final velocities."""
    domain = sim.domain
    blocks = domain.getBlocks()
    w = _WORK.get(domain)
    u = w.last if (w is not None and w.last is not None and w.installed(domain)) else [b.velocity for b in blocks]
    for _ in range(steps):
        srcs = [s.contiguous() for s in source_fn(u)]
        for b, s in zip(blocks, srcs):
            b.setVelocitySource(s)
        domain.UpdateDomainData()
        u = checkpointed_step(sim, srcs, keep_tape)
    return u
