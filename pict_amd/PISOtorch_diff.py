"""PISOtorch_diff -- differentiable counterpart of PISOtorch.

Mirrors the reference's autograd layer (tum-pbs/PICT PISOtorch_diff.py): the
same function names as PISOtorch, the solver-wrapper contract
(`_linear_solve_wrapper`, :245-293: zero-RHS shortcut, LinsolveError on
non-convergence, FP64 retry) and `linear_solve_GPU` (:295-374) whose backward
solves the adjoint system on A^T and forms dA = -lambda x^T on the pattern.
"""
import logging

import torch

from . import PISOtorch

_LOG = logging.getLogger("PISOtorch_diff")


def _get_solver_tolerance(tol=None, dtype=torch.float32):
    """PISOtorch_diff.py:183-189."""
    if tol is None:
        if dtype == torch.float64:
            tol = 1e-8
        elif dtype == torch.float32:
            tol = 1e-5
    return tol


def _get_solver_tolerance_torch(tol=None, dtype=torch.float32):
    return torch.tensor([_get_solver_tolerance(tol, dtype)], dtype=dtype)


class LinsolveError(RuntimeError):
    pass


def _check_solver_return_infos(solver_infos, transpose, use_BiCG, tol, max_iter, return_best_result, is_FWD=None,
                               debug_out=False):
    """PISOtorch_diff.py:199-243."""
    fwd_str = "" if is_FWD is None else ("FWD, " if is_FWD else "BWD, ")
    if any(not i.isFiniteResidual for i in solver_infos):
        bad = [k for k, i in enumerate(solver_infos) if not i.isFiniteResidual]
        raise LinsolveError("Linear solve (%sT:%s, BiCG:%s, tol:%.03e) reported non-finite residual for RHS %s of %d. "
                            "Solver infos:\n%s." % (fwd_str, transpose, use_BiCG, tol, bad, len(solver_infos),
                                                    "\n".join("\tRHS %d: %s" % (k, i)
                                                              for k, i in enumerate(solver_infos))))
    if any(not i.converged for i in solver_infos):
        bad = [k for k, i in enumerate(solver_infos) if not i.converged]
        msg = "Linear solve (%sT:%s, BiCG:%s, tol:%.03e) RHS %s of %d did not converge after %d iterations." % (
            fwd_str, transpose, use_BiCG, tol, bad, len(solver_infos), max_iter)
        if return_best_result:
            _LOG.warning(msg)
        else:
            raise LinsolveError(msg + " Solver infos:\n%s" % "\n".join(
                "\tRHS %d: %s" % (k, i) for k, i in enumerate(solver_infos)))


def _linear_solve_wrapper(csrMat, rhs, result, maxit_torch, tol_torch, convergence_criterion, use_BiCG,
                          matrix_rank_deficient, residual_reset_step, transpose, print_residual, return_best_result,
                          is_FWD=None, debug_out=False, double_fallback=False, BiCG_with_preconditioner=True,
                          BiCG_precondition_fallback=False):
    """PISOtorch_diff.py:245-293.  The reference's host-side `rhs.eq(0).all()` test (a device->host sync) is
    done inside the solver kernel (zero_rhs_shortcut): an all-zero RHS yields result = 0 and no error."""
    tol = float(tol_torch[0])
    maxit = int(maxit_torch[0])
    opts = PISOtorch._solve_opts(maxit, tol, convergence_criterion, use_BiCG, matrix_rank_deficient,
                                 residual_reset_step, transpose, print_residual, return_best_result,
                                 BiCG_with_preconditioner, zero_rhs_shortcut=True)
    infos = PISOtorch._solve(csrMat, rhs, result, opts)

    def not_solved(sinfos):
        return (not return_best_result and any(not i.converged for i in sinfos)) or \
               (return_best_result and any(not i.isFiniteResidual for i in sinfos))

    if double_fallback and rhs.dtype == torch.float32 and not_solved(infos):
        _LOG.warning("Single precision solve failed, trying double precision.")
        dp = torch.float64
        result_dp = torch.zeros_like(result, dtype=dp)
        opts.tolerance = tol
        infos = PISOtorch._solve(csrMat.toType(dp), rhs.to(dp), result_dp, opts)
        result.copy_(result_dp.to(result.dtype))
    if BiCG_precondition_fallback and use_BiCG and not BiCG_with_preconditioner and not_solved(infos):
        _LOG.warning("Not preconditioned BiCG solve failed, trying preconditioned.")
        _LOG.info("Solver infos:\n%s", [str(i) for i in infos])
        result.zero_()  # may hold non-finite values after the failed solve
        opts.bicg_with_preconditioner = 1
        infos = PISOtorch._solve(csrMat, rhs, result, opts)
    _check_solver_return_infos(infos, transpose, use_BiCG, tol, maxit, return_best_result, is_FWD=is_FWD,
                               debug_out=debug_out)
    return infos


class _LinearSolveFunction(torch.autograd.Function):
    """x = A^-1 b (PISOtorch_diff.py:316-372); `cfg` = (A detached, maxit, tol, criterion, use_BiCG, transpose,
    return_best_result, double_fallback, BiCG_with_preconditioner, BiCG_precondition_fallback)."""

    @staticmethod
    def forward(ctx, cfg, A_val, b):
        A, maxit_torch, tol_torch, conv, use_BiCG, transpose, best, dfb, pre, pfb = cfg
        x = torch.zeros_like(b)
        _linear_solve_wrapper(A, b.contiguous(), x, maxit_torch, tol_torch, conv, use_BiCG, False, 0, transpose,
                              False, best, is_FWD=True, double_fallback=dfb, BiCG_with_preconditioner=pre,
                              BiCG_precondition_fallback=pfb)
        ctx.cfg = cfg
        ctx.save_for_backward(A_val, x)
        return x

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, grad_x):
        A, maxit_torch, tol_torch, conv, use_BiCG, transpose, best, dfb, pre, pfb = ctx.cfg
        grad_b = None
        grad_A_val = None
        if ctx.needs_input_grad[2] or ctx.needs_input_grad[1]:
            _, x = ctx.saved_tensors
            grad_b = torch.zeros_like(grad_x)
            _linear_solve_wrapper(A, grad_x.contiguous(), grad_b, maxit_torch, tol_torch, conv, use_BiCG, False, 0,
                                  not transpose, False, best, is_FWD=False, double_fallback=dfb,
                                  BiCG_with_preconditioner=pre, BiCG_precondition_fallback=pfb)
            if ctx.needs_input_grad[1]:
                grad_A = A.WithZeroValue()
                grad_A_val = grad_A.value
                n = grad_A.getRows()
                if n == grad_b.size(0):
                    PISOtorch.SparseOuterProduct(grad_b, x, grad_A)
                else:
                    for gb, xs in zip(torch.split(grad_b, n), torch.split(x, n)):
                        part = grad_A.WithZeroValue()
                        PISOtorch.SparseOuterProduct(gb.contiguous(), xs.contiguous(), part)
                        grad_A_val += part.value
                grad_A_val *= -1
        return None, grad_A_val, grad_b


def linear_solve_GPU(csrMat, rhs, transpose=False, use_BiCG=False, tol=None, max_iter=5000, return_best_result=False,
                     double_fallback=False, BiCG_with_preconditioner=True, BiCG_precondition_fallback=False):
    """PISOtorch_diff.py:295-374: x = A^-1 b with gradients for b (adjoint solve on A^T) and A (-lambda x^T on
    the pattern, summed over batched right-hand sides)."""
    # a detached copy of the values (later stages may overwrite csrMat.value in place); the pattern is immutable
    A = PISOtorch.CSRmatrix(csrMat.value.detach().clone(), csrMat.index, csrMat.row)
    cfg = (A, torch.IntTensor([max_iter]), _get_solver_tolerance_torch(tol, rhs.dtype),
           PISOtorch.ConvergenceCriterion.NORM2_NORMALIZED, use_BiCG, transpose, return_best_result, double_fallback,
           BiCG_with_preconditioner, BiCG_precondition_fallback)
    return _LinearSolveFunction.apply(cfg, csrMat.value, rhs)


# ----------------------------------------------------------------------------
# gradient bookkeeping (PISOtorch_diff.py:13-181)
# ----------------------------------------------------------------------------

_EXCLUDED_GRADIENTS = {}


def exclude_gradient(function_name, grad_name, is_gradient_input):
    """Zero a gradient input (is_gradient_input) or drop a gradient output of a stage (PISOtorch_diff.py:14-20)."""
    if function_name not in _EXCLUDED_GRADIENTS:
        raise ValueError("Function '%s' does not exist" % (function_name,))
    if not grad_name.endswith("_GRAD"):
        raise ValueError("Only gradients can be excluded (name should end with '_GRAD').")
    _EXCLUDED_GRADIENTS[function_name][0 if is_gradient_input else 1].add(grad_name)


def clear_excluded_gradients():
    for inputs, outputs in _EXCLUDED_GRADIENTS.values():
        inputs.clear()
        outputs.clear()


def _set_gradient_input(var_grad, grad_setter, function_name, grad_name):
    if grad_name in _EXCLUDED_GRADIENTS[function_name][0]:
        grad_setter(torch.zeros_like(var_grad))
    else:
        # the backward kernels may accumulate into their gradient inputs: never hand them autograd's own buffer
        grad_setter(var_grad.contiguous().clone())


def check_non_ortho_rhs(non_ortho_flags):
    """NON_ORTHO_DIRECT_RHS (2) or NON_ORTHO_DIAGONAL_RHS (8) set (PISOtorch_diff.py:34-42)."""
    return (non_ortho_flags & 2) > 0 or (non_ortho_flags & 8) > 0


def _domain_entries(domain):
    """(section, name, getter, setter, clear, on_cpu) for every tensor flatten_domain knows, in its order
    (PISOtorch_diff.py:44-118)."""
    d = domain
    yield ("D", "C", lambda: d.C.value, d.C.setValue, None, False)
    yield ("D", "C_GRAD", lambda: d.CGrad.value, d.CGrad.setValue, None, False)
    yield ("D", "A", lambda: d.A, d.setA, None, False)
    yield ("D", "A_GRAD", lambda: d.AGrad, d.setAGrad, None, False)
    yield ("D", "P", lambda: d.P.value, d.P.setValue, None, False)
    yield ("D", "P_GRAD", lambda: d.PGrad.value, d.PGrad.setValue, None, False)
    yield ("D", "VISCOSITY", lambda: d.viscosity, d.setViscosity, None, False)
    yield ("D", "VISCOSITY_GRAD", lambda: d.viscosityGrad, None, None, True)
    yield ("D", "PASSIVE_SCALAR_VISCOSITY", lambda: d.passiveScalarViscosity, d.setScalarViscosity,
           d.clearScalarViscosity, False)
    yield ("D", "PASSIVE_SCALAR_VISCOSITY_GRAD", lambda: d.passiveScalarViscosityGrad,
           d.setPassiveScalarViscosityGrad, None, False)
    yield ("D", "PASSIVE_SCALAR_RESULT", lambda: d.scalarResult, d.setScalarResult, None, False)
    yield ("D", "PASSIVE_SCALAR_RESULT_GRAD", lambda: d.scalarResultGrad, d.setScalarResultGrad, None, False)
    yield ("D", "VELOCITY_RESULT", lambda: d.velocityResult, d.setVelocityResult, None, False)
    yield ("D", "VELOCITY_RESULT_GRAD", lambda: d.velocityResultGrad, d.setVelocityResultGrad, None, False)
    yield ("D", "PRESSURE_RESULT", lambda: d.pressureResult, d.setPressureResult, None, False)
    yield ("D", "PRESSURE_RESULT_GRAD", lambda: d.pressureResultGrad, d.setPressureResultGrad, None, False)
    yield ("D", "PASSIVE_SCALAR_RHS", lambda: d.scalarRHS, d.setScalarRHS, None, False)
    yield ("D", "PASSIVE_SCALAR_RHS_GRAD", lambda: d.scalarRHSGrad, d.setScalarRHSGrad, None, False)
    yield ("D", "VELOCITY_RHS", lambda: d.velocityRHS, d.setVelocityRHS, None, False)
    yield ("D", "VELOCITY_RHS_GRAD", lambda: d.velocityRHSGrad, d.setVelocityRHSGrad, None, False)
    yield ("D", "PRESSURE_RHS", lambda: d.pressureRHS, d.setPressureRHS, None, False)
    yield ("D", "PRESSURE_RHS_GRAD", lambda: d.pressureRHSGrad, d.setPressureRHSGrad, None, False)
    yield ("D", "PRESSURE_RHS_DIV", lambda: d.pressureRHSdiv, d.setPressureRHSdiv, None, False)
    yield ("D", "PRESSURE_RHS_DIV_GRAD", lambda: d.pressureRHSdivGrad, d.setPressureRHSdivGrad, None, False)
    for bi, b in enumerate(d.getBlocks()):
        sec = ("B", bi)
        yield (sec, "VISCOSITY_BLOCK", lambda b=b: b.viscosity, b.setViscosity, b.clearViscosity, False)
        yield (sec, "VISCOSITY_BLOCK_GRAD", lambda b=b: b.viscosityGrad, b.setViscosityGrad, None, False)
        yield (sec, "VELOCITY", lambda b=b: b.velocity, b.setVelocity, None, False)
        yield (sec, "VELOCITY_GRAD", lambda b=b: b.velocityGrad, b.setVelocityGrad, None, False)
        yield (sec, "VELOCITY_SOURCE", lambda b=b: b.velocitySource, b.setVelocitySource, b.clearVelocitySource,
               False)
        yield (sec, "VELOCITY_SOURCE_GRAD", lambda b=b: b.velocitySourceGrad, b.setVelocitySourceGrad, None, False)
        yield (sec, "PASSIVE_SCALAR", lambda b=b: b.passiveScalar, b.setPassiveScalar, None, False)
        yield (sec, "PASSIVE_SCALAR_GRAD", lambda b=b: b.passiveScalarGrad, b.setPassiveScalarGrad, None, False)
        yield (sec, "PRESSURE", lambda b=b: b.pressure, b.setPressure, None, False)
        yield (sec, "PRESSURE_GRAD", lambda b=b: b.pressureGrad, b.setPressureGrad, None, False)
        for fi, fb in b.getFixedBoundaries():
            bsec = ("F", bi, fi)
            yield (bsec, "BOUNDARY_VELOCITY", lambda fb=fb: fb.velocity, fb.setVelocity, None, False)
            yield (bsec, "BOUNDARY_VELOCITY_GRAD", lambda fb=fb: fb.velocityGrad, fb.setVelocityGrad, None, False)
            yield (bsec, "BOUNDARY_PASSIVE_SCALAR", lambda fb=fb: fb.passiveScalar, fb.setPassiveScalar, None, False)
            yield (bsec, "BOUNDARY_PASSIVE_SCALAR_GRAD", lambda fb=fb: fb.passiveScalarGrad, fb.setPassiveScalarGrad,
                   None, False)


def _entries(domain):
    """_domain_entries as a list, cached on the domain for its current block and boundary objects."""
    key = tuple((id(b), tuple(id(bd) for bd in b.boundaries)) for b in domain.getBlocks())
    cached = getattr(domain, "_diff_entries", None)
    if cached is None or cached[0] != key:
        cached = (key, list(_domain_entries(domain)))
        domain._diff_entries = cached
    return cached[1]


def flatten_domain(domain, tensor_filter, clone=False, empty=False, exclusion_list=()):
    """The domain's tensors whose names pass `tensor_filter` as a flat list, plus the index map needed to put them
    back (PISOtorch_diff.py:44-118).  Absent optional tensors are None; `empty` or `exclusion_list` entries too."""
    index, tensors = {}, []
    for sec, name, get, _, _, on_cpu in _entries(domain):
        if tensor_filter is not None and name not in tensor_filter:
            continue
        t = None if (empty or name in exclusion_list) else get()
        if t is not None and clone:
            t = t.clone()
        if t is not None and on_cpu:
            t = t.cpu()
        index[(sec, name)] = len(tensors)
        tensors.append(t)
    return index, tensors


def set_domain_tensors_from_flat(domain, domain_dict, tensor_list, tensor_filter=None, clone=False):
    """Inverse of flatten_domain (PISOtorch_diff.py:120-181): set the listed tensors back on the domain (clearing
    optional ones that were None)."""
    for sec, name, _, setter, clear, _ in _entries(domain):
        if (sec, name) not in domain_dict or setter is None:
            continue
        if tensor_filter is not None and name not in tensor_filter:
            continue
        t = tensor_list[domain_dict[(sec, name)]]
        if t is not None:
            setter(t.clone() if clone else t)
        elif clear is not None:
            clear()


class _StageFunction(torch.autograd.Function):
    """autograd.Function of one stage; the stage is described by `spec` = (name, tracked, saved, forward,
    backward, outputs): `tracked` inputs (flatten_domain names) are its differentiable inputs, `saved` tensors are
    restored on the domain before the backward kernel runs, `outputs(domain)` are returned."""

    @staticmethod
    def forward(ctx, spec, domain, time_step, *tracked_tensors):
        name, tracked, saved, forward, backward, outputs = spec
        forward(domain)
        ctx.spec = spec
        ctx.domain = domain
        if saved:
            ctx.saved_dict, tensors = flatten_domain(domain, saved)
        else:
            ctx.saved_dict, tensors = None, []
        ctx.saved_none = [t is None for t in tensors]
        ctx.save_for_backward(*[t for t in tensors if t is not None], time_step)
        outs = outputs(domain)
        return outs if isinstance(outs, tuple) else (outs,)

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, *out_grads):
        name, tracked, saved, forward, backward, outputs = ctx.spec
        domain = ctx.domain
        grad_names = [t + "_GRAD" for t in tracked]
        if any(ctx.needs_input_grad[3:]):
            st = list(ctx.saved_tensors)
            time_step = st.pop()
            if ctx.saved_dict is not None:
                it = iter(st)
                tensors = [None if none else next(it) for none in ctx.saved_none]
                set_domain_tensors_from_flat(domain, ctx.saved_dict, tensors)
            backward(domain, time_step, out_grads)
            _, grads = flatten_domain(domain, grad_names, exclusion_list=_EXCLUDED_GRADIENTS[name][1])
        else:
            _, grads = flatten_domain(domain, grad_names, empty=True)
        return (None, None, None, *grads)


def _stage(fn_name, tracked, saved, forward, backward, outputs):
    return (fn_name, tracked, saved, forward, backward, outputs)


def _apply(spec, domain, time_step, tracked):
    _, tensors = flatten_domain(domain, tracked)
    return _StageFunction.apply(spec, domain, time_step, *tensors)


_EXCLUDED_GRADIENTS["SetupAdvectionMatrix"] = (set(), set())


def SetupAdvectionMatrix(domain, time_step, non_ortho_flags, forPassiveScalar=False, passiveScalarChannel=0):
    """PISOtorch_diff.py:377-441: (A, C.value) from block velocity, boundary velocity and viscosities (the velocity
    matrix), or the matrix of passive-scalar channel `passiveScalarChannel` (with the scalar viscosity)."""
    tracked = ["VELOCITY", "BOUNDARY_VELOCITY", "VISCOSITY"]
    tracked.append("PASSIVE_SCALAR_VISCOSITY" if forPassiveScalar else "VISCOSITY_BLOCK")
    name = "SetupAdvectionMatrix"

    def fwd(d):
        d.CreateA()
        d.C.CreateValue()
        d.UpdateDomainData()
        PISOtorch.SetupAdvectionMatrix(d, time_step, non_ortho_flags, forPassiveScalar=forPassiveScalar,
                                       passiveScalarChannel=passiveScalarChannel)

    def bwd(d, dt, grads):
        _set_gradient_input(grads[0], d.setAGrad, name, "A_GRAD")
        _set_gradient_input(grads[1], d.CGrad.setValue, name, "C_GRAD")
        d.CreateViscosityGrad()
        d.CreateVelocityGradOnBlocks()
        d.CreateVelocityGradOnBoundaries()
        d.UpdateDomainData()
        PISOtorch.SetupAdvectionMatrixGrad(d, dt, non_ortho_flags, forPassiveScalar=forPassiveScalar,
                                           passiveScalarChannel=passiveScalarChannel)

    fn = _stage(name, tracked, [], fwd, bwd, lambda d: (d.A, d.C.value))
    return _apply(fn, domain, time_step, tracked)


_EXCLUDED_GRADIENTS["SetupAdvectionVelocity"] = (set(), set())


def SetupAdvectionVelocity(domain, time_step, non_ortho_flags, apply_pressure_gradient=False):
    """PISOtorch_diff.py:531-614: velocityRHS from block velocity, velocity source, boundary velocity, viscosities
    (and the velocity result, through the explicit non-orthogonal terms)."""
    is_non_ortho = check_non_ortho_rhs(non_ortho_flags)
    tracked = ["VELOCITY", "VELOCITY_SOURCE", "BOUNDARY_VELOCITY", "VISCOSITY", "VISCOSITY_BLOCK"]
    saved = ["BOUNDARY_VELOCITY", "VISCOSITY", "VISCOSITY_BLOCK", "VELOCITY_SOURCE"]
    if is_non_ortho:
        tracked.append("VELOCITY_RESULT")
        saved.append("VELOCITY_RESULT")
    name = "SetupAdvectionVelocity"

    def fwd(d):
        d.CreateVelocityRHS()
        d.UpdateDomainData()
        PISOtorch.SetupAdvectionVelocity(d, time_step, non_ortho_flags, apply_pressure_gradient)

    def bwd(d, dt, grads):
        _set_gradient_input(grads[0], d.setVelocityRHSGrad, name, "VELOCITY_RHS_GRAD")
        d.CreateViscosityGrad()
        d.CreateVelocityGradOnBlocks()
        d.CreateVelocityGradOnBoundaries()
        d.CreateVelocitySourceGradOnBlocks()
        if is_non_ortho:
            d.CreateVelocityResultGrad()
        d.UpdateDomainData()
        PISOtorch.SetupAdvectionVelocityGrad(d, dt, non_ortho_flags, apply_pressure_gradient)

    fn = _stage(name, tracked, saved, fwd, bwd, lambda d: d.velocityRHS)
    return _apply(fn, domain, time_step, tracked)[0]


class _CopyFunction(torch.autograd.Function):
    """autograd.Function of a Copy*Result* stage; `spec` = (forward, backward, outputs)."""

    @staticmethod
    def forward(ctx, spec, domain, *inputs):
        forward, backward, outputs = spec
        ctx.spec = spec
        ctx.domain = domain
        forward(domain)
        outs = outputs(domain)
        return outs if isinstance(outs, tuple) else (outs,)

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, *grads):
        if any(ctx.needs_input_grad[2:]):
            res = ctx.spec[1](ctx.domain, grads)
        else:
            res = [None] * (len(ctx.needs_input_grad) - 2)
        return (None, None, *res)


def _copy_stage(forward, backward, tensors_in, outputs):
    return (forward, backward, outputs)


def CopyVelocityResultToBlocks(domain):
    """PISOtorch_diff.py:697-729: block velocities from velocityResult."""
    def fwd(d):
        d.CreateVelocityOnBlocks()
        d.UpdateDomainData()
        PISOtorch.CopyVelocityResultToBlocks(d)

    def bwd(d, grads):
        for b, g in zip(d.getBlocks(), grads):
            b.setVelocityGrad(g.contiguous())
        d.CreateVelocityResultGrad()
        d.UpdateDomainData()
        PISOtorch.CopyVelocityResultGradFromBlocks(d)
        return [d.velocityResultGrad]

    fn = _copy_stage(fwd, bwd, None, lambda d: tuple(b.velocity for b in d.getBlocks()))
    _CopyFunction.apply(fn, domain, domain.velocityResult)


def CopyVelocityResultFromBlocks(domain):
    """PISOtorch_diff.py:731-763: velocityResult from the block velocities."""
    def fwd(d):
        d.CreateVelocityResult()
        d.UpdateDomainData()
        PISOtorch.CopyVelocityResultFromBlocks(d)

    def bwd(d, grads):
        d.setVelocityResultGrad(grads[0].contiguous())
        d.CreateVelocityGradOnBlocks()
        d.UpdateDomainData()
        PISOtorch.CopyVelocityResultGradToBlocks(d)
        return [b.velocityGrad for b in d.getBlocks()]

    fn = _copy_stage(fwd, bwd, None, lambda d: d.velocityResult)
    _CopyFunction.apply(fn, domain, *[b.velocity for b in domain.getBlocks()])


def CopyPressureResultToBlocks(domain):
    """PISOtorch_diff.py:765-799: block pressures from pressureResult."""
    def fwd(d):
        d.CreatePressureOnBlocks()
        d.UpdateDomainData()
        PISOtorch.CopyPressureResultToBlocks(d)

    def bwd(d, grads):
        for b, g in zip(d.getBlocks(), grads):
            b.setPressureGrad(g.contiguous())
        d.CreatePressureResultGrad()
        d.UpdateDomainData()
        PISOtorch.CopyPressureResultGradFromBlocks(d)
        return [d.pressureResultGrad]

    fn = _copy_stage(fwd, bwd, None, lambda d: tuple(b.pressure for b in d.getBlocks()))
    _CopyFunction.apply(fn, domain, domain.pressureResult)


def CopyScalarResultToBlocks(domain):
    """PISOtorch_diff.py:616-651: block passive scalars from scalarResult."""
    def fwd(d):
        d.CreatePassiveScalarOnBlocks()
        d.UpdateDomainData()
        PISOtorch.CopyScalarResultToBlocks(d)

    def bwd(d, grads):
        for b, g in zip(d.getBlocks(), grads):
            b.setPassiveScalarGrad(g.contiguous())
        d.CreateScalarResultGrad()
        d.UpdateDomainData()
        PISOtorch.CopyScalarResultGradFromBlocks(d)
        return [d.scalarResultGrad]

    fn = _copy_stage(fwd, bwd, None, lambda d: tuple(b.passiveScalar for b in d.getBlocks()))
    _CopyFunction.apply(fn, domain, domain.scalarResult)


def CopyScalarResultFromBlocks(domain):
    """PISOtorch_diff.py:653-695: scalarResult from the block passive scalars (its backward is
    CopyScalarResultGradToBlocks, which the reference's wrapper leaves unimplemented)."""
    def fwd(d):
        d.CreateScalarResult()
        d.UpdateDomainData()
        PISOtorch.CopyScalarResultFromBlocks(d)

    def bwd(d, grads):
        d.setScalarResultGrad(grads[0].contiguous())
        d.CreatePassiveScalarGradOnBlocks()
        d.UpdateDomainData()
        PISOtorch.CopyScalarResultGradToBlocks(d)
        return [b.passiveScalarGrad for b in d.getBlocks()]

    fn = _copy_stage(fwd, bwd, None, lambda d: d.scalarResult)
    _CopyFunction.apply(fn, domain, *[b.passiveScalar for b in domain.getBlocks()])


_EXCLUDED_GRADIENTS["SetupAdvectionScalar"] = (set(), set())


def SetupAdvectionScalar(domain, time_step, non_ortho_flags):
    """PISOtorch_diff.py:444-529: scalarRHS from the block passive scalars, boundary scalars and velocities, the
    viscosities (and the scalar result, through the explicit non-orthogonal terms)."""
    is_non_ortho = check_non_ortho_rhs(non_ortho_flags)
    tracked = ["PASSIVE_SCALAR", "BOUNDARY_PASSIVE_SCALAR", "BOUNDARY_VELOCITY", "VISCOSITY",
               "PASSIVE_SCALAR_VISCOSITY"]
    saved = ["BOUNDARY_PASSIVE_SCALAR", "BOUNDARY_VELOCITY", "VISCOSITY", "PASSIVE_SCALAR_VISCOSITY"]
    if is_non_ortho:
        tracked.append("PASSIVE_SCALAR_RESULT")
        saved.append("PASSIVE_SCALAR_RESULT")
    name = "SetupAdvectionScalar"

    def fwd(d):
        d.CreateScalarRHS()
        d.UpdateDomainData()
        PISOtorch.SetupAdvectionScalar(d, time_step, non_ortho_flags)

    def bwd(d, dt, grads):
        _set_gradient_input(grads[0], d.setScalarRHSGrad, name, "PASSIVE_SCALAR_RHS_GRAD")
        d.CreateViscosityGrad()
        d.CreatePassiveScalarGradOnBlocks()
        d.CreatePassiveScalarGradOnBoundaries()
        d.CreateVelocityGradOnBoundaries()
        if is_non_ortho:
            d.CreateScalarResultGrad()
        d.UpdateDomainData()
        PISOtorch.SetupAdvectionScalarGrad(d, dt, non_ortho_flags)

    fn = _stage(name, tracked, saved, fwd, bwd, lambda d: d.scalarRHS)
    return _apply(fn, domain, time_step, tracked)[0]


_PRESSURE_TRACKED = ["VELOCITY", "VELOCITY_SOURCE", "VELOCITY_RESULT", "BOUNDARY_VELOCITY", "VISCOSITY",
                     "VISCOSITY_BLOCK", "A", "C"]
_PRESSURE_SAVED = ["BOUNDARY_VELOCITY", "VISCOSITY", "VISCOSITY_BLOCK", "A", "C", "VELOCITY_RESULT", "PRESSURE_RHS",
                   "VELOCITY_SOURCE"]


def _pressure_lists(non_ortho_flags):
    """Tracked and saved tensors of the pressure RHS stages: with explicit non-orthogonal terms the RHS also
    depends on the pressure result (PISOtorch_diff.py:807-818, :985-994).  It is saved for the backward pass here
    (the reference leaves it out of its saved list and reads whatever the domain holds then)."""
    if not check_non_ortho_rhs(non_ortho_flags):
        return _PRESSURE_TRACKED, _PRESSURE_SAVED
    return _PRESSURE_TRACKED + ["PRESSURE_RESULT"], _PRESSURE_SAVED + ["PRESSURE_RESULT"]


def _pressure_rhs_grad_setup(d, name, prhs_grad, prhs_div_grad, non_ortho_flags=0):
    _set_gradient_input(prhs_grad, d.setPressureRHSGrad, name, "PRESSURE_RHS_GRAD")
    _set_gradient_input(prhs_div_grad, d.setPressureRHSdivGrad, name, "PRESSURE_RHS_DIV_GRAD")
    d.CreateAGrad()
    d.CGrad.CreateValue()
    d.CreateViscosityGrad()
    d.CreateVelocityResultGrad()
    d.CreateVelocityGradOnBlocks()
    d.CreateVelocitySourceGradOnBlocks()
    d.CreateVelocityGradOnBoundaries()
    if check_non_ortho_rhs(non_ortho_flags):
        d.CreatePressureResultGrad()


_EXCLUDED_GRADIENTS["SetupPressureCorrection"] = (set(), set())


def SetupPressureCorrection(domain, time_step, non_ortho_flags, use_face_transform=False, timeStepNorm=False):
    """PISOtorch_diff.py:802-894: (P.value, pressureRHS, pressureRHSdiv)."""
    name = "SetupPressureCorrection"
    tracked, saved = _pressure_lists(non_ortho_flags)

    def fwd(d):
        d.CreatePressureRHS()
        d.CreatePressureRHSdiv()
        d.P.CreateValue()
        d.UpdateDomainData()
        PISOtorch.SetupPressureCorrection(d, time_step, non_ortho_flags, use_face_transform, timeStepNorm=timeStepNorm)

    def bwd(d, dt, grads):
        if use_face_transform:
            raise NotImplementedError("SetupPressureCorrection: Gradient does not support face transformations.")
        _set_gradient_input(grads[0], d.PGrad.setValue, name, "P_GRAD")
        _pressure_rhs_grad_setup(d, name, grads[1], grads[2], non_ortho_flags)
        d.UpdateDomainData()
        PISOtorch.SetupPressureCorrectionGrad(d, dt, non_ortho_flags, use_face_transform, timeStepNorm)

    fn = _stage(name, tracked, saved, fwd, bwd, lambda d: (d.P.value, d.pressureRHS, d.pressureRHSdiv))
    _apply(fn, domain, time_step, tracked)


_EXCLUDED_GRADIENTS["SetupPressureMatrix"] = (set(), set())


def SetupPressureMatrix(domain, time_step, non_ortho_flags, use_face_transform=False):
    """PISOtorch_diff.py:897-964: P.value from A."""
    name = "SetupPressureMatrix"

    def fwd(d):
        d.P.CreateValue()
        d.UpdateDomainData()
        PISOtorch.SetupPressureMatrix(d, time_step, non_ortho_flags, use_face_transform)

    def bwd(d, dt, grads):
        if use_face_transform:
            raise NotImplementedError("SetupPressureMatrix: Gradient does not support face transformations.")
        _set_gradient_input(grads[0], d.PGrad.setValue, name, "P_GRAD")
        d.CreateAGrad()
        d.UpdateDomainData()
        PISOtorch.SetupPressureMatrixGrad(d, dt, non_ortho_flags, use_face_transform)

    fn = _stage(name, ["A"], ["A"], fwd, bwd, lambda d: d.P.value)
    _apply(fn, domain, time_step, ["A"])


_EXCLUDED_GRADIENTS["SetupPressureRHS"] = (set(), set())


def SetupPressureRHS(domain, time_step, non_ortho_flags, use_face_transform=False, timeStepNorm=False):
    """PISOtorch_diff.py:967-1064: (pressureRHS, pressureRHSdiv)."""
    name = "SetupPressureRHS"
    tracked, saved = _pressure_lists(non_ortho_flags)

    def fwd(d):
        d.CreatePressureRHS()
        d.CreatePressureRHSdiv()
        d.UpdateDomainData()
        PISOtorch.SetupPressureRHS(d, time_step, non_ortho_flags, use_face_transform, timeStepNorm=timeStepNorm)

    def bwd(d, dt, grads):
        if use_face_transform:
            raise NotImplementedError("SetupPressureRHS: Gradient does not support face transformations.")
        _pressure_rhs_grad_setup(d, name, grads[0], grads[1], non_ortho_flags)
        d.UpdateDomainData()
        PISOtorch.SetupPressureRHSGrad(d, dt, non_ortho_flags, use_face_transform, timeStepNorm)

    fn = _stage(name, tracked, saved, fwd, bwd, lambda d: (d.pressureRHS, d.pressureRHSdiv))
    _apply(fn, domain, time_step, tracked)


_EXCLUDED_GRADIENTS["SetupPressureRHSdiv"] = (set(), set())


def SetupPressureRHSdiv(domain, time_step, non_ortho_flags, use_face_transform=False, timeStepNorm=False):
    """PISOtorch_diff.py:1067-1157: pressureRHSdiv from pressureRHS (and the boundary velocities; with explicit
    non-orthogonal terms also from the pressure result and A)."""
    name = "SetupPressureRHSdiv"
    is_non_ortho = check_non_ortho_rhs(non_ortho_flags)
    tracked = ["PRESSURE_RHS", "BOUNDARY_VELOCITY"]
    saved = []
    if is_non_ortho:
        tracked += ["PRESSURE_RESULT", "A"]
        saved = ["PRESSURE_RESULT", "A"]

    def fwd(d):
        d.CreatePressureRHSdiv()
        d.UpdateDomainData()
        PISOtorch.SetupPressureRHSdiv(d, time_step, non_ortho_flags, use_face_transform, timeStepNorm=timeStepNorm)

    def bwd(d, dt, grads):
        if use_face_transform:
            raise NotImplementedError("SetupPressureRHSdiv: Gradient does not support face transformations.")
        _set_gradient_input(grads[0], d.setPressureRHSdivGrad, name, "PRESSURE_RHS_DIV_GRAD")
        if is_non_ortho:
            d.CreateAGrad()
        d.CreatePressureRHSGrad()
        d.CreateVelocityGradOnBoundaries()
        if is_non_ortho:
            d.CreatePressureResultGrad()
        d.UpdateDomainData()
        PISOtorch.SetupPressureRHSdivGrad(d, dt, non_ortho_flags, use_face_transform, timeStepNorm)

    fn = _stage(name, tracked, saved, fwd, bwd, lambda d: d.pressureRHSdiv)
    _apply(fn, domain, time_step, tracked)


_EXCLUDED_GRADIENTS["CorrectVelocity"] = (set(), set())


def CorrectVelocity(domain, time_step, version, timeStepNorm=False):
    """PISOtorch_diff.py:1160-1234: velocityResult from A, block pressures and pressureRHS."""
    name = "CorrectVelocity"
    tracked = ["A", "PRESSURE", "PRESSURE_RHS"]

    def fwd(d):
        d.CreateVelocityResult()
        d.UpdateDomainData()
        PISOtorch.CorrectVelocity(d, time_step, version, timeStepNorm=timeStepNorm)

    def bwd(d, dt, grads):
        if version not in (0, 1):
            raise NotImplementedError("CorrectVelocity: Gradient only implemented for Finite Difference pressure "
                                      "gradients (version 0 or 1).")
        _set_gradient_input(grads[0], d.setVelocityResultGrad, name, "VELOCITY_RESULT_GRAD")
        d.CreateAGrad()
        d.CreatePressureGradOnBlocks()
        d.CreatePressureRHSGrad()
        d.UpdateDomainData()
        PISOtorch.CorrectVelocityGrad(d, dt, timeStepNorm)

    fn = _stage(name, tracked, ["A", "PRESSURE"], fwd, bwd, lambda d: d.velocityResult)
    _apply(fn, domain, time_step, tracked)


class _SubtractMeanFunction(torch.autograd.Function):
    """x - mean(x); its VJP is g - mean(g)."""

    @staticmethod
    def forward(ctx, x):
        return PISOtorch._subtract_mean(x)

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, g):
        return PISOtorch._subtract_mean(g.contiguous())


def _subtract_mean(x):
    return _SubtractMeanFunction.apply(x)


def reset_domain(domain):
    """PISOtorch_diff.py:1238-1242: fresh forward tensors, so that autograd graphs held by the domain are freed."""
    domain.CreateVelocityOnBlocks()
    domain.CreatePressureOnBlocks()
    domain.PrepareSolve()


def detach_domain_fwd(domain):
    domain.DetachFwd()


def detach_domain_grad(domain):
    domain.DetachGrad()


def detach_domain(domain):
    domain.Detach()


def is_tensor_empty(tensor):
    return tensor.dim() == 1 and tensor.size(0) == 0


# Forward-only passthroughs (no gradient in the reference either)
from .PISOtorch import CopyPressureResultFromBlocks  # noqa: E402,F401


class _MatmulFunction(torch.autograd.Function):
    """matmul with its VJP from PISOtorch.matmulGrad (PISOtorch_diff.py:1300-1371): a gradient is refused for an
    inverted output, and for an inverted operand whose gradient is needed."""

    @staticmethod
    def forward(ctx, vectorMatrixA, vectorMatrixB, flags):
        ctx.flags = flags
        ctx.save_for_backward(vectorMatrixA, vectorMatrixB)
        return PISOtorch.matmul(vectorMatrixA, vectorMatrixB, *flags)

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, result_grad):
        if not any(ctx.needs_input_grad[:2]):
            return None, None, None
        transposeA, invertA, transposeB, invertB, transposeOutput, invertOutput = ctx.flags
        if invertOutput:
            raise NotImplementedError("Can not compute gradients for inverted matrices (output).")
        if ctx.needs_input_grad[0] and invertA:
            raise NotImplementedError("Can not compute gradients for inverted matrices (input A).")
        if ctx.needs_input_grad[1] and invertB:
            raise NotImplementedError("Can not compute gradients for inverted matrices (input B).")
        a, b = ctx.saved_tensors
        ga, gb = PISOtorch.matmulGrad(a, b, result_grad.contiguous(), *ctx.flags)
        return (ga if ctx.needs_input_grad[0] else None), (gb if ctx.needs_input_grad[1] else None), None


def matmul(vectorMatrixA, vectorMatrixB, transposeA=False, invertA=False, transposeB=False, invertB=False,
           transposeOutput=False, invertOutput=False):
    """Differentiable per-cell product of vectors / flat row-major matrices in the channel dimension
    (PISOtorch_diff.py:1300-1371): a matrix if both are matrices, a vector if exactly one is, a scalar otherwise."""
    flags = (bool(transposeA), bool(invertA), bool(transposeB), bool(invertB), bool(transposeOutput),
             bool(invertOutput))
    return _MatmulFunction.apply(vectorMatrixA, vectorMatrixB, flags)
