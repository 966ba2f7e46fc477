"""PISOtorch_simulation -- the PISO step driver over PISOtorch / PISOtorch_diff.

Mirrors `Simulation` of the reference (tum-pbs/PICT PISOtorch_simulation.py:388-1518):
the same constructor arguments and defaults, the same call order in
`_PISO_split_step` (:1054-1327), the same solver defaults (BiCGStab for
advection, CG for pressure, tolerance None -> 1e-5 FP32 / 1e-8 FP64, maxit 5000,
pressure mean subtraction) and the same `prep_fn` hook points.

Image/VTK output, adaptive CFL stepping and the scipy solve path are host
conveniences outside the PISO step and are not reproduced: the corresponding
arguments are accepted and must stay at their defaults.
"""
import ctypes
import logging
import numbers
from contextlib import nullcontext

import numpy as np
import torch

from . import PISOtorch
from . import PISOtorch_diff

_LOG = logging.getLogger("PISOsim")
cpu_device = torch.device("cpu")

# Asynchronous native calls hold one of the library's 16 pending-step slots until their solver results are
# collected (pict_step_wait); the slots are shared by every Simulation of the process.
_MAX_PENDING_STEPS = 16
_deferring = []  # Simulations that hold pending checks, in the order they queued their first one


def _pending_steps_total():
    return sum(len(s._pending_checks) for s in _deferring)


def tensor_as_np(tensor):
    return tensor.detach().cpu().numpy()


def ntonp(t):
    return t.detach().cpu().numpy() if isinstance(t, torch.Tensor) else np.asarray(t)


# --- module-level helpers (PISOtorch_simulation.py:55-386) ---

def get_max_time_step(domain, time_step_target, CFL_cond=0.8, with_transformations=True):
    """Largest time step <= time_step_target with max|u| dt <= CFL_cond, as (time step, substeps) (:55-65)."""
    max_vel = float(ntonp(domain.getMaxVelocity(True, with_transformations)))
    limit = CFL_cond / max_vel if max_vel > 0 else float("inf")
    if limit >= time_step_target:
        return time_step_target, 1
    substeps = int(np.ceil(time_step_target / limit))
    return time_step_target / substeps, substeps


def getVelocityResultMaxMag(domain):
    """Largest per-cell velocity magnitude of domain.velocityResult (:67-73)."""
    vel = domain.velocityResult.reshape(domain.getSpatialDims(), domain.getTotalSize())
    if float(vel.abs().max()) > 0:
        return float(torch.linalg.vector_norm(vel, dim=0).max())
    return 0


def getVelocityResultMaxVel(domain):
    """Largest velocity component magnitude of domain.velocityResult (:75-78)."""
    return float(domain.velocityResult.abs().max())


def _face_slice(t, face):
    """The layer of cells next to `face` (0 = -x, 1 = +x, 2 = -y, ...) of an NC[D]HW tensor, one cell thick."""
    dim = t.dim() - 1 - (face >> 1)
    n = t.size(dim)
    return t.narrow(dim, n - 1 if face & 1 else 0, 1)


def get_fixed_boundary_flux(bound, bound_idx):
    """Total boundary-normal flux through a FixedBoundary, in the boundary's own orientation (:80-90)."""
    if not isinstance(bound, PISOtorch.FixedBoundary):
        raise AssertionError("get_fixed_boundary_flux needs a FixedBoundary")
    if bound.velocityType != PISOtorch.BoundaryConditionType.DIRICHLET:
        raise ValueError("Only DIRICHLET boundaries are supportet")
    return torch.sum(bound.GetFluxes())


def get_fixed_boundary_fluxes(list_idx_bound):
    """Net outward flux through a list of (face index, FixedBoundary): a lower face's flux counts negative,
    an upper face's positive (:92-107)."""
    domain = list_idx_bound[0][1].getParentDomain()
    total = torch.zeros([1], dtype=domain.getDtype(), device=domain.getDevice())
    for face, bound in list_idx_bound:
        if not isinstance(bound, PISOtorch.FixedBoundary):
            raise TypeError("boundary type not supported.")
        flux = get_fixed_boundary_flux(bound, face)
        total = total + flux if face & 1 else total - flux
    return total


def get_advective_velocity(velms, velm_idx, bound, bound_idx):
    """The boundary-normal advection speed, in computational units, with which an outflow boundary is advected:
    velms is one [1, dims] tensor for every boundary or a list with one tensor per boundary, static [1, dims] or
    varying [1, dims, <boundary shape>].  With a transformed boundary the speed is row `axis` of the inverse
    transform times the velocity, per boundary cell; otherwise it is the velocity's `axis` component
    (:145-172).  Returns [1, 1] (static, untransformed) or [1, 1, <boundary shape>]."""
    velm = velms if isinstance(velms, torch.Tensor) else velms[velm_idx]
    dims = bound.getSpatialDims()
    axis = bound_idx >> 1
    size = bound.getSizes()
    shape = [size[dims - 1 - d] for d in range(dims)]  # z, y, x
    static = velm.dim() == 2
    if not static and list(velm.shape[2:]) != shape:
        raise AssertionError("varying velm size must match boundary")
    if not bound.hasTransform():
        return velm[:, axis:axis + 1]
    if static:
        velm = velm.reshape([1, dims] + [1] * dims).expand([1, dims] + shape)
    row = bound.transform[..., dims * dims + axis * dims: dims * dims + (axis + 1) * dims]  # [1, z, y, x, dims]
    return (row * torch.movedim(velm, 1, -1)).sum(-1).unsqueeze(1)


def balance_boundary_fluxes(domain, free_bounds, tol=None):
    """Scales the velocities of the FixedBoundaries in free_bounds (every component) so that the net flux over
    all fixed boundaries of the domain vanishes, unless it already does within 1% of the solver tolerance
    (:174-207)."""
    with torch.no_grad():
        faces = [fb for block in domain.getBlocks() for fb in block.getFixedBoundaries()]
        fixed = [fb for fb in faces if not any(fb[1] is b for b in free_bounds)]
        free = [fb for fb in faces if any(fb[1] is b for b in free_bounds)]
        fixed_flux = get_fixed_boundary_fluxes(fixed) if fixed else \
            torch.zeros([1], dtype=domain.getDtype(), device=domain.getDevice())
        free_flux = get_fixed_boundary_fluxes(free) if free else torch.zeros_like(fixed_flux)
        atol = PISOtorch_diff._get_solver_tolerance(tol, dtype=domain.getDtype()) * 0.01
        if torch.allclose(fixed_flux + free_flux, torch.zeros_like(fixed_flux), atol=atol):
            return
        scale = -fixed_flux / free_flux
        for _, bound in free:
            bound.setVelocity(bound.velocity * scale)


def update_advective_boundaries(domain, bounds, velms, dt, tol=None):
    """Advective (convective) outflow: each boundary in `bounds` (varying FixedBoundaries of the domain) moves
    towards the cells next to it, u_b <- u_b - t (u_b - u_cell) with t = 1 - 1/(1 + 2 dt a), a the advection
    speed of get_advective_velocity; a varying Dirichlet passive scalar on the boundary follows the same rule.
    Then balance_boundary_fluxes rescales these boundaries so that the domain's boundary flux is balanced
    (:210-311)."""
    dims = domain.getSpatialDims()
    # argument checks are assertions in the reference: the same exception type here
    if not (isinstance(bounds, list) and len(bounds) > 0):
        raise AssertionError("bounds must be a non-empty list of boundaries")

    def _ok(v, varying):
        return (isinstance(v, torch.Tensor) and v.size(0) == 1 and v.size(1) == dims
                and (v.dim() == 2 or (varying and v.dim() == dims + 2)))

    if not (_ok(velms, False) or (isinstance(velms, list) and len(velms) == len(bounds)
                                  and all(_ok(v, True) for v in velms))):
        raise AssertionError("velms tensors have wrong shape")
    with torch.no_grad():
        found = []
        for block in domain.getBlocks():
            for face in range(2 * dims):
                bound = block.getBoundary(face)
                if isinstance(bound, PISOtorch.FixedBoundary) and any(bound is b for b in bounds):
                    found.append((block, face, bound))
        if len(found) != len(bounds):
            raise ValueError("Bounds passed as advective outflow are not (all) part of the domain.")
        if isinstance(dt, torch.Tensor):
            dt = dt.to(domain.getDevice())
        for block, face, bound in found:
            idx = next(i for i, b in enumerate(bounds) if b is bound)
            alpha = dt * 2 * get_advective_velocity(velms, idx, bound, face)
            t = 1 - 1 / (1 + alpha)
            if bound.isVelocityStatic:
                raise ValueError("outflow boundaries must be varying. Use FixedBoundary.makeVelocityVarying() to "
                                 "make the velocity varying.")
            scalar = None
            if block.hasPassiveScalar() and bound.hasPassiveScalar():
                if bound.isPassiveScalarStatic():
                    raise ValueError("outflow boundary passive scalar must be varying.")
                if any(c != PISOtorch.BoundaryConditionType.DIRICHLET for c in bound.passiveScalarTypes):
                    raise ValueError("passive scalar outflow boundary condition must be Dirichlet.")
                scalar = bound.passiveScalar
            vel = bound.velocity
            bound.setVelocity(vel - t * (vel - _face_slice(block.velocity, face)))
            if scalar is not None:
                bound.setPassiveScalar(scalar - t * (scalar - _face_slice(block.passiveScalar, face)))
    balance_boundary_fluxes(domain, bounds, tol=tol)


def update_advective_boundaries_static(domain, bounds, velms, dt):
    """The reference's static variant (:313-347) acts only on the deprecated Varying/StaticDirichletBoundary
    types, which this package does not have; over FixedBoundaries it changes nothing, and so it does here."""
    return None


def append_prep_fn(prep_fn, name, fn):
    """Adds fn after the hooks already registered under `name` (:349-361)."""
    return _add_prep_fn(prep_fn, name, fn, False)


def prepend_prep_fn(prep_fn, name, fn):
    """Adds fn before the hooks already registered under `name` (:363-375)."""
    return _add_prep_fn(prep_fn, name, fn, True)


def _add_prep_fn(prep_fn, name, fn, first):
    if not isinstance(prep_fn, dict):
        raise TypeError("prep_fn must be a dict")
    fns = prep_fn.get(name, [])
    fns = list(fns) if isinstance(fns, (list, tuple)) else [fns]
    fns.insert(0, fn) if first else fns.append(fn)
    prep_fn[name] = fns
    return prep_fn


def add_pressure_ts_norm(prep_fn):
    """Hooks that solve for dt * p instead of p: the pressure RHS divergence is multiplied by the time step after
    the pressure setup and the result divided by it again after the pressure solve (:377-386)."""
    def scale_rhs(domain, time_step, **kwargs):
        domain.setPressureRHSdiv(domain.pressureRHSdiv * time_step.to(domain.getDevice()))
        domain.UpdateDomainData()

    def unscale_result(domain, time_step, **kwargs):
        domain.setPressureResult(domain.pressureResult / time_step.to(domain.getDevice()))
        domain.UpdateDomainData()
    prepend_prep_fn(prep_fn, "POST_PRESSURE_SETUP", scale_rhs)
    prepend_prep_fn(prep_fn, "POST_PRESSURE_NON_ORTHO", unscale_result)
    return prep_fn


class StopHandler:
    """PISOtorch_simulation.py:27-49 (without signal registration side effects unless asked)."""

    def __init__(self, log=None):
        self._log = log
        self.reset()

    def stop(self):
        self._stop = True

    def __call__(self):
        return self._stop

    def reset(self):
        self._stop = False


def _never_stop():
    return False


class Simulation:
    # must match PISO_multiblock_cuda.h:9-13
    NON_ORTHO_DIRECT_MATRIX = 1
    NON_ORTHO_DIRECT_RHS = 2
    NON_ORTHO_DIAGONAL_MATRIX = 4
    NON_ORTHO_DIAGONAL_RHS = 8
    NON_ORTHO_CENTER_MATRIX = 16
    _NON_ORTHO_MODE = NON_ORTHO_CENTER_MATRIX | NON_ORTHO_DIRECT_MATRIX | NON_ORTHO_DIAGONAL_RHS

    def __init__(self, domain=None, *, time_step=1.0, substeps=1, corrector_steps=2, density_viscosity=None,
                 adaptive_CFL=0.8, prep_fn=None, advection_use_BiCG=True, pressure_use_BiCG=False,
                 scipy_solve_advection=False, scipy_solve_pressure=False, preconditionBiCG=False,
                 BiCG_precondition_fallback=True, advection_tol=None, pressure_tol=None, convergence_tol=None,
                 solver_double_fallback=False, advect_non_ortho_steps=1, pressure_non_ortho_steps=1,
                 advect_reuse_result=False, pressure_reuse_result=False, normalize_pressure_result=True,
                 pressure_return_best_result=False, advect_passive_scalar=True, pressure_time_step_normalized=False,
                 velocity_corrector="FD", non_orthogonal=True, differentiable=False,
                 exclude_advection_solve_gradients=False, exclude_pressure_solve_gradients=False, log_dir=None,
                 log_interval=0, log_images=True, log_vtk=False, norm_vel=False, block_layout=None,
                 output_mode3D="slice", log_fn=None, output_resampling_coords=None, output_resampling_shape=None,
                 output_resampling_fill_max_steps=0, save_domain_name=None, stop_fn=_never_stop, native_step=True,
                 pressure_refine_tol=None, pressure_refine_inner_tol=1e-5, pressure_refine_max_sweeps=8):
        if density_viscosity is not None:
            raise NotImplementedError("Simulation.density_viscosity is superseded by Domain.scalarViscosity.")
        if scipy_solve_advection or scipy_solve_pressure:
            raise NotImplementedError("the scipy solve path is not part of this build")
        self.domain = domain
        self.prep_fn = prep_fn
        self.non_orthogonal = non_orthogonal
        self.time_step = time_step
        self.substeps = substeps
        self.corrector_steps = corrector_steps
        self.convergence_tol = convergence_tol
        self.adaptive_CFL = adaptive_CFL
        self.scipy_solve_advection = False
        self.advection_use_BiCG = advection_use_BiCG
        self.advect_non_ortho_steps = advect_non_ortho_steps
        self._advection_tol = advection_tol
        self.advect_passive_scalar = advect_passive_scalar
        self.advect_reuse_result = advect_reuse_result
        self.scipy_solve_pressure = False
        self.pressure_use_BiCG = pressure_use_BiCG
        self.pressure_non_ortho_steps = pressure_non_ortho_steps
        self._pressure_tol = pressure_tol
        self.normalize_pressure_result = normalize_pressure_result
        self.pressure_return_best_result = pressure_return_best_result
        self.pressure_time_step_normalized = pressure_time_step_normalized
        self.pressure_reuse_result = pressure_reuse_result
        self.solver_double_fallback = solver_double_fallback
        self.linear_solve_max_iterations = 5000
        self.preconditionBiCG = preconditionBiCG
        self.BiCG_precondition_fallback = BiCG_precondition_fallback
        self._velocity_corrector_versions = {"FD": 1, "FVM_CENTER": 5, "FVM_FACE": 6}
        self.velocity_corrector = velocity_corrector
        self.log_dir = log_dir
        self.log_interval = log_interval
        self.log_images = log_images
        self.log_vtk = log_vtk
        self.norm_vel = norm_vel
        self.block_layout = block_layout
        self.log_fn = log_fn
        self.save_domain_name = save_domain_name
        self.differentiable = differentiable
        self.exclude_advection_solve_gradients = exclude_advection_solve_gradients
        self.exclude_pressure_solve_gradients = exclude_pressure_solve_gradients
        self.stop_fn = stop_fn
        # run eligible steps as one native call (pict_piso_step); not a reference option
        self.native_step = native_step
        # FP64 iterative refinement of the pressure solve of a non-differentiable step to this relative residual
        # (PISOtorch.SolveLinearRefined; not a reference option): each correction solve runs in the domain dtype
        # to pressure_refine_inner_tol on its unit-RMS right-hand side
        self.pressure_refine_tol = pressure_refine_tol
        self.pressure_refine_inner_tol = pressure_refine_inner_tol
        self.pressure_refine_max_sweeps = pressure_refine_max_sweeps
        self.solver_infos = []  # (kind, [LinearSolverResultInfo]) of the last step, for iteration-count logging
        self.record_solver_infos = False
        # native steps (and checkpointed reverse passes) queued without the host wait for their solver results;
        # flush_solver_checks() waits for them and raises what the waiting calls would have raised
        self.defer_solver_checks = False
        self._pending_checks = []
        self.reset_step_counters()

    # --- properties (validation as in :552-835) ---
    @property
    def domain(self):
        return self._domain

    @domain.setter
    def domain(self, domain):
        if domain is not None:
            if not isinstance(domain, PISOtorch.Domain):
                raise TypeError("domain must be a PISOtorch.Domain object or None.")
            if not domain.IsInitialized():
                raise RuntimeError("domain must be initilized. Call domain.PrepareSolve() before assignment.")
        self._domain = domain

    def _check_domain(self):
        if self._domain is None:
            raise ValueError("no domain set.")

    def _get_dtype(self):
        self._check_domain()
        if self.domain.getNumBlocks() == 0:
            raise RuntimeError("Block required to infer dtype")
        return self.domain.getBlock(0).velocity.dtype

    @property
    def differentiable(self):
        return self._differentiable

    @differentiable.setter
    def differentiable(self, d):
        if not isinstance(d, bool):
            raise TypeError("differentiable must be bool.")
        self._differentiable = d
        self._backend = PISOtorch_diff if d else PISOtorch

    @property
    def non_orthogonal(self):
        return self._non_orthogonal

    @non_orthogonal.setter
    def non_orthogonal(self, v):
        if not isinstance(v, bool):
            raise TypeError("non_orthogonal must be bool.")
        self._non_orthogonal = v
        self._non_ortho_flags = Simulation._NON_ORTHO_MODE if v else 0

    @property
    def time_step(self):
        return self._time_step

    @time_step.setter
    def time_step(self, t):
        if not isinstance(t, numbers.Real):
            raise TypeError("time_step must be float.")
        self._time_step = t

    def _get_time_step_torch(self):
        return torch.tensor([self._time_step], device=cpu_device, dtype=self._get_dtype())

    @property
    def substeps(self):
        return self._substeps

    @substeps.setter
    def substeps(self, s):
        if isinstance(s, str) and s.upper() == "ADAPTIVE":
            raise NotImplementedError("adaptive substeps are not part of this build")
        if not isinstance(s, numbers.Integral):
            raise TypeError("substeps must be integral type or 'ADAPTIVE'.")
        if not s > 0:
            raise ValueError("substeps must be positive.")
        self._substeps = s

    @property
    def corrector_steps(self):
        return self._corrector_steps

    @corrector_steps.setter
    def corrector_steps(self, c):
        if not isinstance(c, numbers.Integral):
            raise TypeError("corrector_steps must be integral type.")
        if c < 0:
            raise ValueError("corrector_steps must not be negative.")
        self._corrector_steps = c

    @property
    def advection_tol(self):
        if self._advection_tol is None:
            return PISOtorch_diff._get_solver_tolerance(dtype=self._get_dtype())
        return self._advection_tol

    @advection_tol.setter
    def advection_tol(self, t):
        if t is not None and not (isinstance(t, numbers.Real) and t > 0):
            raise ValueError("advection_tol must be positive float or None.")
        self._advection_tol = t

    @property
    def pressure_tol(self):
        if self._pressure_tol is None:
            return PISOtorch_diff._get_solver_tolerance(dtype=self._get_dtype())
        return self._pressure_tol

    @pressure_tol.setter
    def pressure_tol(self, t):
        if t is not None and not (isinstance(t, numbers.Real) and t > 0):
            raise ValueError("pressure_tol must be positive float or None.")
        self._pressure_tol = t

    @property
    def velocity_corrector(self):
        return self._velocity_corrector

    @velocity_corrector.setter
    def velocity_corrector(self, v):
        if v not in self._velocity_corrector_versions:
            raise ValueError("velocity_corrector must be one of %s" % list(self._velocity_corrector_versions))
        self._velocity_corrector = v
        self._velocity_corrector_version = self._velocity_corrector_versions[v]

    # --- hooks ---
    @property
    def prep_fn(self):
        return self._prep_fn

    @prep_fn.setter
    def prep_fn(self, fn):
        if fn is not None and not isinstance(fn, dict):
            raise TypeError("prep_fn must be a dict {hook name: callable or list of callables}")
        self._prep_fn = fn

    def _run_prep_fn(self, name, **kwargs):
        if self._prep_fn is not None and name in self._prep_fn:
            fns = self._prep_fn[name]
            if not isinstance(fns, (list, tuple)):
                fns = [fns]
            for fn in fns:
                if fn is not None:
                    fn(**kwargs)

    def _check_stop(self):
        return bool(self.stop_fn())

    def reset_total_step(self):
        self.total_step = 0

    def reset_total_time(self):
        self.total_time = 0

    def end_step(self, time_step=0):
        self.total_step += 1
        self.total_time += float(ntonp(time_step)[0])

    def reset_step_counters(self):
        self.reset_total_step()
        self.reset_total_time()

    # --- linear solves (:861-903) ---
    def linear_solve_GPU(self, csrMat, rhs, x=None, use_BiCG=False, tol=None, max_iter=5000,
                         matrix_rank_deficient=False, residual_reset_step=0, return_best_result=False,
                         double_fallback=False, BiCG_with_preconditioner=True, BiCG_precondition_fallback=False):
        if self._backend is PISOtorch:
            maxit_torch = torch.IntTensor([max_iter])
            tol_torch = PISOtorch_diff._get_solver_tolerance_torch(tol, rhs.dtype)
            if x is None:
                x = torch.zeros_like(rhs)
            infos = PISOtorch_diff._linear_solve_wrapper(
                csrMat, rhs, x, maxit_torch, tol_torch, PISOtorch.ConvergenceCriterion.NORM2_NORMALIZED, use_BiCG,
                matrix_rank_deficient, residual_reset_step, False, False, return_best_result,
                double_fallback=double_fallback, debug_out=False, BiCG_with_preconditioner=BiCG_with_preconditioner,
                BiCG_precondition_fallback=BiCG_precondition_fallback)
            if self.record_solver_infos:
                self.solver_infos.append(("BiCG" if use_BiCG else "CG", infos))
            return x, True
        return self._backend.linear_solve_GPU(csrMat, rhs, transpose=False, use_BiCG=use_BiCG, tol=tol,
                                              max_iter=max_iter, return_best_result=return_best_result,
                                              double_fallback=double_fallback,
                                              BiCG_with_preconditioner=BiCG_with_preconditioner,
                                              BiCG_precondition_fallback=BiCG_precondition_fallback), True

    def linear_solve(self, csrMat, rhs, x=None, use_BiCG=False, tol=None, max_iter=None, matrix_rank_deficient=False,
                     residual_reset_step=0, use_scipy=False, return_best_result=False, refine_tol=None):
        if use_scipy:
            raise NotImplementedError("the scipy solve path is not part of this build")
        if max_iter is None:
            max_iter = self.linear_solve_max_iterations
        if refine_tol is not None and not self.differentiable:
            # the sweeps recompute the residual in FP64 themselves: the correction solves need no resets
            return self._linear_solve_refined(csrMat, rhs, x, use_BiCG, refine_tol, max_iter)
        return self.linear_solve_GPU(csrMat, rhs, x, use_BiCG, tol, max_iter, matrix_rank_deficient,
                                     residual_reset_step, return_best_result,
                                     double_fallback=self.solver_double_fallback,
                                     BiCG_with_preconditioner=self.preconditionBiCG,
                                     BiCG_precondition_fallback=self.BiCG_precondition_fallback)

    def _linear_solve_refined(self, csrMat, rhs, x, use_BiCG, refine_tol, max_iter, residual_reset_step=0):
        """FP64-refined solve: the FP64 solution rounded to the domain dtype."""
        x64 = torch.zeros(rhs.numel(), dtype=torch.float64, device=rhs.device) if x is None else \
            x.detach().to(torch.float64).contiguous()
        info = PISOtorch.SolveLinearRefined(csrMat, rhs.detach().contiguous(), x64, max_iter,
                                            self.pressure_refine_inner_tol, refine_tol,
                                            self.pressure_refine_max_sweeps, useBiCG=use_BiCG,
                                            residualResetSteps=residual_reset_step)
        if self.record_solver_infos:
            self.solver_infos.append(("BiCG" if use_BiCG else "CG", [info]))
        if not (info.converged and info.isFiniteResidual):
            raise PISOtorch_diff.LinsolveError("FP64-refined solve reached relative residual %.3e (target %.1e)"
                                               % (info.finalResidual, refine_tol))
        return x64.to(rhs.dtype), True

    def _advect_scalar(self, step, time_step, non_ortho_flags, reuse_result):
        """Passive-scalar advection of one step (PISOtorch_simulation.py:1085-1167): one advection matrix per
        channel when the channels differ (scalar viscosity per channel or boundary types per channel), else one
        matrix and one multi-right-hand-side solve.  Returns (whether that matrix serves the velocity advection too,
        solve_ok, whether the step ends here: a failed solve or the stop check of the non-orthogonal loop)."""
        domain = self.domain
        B = self._backend
        split = not (domain.isPassiveScalarViscosityStatic() and domain.isAllFixedBoundariesPassiveScalarTypeStatic())
        channels = domain.getPassiveScalarChannels()
        n = domain.getTotalSize()
        for_velocity = not (domain.hasPassiveScalarViscosity() or domain.hasBlockViscosity() or split)
        matrices = []
        if split:
            for channel in range(channels):
                B.SetupAdvectionMatrix(domain, time_step, non_ortho_flags, forPassiveScalar=True,
                                       passiveScalarChannel=channel)
                matrices.append(domain.C.clone())
        else:
            B.SetupAdvectionMatrix(domain, time_step, non_ortho_flags, forPassiveScalar=True, passiveScalarChannel=0)
        solve_ok = True
        no_grad = torch.no_grad() if self.exclude_advection_solve_gradients else nullcontext()
        if not self.non_orthogonal:
            B.SetupAdvectionScalar(domain, time_step, 0)
            self._run_prep_fn("POST_SCALAR_SETUP", domain=domain, no_step=0, local_step=step, time_step=time_step,
                              total_step=self.total_step, total_time=self.total_time)
            with no_grad:
                if split:
                    rhs = torch.split(domain.scalarRHS, n)
                    parts = []
                    for channel in range(channels):
                        r, solve_ok = self.linear_solve(matrices[channel], rhs[channel], x=None,
                                                        use_BiCG=self.advection_use_BiCG, tol=self.advection_tol)
                        parts.append(r)
                    scalar_result = torch.cat(parts)
                else:
                    scalar_result, solve_ok = self.linear_solve(domain.C, domain.scalarRHS, x=None,
                                                                use_BiCG=self.advection_use_BiCG,
                                                                tol=self.advection_tol)
            domain.setScalarResult(scalar_result)
            domain.UpdateDomainData()
        else:
            B.CopyScalarResultFromBlocks(domain)  # the explicit non-orthogonal RHS terms read it
            for no_step in range(self.advect_non_ortho_steps):
                B.SetupAdvectionScalar(domain, time_step, non_ortho_flags)
                self._run_prep_fn("POST_SCALAR_SETUP", domain=domain, no_step=no_step, local_step=step,
                                  time_step=time_step, total_step=self.total_step, total_time=self.total_time)
                with no_grad:
                    reuse = no_step > 0 and reuse_result
                    if split:
                        xs = torch.split(domain.scalarResult, n) if reuse else [None] * channels
                        rhs = torch.split(domain.scalarRHS, n)
                        parts = []
                        for channel in range(channels):
                            x = xs[channel].contiguous() if xs[channel] is not None else None
                            r, solve_ok = self.linear_solve(matrices[channel], rhs[channel], x=x,
                                                            use_BiCG=self.advection_use_BiCG, tol=self.advection_tol)
                            parts.append(r)
                        scalar_result = torch.cat(parts)
                    else:
                        x = domain.scalarResult if reuse else None
                        scalar_result, solve_ok = self.linear_solve(domain.C, domain.scalarRHS, x=x,
                                                                    use_BiCG=self.advection_use_BiCG,
                                                                    tol=self.advection_tol)
                domain.setScalarResult(scalar_result)
                domain.UpdateDomainData()
                if not solve_ok or self._check_stop():
                    return for_velocity, solve_ok, True
        B.CopyScalarResultToBlocks(domain)
        return for_velocity, solve_ok, False

    # --- the PISO step (:1054-1327) ---
    def _PISO_split_step(self, iterations, time_step=None, native=True):
        self._check_domain()
        solve_ok = True
        if time_step is None:
            time_step = self._get_time_step_torch()
        advect_use_prev_result = self.advect_reuse_result and not self.differentiable
        advect_non_ortho_reuse_result = self.advect_reuse_result and not self.differentiable
        pressure_reuse_result = self.pressure_reuse_result and not self.differentiable
        pressure_use_face_transform = False
        vcv = self._velocity_corrector_version
        if isinstance(iterations, torch.Tensor):
            iterations = iterations.numpy()[0]
        assert iterations > 0
        domain = self.domain
        non_ortho_flags = self._non_ortho_flags
        B = self._backend
        self.solver_infos = []
        if native and self._native_step_eligible():
            return self._native_steps(iterations, time_step)

        for step in range(iterations):
            if self.convergence_tol is not None:
                B.CopyVelocityResultFromBlocks(domain)
                last_vel = domain.velocityResult.clone().detach()
            self._run_prep_fn("PRE", domain=domain, local_step=step, time_step=time_step, total_step=self.total_step,
                              total_time=self.total_time)
            domain.UpdateDomainData()

            advection_matrix_for_velocity = False
            if self.advect_passive_scalar and domain.hasPassiveScalar():
                advection_matrix_for_velocity, solve_ok, stop = self._advect_scalar(
                    step, time_step, non_ortho_flags, advect_non_ortho_reuse_result)
                if not solve_ok or stop:
                    return solve_ok

            apply_pressure_gradient = False
            self._run_prep_fn("PRE_VELOCITY_SETUP", domain=domain, local_step=step, time_step=time_step,
                              total_step=self.total_step, total_time=self.total_time)
            if not advection_matrix_for_velocity:
                B.SetupAdvectionMatrix(domain, time_step, non_ortho_flags)
            if not self.non_orthogonal:
                B.SetupAdvectionVelocity(domain, time_step, 0, apply_pressure_gradient)
                self._run_prep_fn("POST_VELOCITY_SETUP", domain=domain, no_step=0, local_step=step,
                                  time_step=time_step, total_step=self.total_step, total_time=self.total_time)
                with torch.no_grad() if self.exclude_advection_solve_gradients else nullcontext():
                    x = None if (not advect_use_prev_result) else domain.velocityResult
                    velocityResult, solve_ok = self.linear_solve(domain.C, domain.velocityRHS, x=x,
                                                                 use_BiCG=self.advection_use_BiCG,
                                                                 tol=self.advection_tol)
                domain.setVelocityResult(velocityResult)
                domain.UpdateDomainData()
            else:
                B.CopyVelocityResultFromBlocks(domain)
                for no_step in range(self.advect_non_ortho_steps):
                    B.SetupAdvectionVelocity(domain, time_step, non_ortho_flags, apply_pressure_gradient)
                    self._run_prep_fn("POST_VELOCITY_SETUP", domain=domain, no_step=no_step, local_step=step,
                                      time_step=time_step, total_step=self.total_step, total_time=self.total_time)
                    with torch.no_grad() if self.exclude_advection_solve_gradients else nullcontext():
                        x = None if (no_step == 0 or not advect_non_ortho_reuse_result) else domain.velocityResult
                        velocityResult, solve_ok = self.linear_solve(domain.C, domain.velocityRHS, x=x,
                                                                     use_BiCG=self.advection_use_BiCG,
                                                                     tol=self.advection_tol)
                    domain.setVelocityResult(velocityResult)
                    domain.UpdateDomainData()
                    if not solve_ok or self._check_stop():
                        return solve_ok
            if not solve_ok:
                return solve_ok

            self._run_prep_fn("POST_PREDICTION", domain=domain, local_step=step, time_step=time_step,
                              total_step=self.total_step, total_time=self.total_time)
            if self._check_stop():
                break

            for cstep in range(self.corrector_steps):
                if not self.non_orthogonal:
                    B.SetupPressureCorrection(domain, time_step, 0, pressure_use_face_transform,
                                              timeStepNorm=self.pressure_time_step_normalized)
                    self._run_prep_fn("POST_PRESSURE_SETUP", domain=domain, local_step=step, time_step=time_step,
                                      total_step=self.total_step, total_time=self.total_time)
                    with torch.no_grad() if self.exclude_pressure_solve_gradients else nullcontext():
                        pressureResult, solve_ok = self.linear_solve(domain.P, domain.pressureRHSdiv, x=None,
                                                                     use_BiCG=self.pressure_use_BiCG,
                                                                     tol=self.pressure_tol,
                                                                     return_best_result=self.pressure_return_best_result,
                                                                     refine_tol=self.pressure_refine_tol)
                    if not solve_ok:
                        return solve_ok
                    if self.normalize_pressure_result:
                        pressureResult = B._subtract_mean(pressureResult)
                    domain.setPressureResult(pressureResult)
                    domain.UpdateDomainData()
                    self._run_prep_fn("POST_PRESSURE_RESULT", domain=domain, local_step=step, time_step=time_step,
                                      total_step=self.total_step, total_time=self.total_time)
                else:
                    B.SetupPressureMatrix(domain, time_step, non_ortho_flags, pressure_use_face_transform)
                    for pstep in range(self.pressure_non_ortho_steps):
                        if pstep == 0:
                            B.SetupPressureRHS(domain, time_step, non_ortho_flags, pressure_use_face_transform,
                                               timeStepNorm=self.pressure_time_step_normalized)
                        else:
                            B.SetupPressureRHSdiv(domain, time_step, non_ortho_flags, pressure_use_face_transform,
                                                  timeStepNorm=self.pressure_time_step_normalized)
                        self._run_prep_fn("POST_PRESSURE_SETUP", domain=domain, local_step=step, time_step=time_step,
                                          total_step=self.total_step, total_time=self.total_time)
                        with torch.no_grad() if self.exclude_pressure_solve_gradients else nullcontext():
                            x = None if (pstep == 0 or not pressure_reuse_result) else domain.pressureResult
                            pressureResult, solve_ok = self.linear_solve(
                                domain.P, domain.pressureRHSdiv, x=x, matrix_rank_deficient=False,
                                residual_reset_step=100, use_BiCG=self.pressure_use_BiCG, tol=self.pressure_tol,
                                return_best_result=self.pressure_return_best_result,
                                refine_tol=self.pressure_refine_tol)
                        if self.normalize_pressure_result:
                            pressureResult = B._subtract_mean(pressureResult)
                        domain.setPressureResult(pressureResult)
                        domain.UpdateDomainData()
                        self._run_prep_fn("POST_PRESSURE_RESULT", domain=domain, local_step=step,
                                          time_step=time_step, total_step=self.total_step, total_time=self.total_time)
                        if not solve_ok:
                            return solve_ok
                        if self._check_stop():
                            break
                self._run_prep_fn("POST_PRESSURE_NON_ORTHO", domain=domain, local_step=step, time_step=time_step,
                                  total_step=self.total_step, total_time=self.total_time)
                B.CopyPressureResultToBlocks(domain)
                B.CorrectVelocity(domain, time_step, version=vcv, timeStepNorm=self.pressure_time_step_normalized)
                self._run_prep_fn("POST_VELOCITY_CORRECTION", domain=domain, local_step=step, time_step=time_step,
                                  total_step=self.total_step, total_time=self.total_time)
                if self._check_stop():
                    break

            B.CopyVelocityResultToBlocks(domain)
            self._run_prep_fn("POST", domain=domain, local_step=step, time_step=time_step, total_step=self.total_step,
                              total_time=self.total_time)
            if self.convergence_tol is not None:
                step_max_diff = torch.max(torch.abs(last_vel - domain.velocityResult)).cpu().numpy()
                if step_max_diff < self.convergence_tol:
                    solve_ok = False
            if not solve_ok or self._check_stop():
                break
            self.end_step(time_step)
        return solve_ok

    # --- native step (pict_piso_step): the orthogonal branch above as one library call ---
    def _native_step_eligible(self):
        """True when _PISO_split_step would run exactly the stage sequence pict_piso_step queues: orthogonal,
        not differentiable, no passive scalar, no hooks, no early-stop checks, every solve from zero, no solver
        fallbacks and a finite-difference velocity corrector."""
        d = self.domain
        return (self.native_step and not self.differentiable and not self.non_orthogonal
                and not (self.advect_passive_scalar and d.hasPassiveScalar())
                and not self._prep_fn and self.stop_fn is _never_stop and self.convergence_tol is None
                and not self.advect_reuse_result and not self.pressure_reuse_result
                and not self.solver_double_fallback and not self.preconditionBiCG and self.pressure_refine_tol is None
                and self._velocity_corrector_version in (0, 1) and d.getDevice().type == "cuda")

    def _native_solve_opts(self, use_BiCG, tol, return_best):
        dtype = self._get_dtype()
        return PISOtorch._solve_opts(self.linear_solve_max_iterations,
                                     PISOtorch_diff._get_solver_tolerance(tol, dtype),
                                     PISOtorch.ConvergenceCriterion.NORM2_NORMALIZED, use_BiCG, False, 0, False,
                                     False, return_best, False, zero_rhs_shortcut=True)

    def _native_opts(self, time_step):
        from . import _abi
        o = _abi.StepOpts()
        o.time_step = float(ntonp(time_step)[0])
        o.corrector_steps = self.corrector_steps
        o.velocity_corrector = self._velocity_corrector_version
        o.time_step_norm = int(bool(self.pressure_time_step_normalized))
        o.normalize_pressure = int(bool(self.normalize_pressure_result))
        o.advection = self._native_solve_opts(self.advection_use_BiCG, self.advection_tol, False)
        o.pressure = self._native_solve_opts(self.pressure_use_BiCG, self.pressure_tol,
                                             self.pressure_return_best_result)
        return o

    def _defer_check(self, fn, o, adjoint, *args):
        """Queue an asynchronous native call (`fn`, its arguments before the ticket) and keep its ticket.
        When the process's pending-step slots are all taken, every Simulation holding some flushes first (its
        own errors are raised by its own flush)."""
        from . import _abi
        if _pending_steps_total() >= _MAX_PENDING_STEPS:
            first_err = None
            for sim in list(_deferring):
                try:
                    sim.flush_solver_checks()
                except Exception as e:  # collect every other Simulation's slots before raising
                    if sim is self and first_err is None:
                        first_err = e
                    elif sim is not self:
                        sim._flush_error = e
            if first_err is not None:
                raise first_err
        t = ctypes.c_int32(-1)
        _abi.check(fn(*args, ctypes.byref(t), PISOtorch._stream(self.domain.getDevice())),
                   "PISO step reverse pass" if adjoint else "PISO step")
        if not self._pending_checks and self not in _deferring:
            _deferring.append(self)
        self._pending_checks.append((t.value, o, adjoint, self.domain.getSpatialDims()))

    def flush_solver_checks(self):
        """Wait for the native calls queued with defer_solver_checks and check their solver results in order,
        raising the first error (LinsolveError, or a solver barrier error) the waiting calls would have raised.
        Whoever sets defer_solver_checks calls this before reading the solver results (Simulation.run does after
        every major step).  An error found while another Simulation freed the shared slots is raised here."""
        err, self._flush_error = getattr(self, "_flush_error", None), None
        if not self._pending_checks:
            if err is not None:
                raise err
            return
        from . import _abi
        lib = _abi.load_library()
        pending, self._pending_checks = self._pending_checks, []
        if self in _deferring:
            _deferring.remove(self)
        for ticket, o, adjoint, D in pending:
            infos = (_abi.SolveInfo * (D + o.corrector_steps))()
            st = lib.pict_step_wait(ticket, infos)
            if err is not None:
                continue
            try:
                _abi.check(st, "PISO step reverse pass" if adjoint else "PISO step")
                self._native_record(infos, o, None, adjoint)
            except Exception as e:  # every ticket is still collected
                err = e
        if err is not None:
            raise err

    def _native_record(self, infos, o, kms=None, adjoint=False):
        """Solver bookkeeping of one native step or reverse pass: solve log, recorded infos, LinsolveError."""
        domain = self.domain
        D, K = domain.getSpatialDims(), o.corrector_steps
        res = [PISOtorch.LinearSolverResultInfo(i.final_residual, i.used_iterations, i.converged,
                                                i.is_finite_residual) for i in infos]
        adv, prs = res[:D], res[D:]
        akind = "BiCG" if self.advection_use_BiCG else "CG"
        pkind = "BiCG" if self.pressure_use_BiCG else "CG"
        log = PISOtorch.solve_log
        if log is not None:
            es = domain.velocityRHS.element_size()
            n, nnz = domain.getTotalSize(), domain.C.getSize()
            aname = PISOtorch.solver_kernel_name(domain.C, D, o.advection)
            pname = PISOtorch.solver_kernel_name(domain.P, 1, o.pressure)
            log.append((akind, n, nnz, D, es, adv, kms[0] if kms else 0.0, adjoint, aname))
            for k in range(K):
                log.append((pkind, n, nnz, 1, es, [prs[k]], kms[1 + k] if kms else 0.0, adjoint, pname))
        if self.record_solver_infos:
            self.solver_infos.append((akind, adv))
            for k in range(K):
                self.solver_infos.append((pkind, [prs[k]]))
        PISOtorch_diff._check_solver_return_infos(adv, adjoint, self.advection_use_BiCG, o.advection.tolerance,
                                                  o.advection.max_iterations, False, is_FWD=not adjoint)
        for k in range(K):
            PISOtorch_diff._check_solver_return_infos([prs[k]], adjoint, self.pressure_use_BiCG, o.pressure.tolerance,
                                                      o.pressure.max_iterations, self.pressure_return_best_result,
                                                      is_FWD=not adjoint)

    def _native_steps(self, iterations, time_step, tape=None):
        """`iterations` native steps; `tape` (a pict_amd._abi.StepTape) records the last one for its reverse
        pass."""
        from . import _abi
        domain = self.domain
        D, K = domain.getSpatialDims(), self.corrector_steps
        o = self._native_opts(time_step)
        lib = _abi.load_library()
        timing = PISOtorch.solve_log is not None
        if self.defer_solver_checks and not timing:
            for _ in range(iterations):
                domain.UpdateDomainData()
                self._defer_check(lib.pict_piso_step_async, o, False, domain._atlas, ctypes.byref(o),
                                  ctypes.byref(tape) if tape is not None else None)
                self.end_step(time_step)
            return True
        self.flush_solver_checks()
        # a BiCGStab that fails inside the one-call step is retried with the ILU(0) preconditioner by the stage
        # path (PISOtorch_diff.py:276-286): the step's input velocity is kept to redo it there (the block pressure
        # is no input of an eligible step: its solves start from zero and the corrections write the pressure first)
        retry = tape is None and self.BiCG_precondition_fallback and (self.advection_use_BiCG or
                                                                      self.pressure_use_BiCG)
        if timing:
            _abi.check(lib.pict_set_step_timing(1))
        try:
            for _ in range(iterations):
                domain.UpdateDomainData()
                kept = [b.velocity.clone() for b in domain.getBlocks()] if retry else None
                infos = (_abi.SolveInfo * (D + K))()
                _abi.check(lib.pict_piso_step(domain._atlas, ctypes.byref(o),
                                              ctypes.byref(tape) if tape is not None else None, infos,
                                              PISOtorch._stream(domain.getDevice())), "PISO step")
                kms = None
                if timing:
                    ms = (ctypes.c_double * (1 + K))()
                    lib.pict_last_step_solve_ms(ms, 1 + K)
                    kms = list(ms)
                try:
                    self._native_record(infos, o, kms)
                except PISOtorch_diff.LinsolveError:
                    if kept is None:
                        raise
                    _LOG.warning("A solve of the native step failed; redoing the step stage by stage.")
                    for b, v in zip(domain.getBlocks(), kept):
                        b.velocity.copy_(v)
                    if not self._PISO_split_step(1, time_step, native=False):
                        return False
                    continue
                self.end_step(time_step)
        finally:
            if timing:
                _abi.check(lib.pict_set_step_timing(0))
        return True

    def run(self, iterations, static=False, log_domain=True):
        """Simulation.run (:1374-1518) without image/VTK output."""
        self._check_domain()
        if static:
            raise NotImplementedError("static (passive-scalar only) runs are not part of this build")
        domain = self.domain
        flux = float(domain.GetBoundaryFluxBalance().reshape(-1)[0])
        if abs(flux) > float(self.pressure_tol):
            _LOG.warning("Domain boundary is not divergence free (flux balance: %.03e). This can prevent pressure "
                         "solve convergence.", flux)
            return False
        sim_ok = True
        for it in range(1, iterations + 1):
            try:
                sim_ok = self._PISO_split_step(iterations=self.substeps)
                # deferred solver results of this major step are checked before it counts as done
                self.flush_solver_checks()
            except PISOtorch_diff.LinsolveError:
                _LOG.exception("Simulation failed in major step %d (total step %d):", it, self.total_step)
                break
            max_mag = tensor_as_np(domain.getMaxVelocityMagnitude(False))
            if np.isnan(max_mag).any():
                _LOG.warning("NaN encountered in velocity, stopping")
                sim_ok = False
                break
            if self.log_fn is not None:
                self.log_fn(domain=domain, out_dir=self.log_dir, it=it, out_it=0, total_step=self.total_step)
            if self._check_stop() or not sim_ok:
                break
        return sim_ok

    def make_divergence_free(self, iterations=1, max_iter=1000):
        """Projection of the block velocities onto divergence-free fields with A = 1 and time step 1
        (PISOtorch_simulation.py:985-1052)."""
        self._check_domain()
        pressure_reuse_result = self.pressure_reuse_result and not self.differentiable
        vcv = self._velocity_corrector_version
        if isinstance(iterations, torch.Tensor):
            iterations = iterations.numpy()[0]
        domain = self.domain
        B = self._backend
        non_ortho_flags = self._non_ortho_flags
        time_step = torch.tensor([1], device=cpu_device, dtype=domain.A.dtype)
        domain.setA(torch.ones_like(domain.A))
        solve_ok = True
        for step in range(iterations):
            self._run_prep_fn("PRE", domain=domain, local_step=step, time_step=time_step, total_step=self.total_step,
                              total_time=self.total_time)
            domain.UpdateDomainData()
            B.CopyVelocityResultFromBlocks(domain)
            domain.setPressureRHS(domain.velocityResult)
            domain.UpdateDomainData()
            B.SetupPressureMatrix(domain, time_step, non_ortho_flags, False)
            for pstep in range(self.pressure_non_ortho_steps):
                B.SetupPressureRHSdiv(domain, time_step, non_ortho_flags, False,
                                      timeStepNorm=self.pressure_time_step_normalized)
                with torch.no_grad() if self.exclude_pressure_solve_gradients else nullcontext():
                    x = None if (pstep == 0 or not pressure_reuse_result) else domain.pressureResult
                    pressureResult, solve_ok = self.linear_solve(domain.P, domain.pressureRHSdiv, x=x,
                                                                 use_BiCG=self.pressure_use_BiCG, tol=self.pressure_tol,
                                                                 return_best_result=self.pressure_return_best_result,
                                                                 max_iter=max_iter)
                if self.normalize_pressure_result:
                    pressureResult = B._subtract_mean(pressureResult)
                domain.setPressureResult(pressureResult)
                domain.UpdateDomainData()
                if not solve_ok:
                    return solve_ok
                if self._check_stop():
                    break
            B.CopyPressureResultToBlocks(domain)
            B.CorrectVelocity(domain, time_step, version=vcv, timeStepNorm=self.pressure_time_step_normalized)
            B.CopyVelocityResultToBlocks(domain)
            if not solve_ok or self._check_stop():
                return solve_ok
            self.end_step(time_step)
        return solve_ok
