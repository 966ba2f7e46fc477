"""Python harness of the CPU oracle (pict_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.

The oracle reads the same pict_domain_desc as libpict_hip.  A CPU "twin" of a
PISOtorch.Domain (float64, host tensors) is packed with the product's own
descriptor packer and handed to the C restatement of the reference stages.
"""
import ctypes
import os
import subprocess

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")

_libs = {}


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib(bits=64):
    if bits not in _libs:
        path = os.path.join(BUILD, "liboracle_f%d.so" % bits)
        if not os.path.exists(path):
            build()
        _libs[bits] = declare(ctypes.CDLL(path))
    return _libs[bits]


def declare(L):
    """Declares the argument types of the oracle's entry points on a loaded library; returns it."""
    L.oracle_last_error.restype = ctypes.c_char_p
    for n in ("oracle_setup_advection_matrix",):
        getattr(L, n).argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
    L.oracle_setup_advection_velocity.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int]
    L.oracle_setup_pressure_matrix.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.oracle_setup_pressure_rhs_only.argtypes = [ctypes.c_void_p, ctypes.c_double]
    L.oracle_setup_pressure_rhs_div.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int]
    L.oracle_pressure_rhs_add_non_ortho.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.oracle_pressure_rhs_add_non_ortho_grad.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.oracle_correct_velocity.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
    L.oracle_correct_velocity_fvm.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int]
    L.oracle_correct_velocity_v4.argtypes = [ctypes.c_void_p, ctypes.c_double]
    L.oracle_compute_velocity_divergence.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.oracle_spatial_velocity_gradients.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.oracle_sgs_smagorinsky.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p]
    L.oracle_copy_velocity_result.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.oracle_copy_pressure_result.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.oracle_build_csr_pattern.argtypes = [ctypes.c_void_p]
    L.oracle_csr_row_end_closed_form.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                                 ctypes.POINTER(ctypes.c_int)]
    L.oracle_solve_linear.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                      ctypes.c_void_p, ctypes.c_void_p]
    L.oracle_sparse_outer_product.argtypes = [ctypes.c_void_p] * 4
    L.oracle_ilu0.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    L.oracle_pressure_rhs_add_non_ortho_face.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.oracle_pressure_rhs_add_non_ortho_grad_face.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.oracle_coords_to_transforms.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    L.oracle_coords_to_face_transforms.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                                   ctypes.c_void_p]
    L.oracle_set_threads.argtypes = [ctypes.c_int]
    for n in ("oracle_setup_advection_matrix_grad", "oracle_setup_advection_velocity_grad",
              "oracle_setup_pressure_rhs_div_grad", "oracle_correct_velocity_grad"):
        getattr(L, n).argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
    L.oracle_setup_pressure_matrix_grad.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.oracle_setup_pressure_rhs_only_grad.argtypes = [ctypes.c_void_p, ctypes.c_double]
    L.oracle_copy_velocity_result_grad.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.oracle_copy_pressure_result_grad_from_blocks.argtypes = [ctypes.c_void_p]
    for n in ("oracle_setup_advection_matrix_scalar", "oracle_setup_advection_matrix_scalar_grad"):
        getattr(L, n).argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int]
    for n in ("oracle_setup_advection_scalar", "oracle_setup_advection_scalar_grad"):
        getattr(L, n).argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
    L.oracle_copy_scalar_result.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    return L


def _bits(domain):
    return 64 if domain.getDtype() == torch.float64 else 32


def _chk(L, st, what):
    if st != 0:
        raise RuntimeError("oracle %s: %s" % (what, L.oracle_last_error().decode()))


def _desc(domain):
    d, arr = domain._pack()
    domain._oracle_keep = (d, arr)
    return ctypes.byref(d)


# ---------------------------------------------------------------------------
# CPU twin of a domain
# ---------------------------------------------------------------------------

def cpu_twin(domain, dtype=torch.float64):
    """A host-memory copy of `domain` (same structure) in `dtype`, prepared and with the CSR pattern built by the
    oracle.  Transforms are copied from the source domain."""
    from pict_amd import PISOtorch as P

    def cv(t):
        return None if t is None else t.detach().to(device="cpu", dtype=dtype).contiguous()

    channels = domain.getPassiveScalarChannels()
    d = P.Domain(domain.getSpatialDims(), domain.viscosity.to(dtype), domain.name + "_cpu", dtype, "cpu",
                 passiveScalarChannels=channels, _allow_cpu=True)
    if domain.hasPassiveScalarViscosity():
        d.setScalarViscosity(cv(domain.passiveScalarViscosity))
    mapping = {}
    for b in domain.blocks:
        nb = P.Block(d, b.name, velocity=cv(b.velocity), pressure=cv(b.pressure), passiveScalar=cv(b.passiveScalar))
        if b.hasTransform():
            nb.setTransform(cv(b.transform), cv(b.faceTransform))
        if b.viscosity is not None:
            nb.setViscosity(cv(b.viscosity))
        if b.velocitySource is not None:
            nb.setVelocitySource(cv(b.velocitySource))
        d.AddBlock(nb)
        mapping[id(b)] = nb
    for b in domain.blocks:
        nb = mapping[id(b)]
        for f, bd in enumerate(b.boundaries):
            if bd.type == P.BoundaryType.FIXED:
                fb = P.FixedBoundary(cv(bd.velocity), bd.velocityType, cv(bd.passiveScalar),
                                     list(bd.passiveScalarTypes) if bd.passiveScalarTypes else None,
                                     cv(bd.transform), d)
                fb._setSize(list(bd.getSizes()))
                nb.boundaries[f] = fb
            elif bd.type == P.BoundaryType.CONNECTED:
                nb.boundaries[f] = P.ConnectedBoundary(mapping[id(bd.getConnectedBlock())], list(bd.axes), d)
            else:
                nb.boundaries[f] = P.PeriodicBoundary(d)
    d.PrepareSolve()
    build_csr_pattern(d)
    return d


# ---------------------------------------------------------------------------
# stages (same names and arguments as PISOtorch)
# ---------------------------------------------------------------------------

def build_csr_pattern(domain):
    L = lib(_bits(domain))
    _chk(L, L.oracle_build_csr_pattern(_desc(domain)), "pattern")


def csr_row_end_closed_form(domain, block, flat):
    L = lib(_bits(domain))
    rs = ctypes.c_int()
    re = L.oracle_csr_row_end_closed_form(_desc(domain), block, flat, ctypes.byref(rs))
    return re, rs.value


def _dt(t):
    return float(t[0]) if isinstance(t, torch.Tensor) else float(t)


def SetupAdvectionMatrix(domain, timeStep, nonOrthoFlags=0, forPassiveScalar=False, passiveScalarChannel=0):
    L = lib(_bits(domain))
    if forPassiveScalar:
        _chk(L, L.oracle_setup_advection_matrix_scalar(_desc(domain), _dt(timeStep), int(nonOrthoFlags),
                                                       int(passiveScalarChannel)), "scalar advection matrix")
        return
    _chk(L, L.oracle_setup_advection_matrix(_desc(domain), _dt(timeStep), int(nonOrthoFlags)), "advection matrix")


def SetupAdvectionScalar(domain, timeStep, nonOrthoFlags=0):
    L = lib(_bits(domain))
    _chk(L, L.oracle_setup_advection_scalar(_desc(domain), _dt(timeStep), int(nonOrthoFlags)), "scalar rhs")


def CopyScalarResultToBlocks(domain):
    L = lib(_bits(domain))
    _chk(L, L.oracle_copy_scalar_result(_desc(domain), 1, 0), "copy")


def CopyScalarResultFromBlocks(domain):
    L = lib(_bits(domain))
    _chk(L, L.oracle_copy_scalar_result(_desc(domain), 0, 0), "copy")


def SetupAdvectionVelocity(domain, timeStep, nonOrthoFlags=0, applyPressureGradient=False):
    L = lib(_bits(domain))
    _chk(L, L.oracle_setup_advection_velocity(_desc(domain), _dt(timeStep), int(nonOrthoFlags),
                                              int(applyPressureGradient)), "advection rhs")


def SetupPressureMatrix(domain, timeStep, nonOrthoMode=0, useFaceTransform=False):
    """useFaceTransform: the coefficients from the faces' transforms (PISO_multiblock_cuda_kernel.cu:4857-4865, the
    non-orthogonal ones :1962-1972)."""
    L = lib(_bits(domain))
    _chk(L, L.oracle_setup_pressure_matrix(_desc(domain), int(nonOrthoMode), int(bool(useFaceTransform))),
         "pressure matrix")


def SetupPressureRHSdiv(domain, timeStep, nonOrthoMode=0, useFaceTransform=False, timeStepNorm=False):
    """Divergence of the pressure RHS, then its explicit non-orthogonal terms (PISO_multiblock_cuda_kernel.cu
    :5659-5704)."""
    L = lib(_bits(domain))
    _chk(L, L.oracle_setup_pressure_rhs_div(_desc(domain), _dt(timeStep), int(timeStepNorm),
                                            int(bool(useFaceTransform))), "pressure div")
    add = L.oracle_pressure_rhs_add_non_ortho_face if useFaceTransform else L.oracle_pressure_rhs_add_non_ortho
    _chk(L, add(_desc(domain), int(nonOrthoMode)), "pressure non-ortho rhs")


def PressureRHSaddNonOrtho(domain, nonOrthoMode):
    """pressureRHSdiv += the explicit non-orthogonal terms of the pressure Laplacian (k_pressureRHSaddNonOrthoComponents,
    PISO_multiblock_cuda_kernel.cu:5455-5477)."""
    L = lib(_bits(domain))
    _chk(L, L.oracle_pressure_rhs_add_non_ortho(_desc(domain), int(nonOrthoMode)), "pressure non-ortho rhs")


def SetupPressureRHS(domain, timeStep, nonOrthoMode=0, useFaceTransform=False, timeStepNorm=False):
    L = lib(_bits(domain))
    _chk(L, L.oracle_setup_pressure_rhs_only(_desc(domain), _dt(timeStep)), "pressure rhs")
    SetupPressureRHSdiv(domain, timeStep, nonOrthoMode, useFaceTransform, timeStepNorm)


def SetupPressureCorrection(domain, timeStep, nonOrthoMode=0, useFaceTransform=False, timeStepNorm=False):
    SetupPressureMatrix(domain, timeStep, nonOrthoMode)
    SetupPressureRHS(domain, timeStep, nonOrthoMode, useFaceTransform, timeStepNorm)


def CorrectVelocity(domain, timeStep, version=0, timeStepNorm=False):
    """The version switch of _CorrectVelocity (PISO_multiblock_cuda_kernel.cu:6163-6203): 4 the orthogonal flux
    form, 5 and 6 the finite-volume pressure gradient (cell fluxes / face metrics), every other version the
    finite-difference corrector."""
    L = lib(_bits(domain))
    version = int(version)
    if version == 4:
        st = L.oracle_correct_velocity_v4(_desc(domain), _dt(timeStep))
    elif version in (5, 6):
        st = L.oracle_correct_velocity_fvm(_desc(domain), _dt(timeStep), 0 if version == 5 else 3, int(timeStepNorm))
    else:
        st = L.oracle_correct_velocity(_desc(domain), _dt(timeStep), int(timeStepNorm))
    _chk(L, st, "correct velocity")


def ComputeVelocityDivergence(domain):
    L = lib(_bits(domain))
    out = torch.zeros(domain.getTotalSize(), dtype=domain.getDtype())
    _chk(L, L.oracle_compute_velocity_divergence(_desc(domain), ctypes.c_void_p(out.data_ptr())), "divergence")
    return out


def SpatialVelocityGradients(domain):
    """The spatial velocity gradients as one global vector [D, D, N]: [c, i, g] = d u_c / d x_i at cell g."""
    L = lib(_bits(domain))
    D, N = domain.getSpatialDims(), domain.getTotalSize()
    out = torch.zeros(D * D * N, dtype=domain.getDtype())
    _chk(L, L.oracle_spatial_velocity_gradients(_desc(domain), ctypes.c_void_p(out.data_ptr())), "velocity gradients")
    return out.view(D, D, N)


def SGSViscositySmagorinsky(domain, coefficient):
    """The Smagorinsky viscosity of every cell as one global vector [N]."""
    L = lib(_bits(domain))
    out = torch.zeros(domain.getTotalSize(), dtype=domain.getDtype())
    _chk(L, L.oracle_sgs_smagorinsky(_desc(domain), float(coefficient), ctypes.c_void_p(out.data_ptr())), "sgs")
    return out


def CopyVelocityResultToBlocks(domain):
    L = lib(_bits(domain))
    _chk(L, L.oracle_copy_velocity_result(_desc(domain), 1), "copy")


def CopyVelocityResultFromBlocks(domain):
    L = lib(_bits(domain))
    _chk(L, L.oracle_copy_velocity_result(_desc(domain), 0), "copy")


def CopyPressureResultToBlocks(domain):
    L = lib(_bits(domain))
    _chk(L, L.oracle_copy_pressure_result(_desc(domain), 1), "copy")


def CopyPressureResultFromBlocks(domain):
    L = lib(_bits(domain))
    _chk(L, L.oracle_copy_pressure_result(_desc(domain), 0), "copy")


# backward stages (same names and arguments as PISOtorch's *Grad functions)

def SetupAdvectionMatrixGrad(domain, timeStep, nonOrthoFlags=0, forPassiveScalar=False, passiveScalarChannel=0):
    L = lib(_bits(domain))
    if forPassiveScalar:
        _chk(L, L.oracle_setup_advection_matrix_scalar_grad(_desc(domain), _dt(timeStep), int(nonOrthoFlags),
                                                            int(passiveScalarChannel)), "scalar matrix grad")
        return
    _chk(L, L.oracle_setup_advection_matrix_grad(_desc(domain), _dt(timeStep), int(nonOrthoFlags)), "adv matrix grad")


def SetupAdvectionScalarGrad(domain, timeStep, nonOrthoFlags=0):
    L = lib(_bits(domain))
    _chk(L, L.oracle_setup_advection_scalar_grad(_desc(domain), _dt(timeStep), int(nonOrthoFlags)), "scalar rhs grad")


def CopyScalarResultGradFromBlocks(domain):
    L = lib(_bits(domain))
    _chk(L, L.oracle_copy_scalar_result(_desc(domain), 0, 1), "copy")


def CopyScalarResultGradToBlocks(domain):
    L = lib(_bits(domain))
    _chk(L, L.oracle_copy_scalar_result(_desc(domain), 1, 1), "copy")


def SetupAdvectionVelocityGrad(domain, timeStep, nonOrthoFlags=0, applyPressureGradient=False):
    L = lib(_bits(domain))
    _chk(L, L.oracle_setup_advection_velocity_grad(_desc(domain), _dt(timeStep), int(nonOrthoFlags)), "adv rhs grad")


def SetupPressureMatrixGrad(domain, timeStep, nonOrthoMode=0, useFaceTransform=False):
    """PISO_build_pressure_matrix_GRAD (PISO_multiblock_cuda_kernel.cu:4966-5115): it passes false for the face flag to
    its interpolation and scatter of the non-orthogonal coefficients, so the VJP is the cell-metric one either way."""
    L = lib(_bits(domain))
    _chk(L, L.oracle_setup_pressure_matrix_grad(_desc(domain), int(nonOrthoMode)), "pressure matrix grad")


def SetupPressureRHSdivGrad(domain, timeStep, nonOrthoMode=0, useFaceTransform=False, timeStepNorm=False):
    """k_computePressureRHSdivergenceFromFlux_GRAD (:5422-5452) scatters through the cell-metric fluxes whatever
    useFaceTransform says; the non-orthogonal terms' VJP uses the face coefficients (:5491)."""
    L = lib(_bits(domain))
    add = L.oracle_pressure_rhs_add_non_ortho_grad_face if useFaceTransform else L.oracle_pressure_rhs_add_non_ortho_grad
    _chk(L, add(_desc(domain), int(nonOrthoMode)), "non-ortho rhs grad")
    _chk(L, L.oracle_setup_pressure_rhs_div_grad(_desc(domain), _dt(timeStep), int(timeStepNorm)), "div grad")


def SetupPressureRHSGrad(domain, timeStep, nonOrthoMode=0, useFaceTransform=False, timeStepNorm=False):
    SetupPressureRHSdivGrad(domain, timeStep, nonOrthoMode, useFaceTransform, timeStepNorm)
    L = lib(_bits(domain))
    _chk(L, L.oracle_setup_pressure_rhs_only_grad(_desc(domain), _dt(timeStep)), "pressure rhs grad")


def SetupPressureCorrectionGrad(domain, timeStep, nonOrthoMode=0, useFaceTransform=False, timeStepNorm=False):
    SetupPressureRHSGrad(domain, timeStep, nonOrthoMode, useFaceTransform, timeStepNorm)
    SetupPressureMatrixGrad(domain, timeStep, nonOrthoMode, useFaceTransform)


def CorrectVelocityGrad(domain, timeStep, timeStepNorm=False):
    L = lib(_bits(domain))
    _chk(L, L.oracle_correct_velocity_grad(_desc(domain), _dt(timeStep), int(timeStepNorm)), "correct velocity grad")


def CopyVelocityResultGradToBlocks(domain):
    L = lib(_bits(domain))
    _chk(L, L.oracle_copy_velocity_result_grad(_desc(domain), 1), "copy grad")


def CopyVelocityResultGradFromBlocks(domain):
    L = lib(_bits(domain))
    _chk(L, L.oracle_copy_velocity_result_grad(_desc(domain), 0), "copy grad")


def CopyPressureResultGradFromBlocks(domain):
    L = lib(_bits(domain))
    _chk(L, L.oracle_copy_pressure_result_grad_from_blocks(_desc(domain)), "copy grad")


class SolveInfo:
    def __init__(self, s):
        self.finalResidual = s.final_residual
        self.usedIterations = s.used_iterations
        self.converged = bool(s.converged)
        self.isFiniteResidual = bool(s.is_finite_residual)

    def __repr__(self):
        return "SolveInfo(res=%g, it=%d, conv=%d, finite=%d)" % (self.finalResidual, self.usedIterations,
                                                                  self.converged, self.isFiniteResidual)


def SolveLinear(A, RHS, x, maxIterations=1000, tolerance=1e-8, convergenceCriterion=1, useBiCG=False,
                residualResetSteps=0, transposeA=False, returnBestResult=False, withPreconditioner=False):
    """CPU restatement of SolveLinear on host tensors (A: PISOtorch.CSRmatrix on CPU); withPreconditioner: the
    ILU(0)-preconditioned BiCGStab."""
    from pict_amd import _abi
    bits = 64 if A.value.dtype == torch.float64 else 32
    L = lib(bits)
    desc = A._desc()
    o = _abi.SolveOpts()
    o.max_iterations = int(maxIterations)
    o.convergence_criterion = int(convergenceCriterion)
    o.use_bicg = int(useBiCG)
    o.residual_reset_steps = int(residualResetSteps)
    o.transpose_a = int(transposeA)
    o.return_best_result = int(returnBestResult)
    o.bicg_with_preconditioner = int(bool(withPreconditioner))
    o.tolerance = float(tolerance)
    nb = RHS.size(0) // A.getRows()
    infos = (_abi.SolveInfo * nb)()
    _chk(L, L.oracle_solve_linear(ctypes.byref(desc), ctypes.c_void_p(RHS.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                                  nb, ctypes.byref(o), infos), "solve")
    return [SolveInfo(i) for i in infos]


def ILU0(A, transpose=False):
    """The ILU(0) factors of A (of A^T with transpose) on its effective pattern: L strictly below the diagonal (unit
    diagonal implied), U on and above, in the CSR order of A (of A^T's rows with ascending columns)."""
    bits = 64 if A.value.dtype == torch.float64 else 32
    L = lib(bits)
    out = torch.empty_like(A.value)
    _chk(L, L.oracle_ilu0(ctypes.byref(A._desc()), int(bool(transpose)), ctypes.c_void_p(out.data_ptr())), "ilu0")
    return out


def SparseOuterProduct(a, b, pattern):
    bits = 64 if a.dtype == torch.float64 else 32
    L = lib(bits)
    desc = pattern._desc()
    _chk(L, L.oracle_sparse_outer_product(ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()),
                                          ctypes.byref(desc), ctypes.c_void_p(pattern.value.data_ptr())), "outer")


def CoordsToTransforms(coords):
    dims = coords.dim() - 2
    vs = [coords.size(dims + 1 - d) for d in range(dims)] + [1] * (3 - dims)
    out = torch.zeros([1] + [vs[d] - 1 for d in reversed(range(dims))] + [2 * dims * dims + 1], dtype=coords.dtype)
    L = lib(64 if coords.dtype == torch.float64 else 32)
    c = coords.contiguous()
    _chk(L, L.oracle_coords_to_transforms(ctypes.c_void_p(c.data_ptr()), dims, (ctypes.c_int32 * 3)(*vs),
                                          ctypes.c_void_p(out.data_ptr())), "coords")
    return out


def CoordsToFaceTransforms(coords):
    dims = coords.dim() - 2
    vs = [coords.size(dims + 1 - d) for d in range(dims)] + [1] * (3 - dims)
    out = torch.zeros([1, dims] + [vs[d] for d in reversed(range(dims))] + [2 * dims * dims + 1], dtype=coords.dtype)
    L = lib(64 if coords.dtype == torch.float64 else 32)
    c = coords.contiguous()
    _chk(L, L.oracle_coords_to_face_transforms(ctypes.c_void_p(c.data_ptr()), dims, (ctypes.c_int32 * 3)(*vs),
                                               ctypes.c_void_p(out.data_ptr())), "face coords")
    return out


def csr_to_scipy(A):
    import scipy.sparse as sp
    n = A.getRows()
    return sp.csr_matrix((A.value.cpu().double().numpy(), A.index.cpu().numpy(), A.row.cpu().numpy()), shape=(n, n))


# ---------------------------------------------------------------------------
# one PISO step (PISOtorch_simulation.py:1054-1327, orthogonal branch)
# ---------------------------------------------------------------------------

def advect_scalar_step(domain, dt, flags, non_orthogonal, tol, max_iter, infos=None):
    """The driver's passive-scalar advection (PISOtorch_simulation.py:1085-1167, no result reuse): the advection
    matrix per channel when the channels differ (scalar viscosity per channel or boundary types per channel), else
    one matrix for all channels and one multi-right-hand-side BiCGStab; the non-orthogonal branch reads the scalar
    result (the block scalars copied in first).  Returns whether that matrix serves the velocity too (no scalar
    viscosity, no block viscosity, channels not split: the driver then skips the velocity matrix)."""
    n = domain.getTotalSize()
    channels = domain.getPassiveScalarChannels()
    split = not (domain.isPassiveScalarViscosityStatic() and domain.isAllFixedBoundariesPassiveScalarTypeStatic())
    reuse = not (domain.hasPassiveScalarViscosity() or domain.hasBlockViscosity() or split)
    mats = []
    if split:
        for ch in range(channels):
            SetupAdvectionMatrix(domain, dt, flags, True, ch)
            mats.append(domain.C.value.clone())
    else:
        SetupAdvectionMatrix(domain, dt, flags, True, 0)
    if non_orthogonal:
        CopyScalarResultFromBlocks(domain)
    SetupAdvectionScalar(domain, dt, flags if non_orthogonal else 0)
    x = torch.zeros_like(domain.scalarRHS)
    if split:
        keep = domain.C.value.clone()
        for ch in range(channels):
            domain.C.value.copy_(mats[ch])
            xs = torch.zeros(n, dtype=x.dtype)
            inf = SolveLinear(domain.C, domain.scalarRHS[ch * n:(ch + 1) * n].contiguous(), xs, max_iter, tol, 1,
                              useBiCG=True)
            x[ch * n:(ch + 1) * n] = xs
            if infos is not None:
                infos.append(("scalar", inf))
        domain.C.value.copy_(keep)
    else:
        inf = SolveLinear(domain.C, domain.scalarRHS, x, max_iter, tol, 1, useBiCG=True)
        if infos is not None:
            infos.append(("scalar", inf))
    domain.setScalarResult(x)
    CopyScalarResultToBlocks(domain)
    return reuse


def piso_step(domain, time_step, corrector_steps=2, advection_tol=1e-6, pressure_tol=1e-6, max_iter=5000,
              normalize_pressure=True, infos=None, non_orthogonal=False, non_ortho_flags=25, advect_scalar=True,
              corrector_version=1):
    """One PISO step of the orthogonal driver branch on a CPU twin, with the oracle's solvers (non_orthogonal: the
    driver's non-orthogonal branch); passive scalars are advected first when the domain has them; the velocity
    corrector of `corrector_version` (1 finite differences, 5 / 6 finite volumes with cell / face metrics)."""
    if non_orthogonal:
        return piso_step_non_ortho(domain, time_step, corrector_steps, advection_tol, pressure_tol, max_iter,
                                   normalize_pressure, infos, non_ortho_flags, advect_scalar=advect_scalar,
                                   corrector_version=corrector_version)
    dt = torch.tensor([time_step], dtype=domain.getDtype())
    reuse = False
    if advect_scalar and domain.hasPassiveScalar():
        reuse = advect_scalar_step(domain, dt, 0, False, advection_tol, max_iter, infos)
    if not reuse:
        SetupAdvectionMatrix(domain, dt, 0)
    SetupAdvectionVelocity(domain, dt, 0, False)
    x = torch.zeros_like(domain.velocityRHS)
    inf = SolveLinear(domain.C, domain.velocityRHS, x, max_iter, advection_tol, 1, useBiCG=True)
    if infos is not None:
        infos.append(("adv", inf))
    domain.setVelocityResult(x)
    for _ in range(corrector_steps):
        SetupPressureCorrection(domain, dt, 0)
        p = torch.zeros_like(domain.pressureRHSdiv)
        inf = SolveLinear(domain.P, domain.pressureRHSdiv, p, max_iter, pressure_tol, 1, useBiCG=False)
        if infos is not None:
            infos.append(("p", inf))
        if normalize_pressure:
            p = p - torch.mean(p)
        domain.setPressureResult(p.contiguous())
        CopyPressureResultToBlocks(domain)
        CorrectVelocity(domain, dt, corrector_version)
    CopyVelocityResultToBlocks(domain)
    domain._set_changed(False)


def piso_step_non_ortho(domain, time_step, corrector_steps=2, advection_tol=1e-6, pressure_tol=1e-6, max_iter=5000,
                        normalize_pressure=True, infos=None, flags=25, advect_non_ortho_steps=1,
                        pressure_non_ortho_steps=1, advect_scalar=True, corrector_version=1):
    """One PISO step of the driver's non-orthogonal branch (PISOtorch_simulation.py:1085-1300, no result reuse):
    passive scalars first (when the domain has them), then the explicit non-orthogonal terms of the advection RHS
    read the velocity result (the block velocities copied in first), those of the pressure RHS the pressure result
    left by the previous solve; the pressure CG recomputes its residual every 100 iterations."""
    dt = torch.tensor([time_step], dtype=domain.getDtype())
    reuse = False
    if advect_scalar and domain.hasPassiveScalar():
        reuse = advect_scalar_step(domain, dt, flags, True, advection_tol, max_iter, infos)
    if not reuse:
        SetupAdvectionMatrix(domain, dt, flags)
    CopyVelocityResultFromBlocks(domain)
    for _ in range(advect_non_ortho_steps):
        SetupAdvectionVelocity(domain, dt, flags, False)
        x = torch.zeros_like(domain.velocityRHS)
        inf = SolveLinear(domain.C, domain.velocityRHS, x, max_iter, advection_tol, 1, useBiCG=True)
        if infos is not None:
            infos.append(("adv", inf))
        domain.setVelocityResult(x)
    for _ in range(corrector_steps):
        SetupPressureMatrix(domain, dt, flags)
        for pstep in range(pressure_non_ortho_steps):
            if pstep == 0:
                SetupPressureRHS(domain, dt, flags)
            else:
                SetupPressureRHSdiv(domain, dt, flags)
            p = torch.zeros_like(domain.pressureRHSdiv)
            inf = SolveLinear(domain.P, domain.pressureRHSdiv, p, max_iter, pressure_tol, 1, useBiCG=False,
                              residualResetSteps=100)
            if infos is not None:
                infos.append(("p", inf))
            if normalize_pressure:
                p = p - torch.mean(p)
            domain.setPressureResult(p.contiguous())
        CopyPressureResultToBlocks(domain)
        CorrectVelocity(domain, dt, corrector_version)
    CopyVelocityResultToBlocks(domain)
    domain._set_changed(False)


# ---------------------------------------------------------------------------
# adjoint of one PISO step (reverse of piso_step, with the stage VJPs and adjoint solves)
# ---------------------------------------------------------------------------

def _zero_grads(domain):
    domain.CreateAGrad()
    domain.CGrad.CreateValue()
    domain.PGrad.CreateValue()
    domain.CreateViscosityGrad()
    domain.CreateVelocityGradOnBlocks()
    domain.CreateVelocitySourceGradOnBlocks()
    domain.CreateVelocityGradOnBoundaries()
    domain.CreatePressureGradOnBlocks()
    domain.CreateVelocityRHSGrad()
    domain.CreateVelocityResultGrad()
    domain.CreatePressureRHSGrad()
    domain.CreatePressureRHSdivGrad()
    domain.CreatePressureResultGrad()


def piso_step_vjp(domain, time_step, velocity_grad_out, corrector_steps=2, advection_tol=1e-12, pressure_tol=1e-12,
                  max_iter=5000):
    """One forward PISO step (orthogonal branch, as piso_step) followed by its reverse pass.

    velocity_grad_out: per-block gradients of a loss w.r.t. the step's output block velocities.
    Returns (velocity_grad_in per block, velocity_source_grad per block or None, viscosity_grad float): the gradient
    w.r.t. the step's input block velocities, the blocks' velocity sources and the domain viscosity, accumulated
    over every stage in the order PISOtorch_diff's autograd graph visits them (PISOtorch_diff.py:295-1234)."""
    dt = torch.tensor([time_step], dtype=domain.getDtype())
    dtf = float(time_step)
    N = domain.getTotalSize()
    D = domain.getSpatialDims()
    blocks = domain.blocks
    u_in = [b.velocity.clone() for b in blocks]
    # ---- forward, recording what the reverse pass needs ----
    SetupAdvectionMatrix(domain, dt, 0)
    SetupAdvectionVelocity(domain, dt, 0, False)
    C_adv, A_adv = domain.C.value.clone(), domain.A.clone()
    ustar = torch.zeros_like(domain.velocityRHS)
    SolveLinear(domain.C, domain.velocityRHS, ustar, max_iter, advection_tol, 1, useBiCG=True)
    domain.setVelocityResult(ustar.clone())  # CorrectVelocity overwrites the velocity result in place
    rec = []
    vr = ustar
    for _ in range(corrector_steps):
        SetupPressureCorrection(domain, dt, 0)
        prhs = domain.pressureRHS.clone()
        Pv = domain.P.value.clone()
        p = torch.zeros_like(domain.pressureRHSdiv)
        SolveLinear(domain.P, domain.pressureRHSdiv, p, max_iter, pressure_tol, 1, useBiCG=False)
        ph = (p - torch.mean(p)).contiguous()
        domain.setPressureResult(ph)
        CopyPressureResultToBlocks(domain)
        rec.append((vr.clone(), prhs, Pv, p.clone(), [b.pressure.clone() for b in blocks]))
        CorrectVelocity(domain, dt, 1)
        vr = domain.velocityResult.clone()
    CopyVelocityResultToBlocks(domain)
    u_out = [b.velocity.clone() for b in blocks]

    # ---- reverse ----
    g_u = [torch.zeros_like(u) for u in u_in]
    g_src = [torch.zeros_like(b.velocitySource) if b.velocitySource is not None else None for b in blocks]
    g_visc = 0.0
    g_A = torch.zeros(N, dtype=domain.getDtype())
    g_C = torch.zeros_like(C_adv)

    def collect(d):
        nonlocal g_visc
        for i, b in enumerate(d.blocks):
            g_u[i] += b.velocityGrad
            if g_src[i] is not None:
                g_src[i] += b.velocitySourceGrad
        g_visc += float(d.viscosityGrad[0])

    # CopyVelocityResultToBlocks: velocity result gradient = output block gradients
    g_vr = torch.cat([torch.cat([gv[0, c].reshape(-1) for gv in velocity_grad_out]) for c in range(D)]).clone()
    for vr_prev, prhs, Pv, p, pblocks in reversed(rec):
        # CorrectVelocity (inputs: A, block pressure, pressure RHS)
        domain.setPressureRHS(prhs)
        for b, pb in zip(blocks, pblocks):
            b.setPressure(pb)
        domain.setA(A_adv.clone())
        _zero_grads(domain)
        domain.setVelocityResultGrad(g_vr.clone())
        CorrectVelocityGrad(domain, dt)
        g_A += domain.AGrad
        g_prhs = domain.pressureRHSGrad.clone()
        # CopyPressureResultToBlocks, mean subtraction, pressure solve
        CopyPressureResultGradFromBlocks(domain)
        g_ph = domain.pressureResultGrad.clone()
        g_p = g_ph - torch.mean(g_ph)
        lam = torch.zeros_like(g_p)
        domain.P.setValue(Pv.clone())
        if bool((g_p != 0).any()):
            SolveLinear(domain.P, g_p.contiguous(), lam, max_iter, pressure_tol, 1, useBiCG=False, transposeA=True)
        outer = domain.P.WithZeroValue()
        SparseOuterProduct(lam, p, outer)
        gP = -outer.value
        # SetupPressureCorrection (inputs: velocity, source, velocity result, boundaries, viscosity, A, C)
        domain.setVelocityResult(vr_prev)
        domain.setPressureRHS(prhs)
        domain.C.setValue(C_adv.clone())
        domain.setA(A_adv.clone())
        for b, u in zip(blocks, u_in):
            b.setVelocity(u.clone())
        _zero_grads(domain)
        domain.PGrad.setValue(gP.contiguous())
        domain.setPressureRHSGrad(g_prhs)
        domain.setPressureRHSdivGrad(lam.clone())
        SetupPressureCorrectionGrad(domain, dt)
        collect(domain)
        g_A += domain.AGrad
        g_C += domain.CGrad.value
        g_vr = domain.velocityResultGrad.clone()
    # advection solve: C u* = rhs, one right-hand side per component
    lam = torch.zeros_like(g_vr)
    if bool((g_vr != 0).any()):
        domain.C.setValue(C_adv.clone())
        SolveLinear(domain.C, g_vr.contiguous(), lam, max_iter, advection_tol, 1, useBiCG=True, transposeA=True)
    for c in range(D):
        outer = domain.C.WithZeroValue()
        SparseOuterProduct(lam[c * N:(c + 1) * N].contiguous(), ustar[c * N:(c + 1) * N].contiguous(), outer)
        g_C -= outer.value
    # SetupAdvectionVelocity (rhs) and SetupAdvectionMatrix (A, C)
    for b, u in zip(blocks, u_in):
        b.setVelocity(u.clone())
    _zero_grads(domain)
    domain.setVelocityRHSGrad(lam.clone())
    SetupAdvectionVelocityGrad(domain, dt)
    collect(domain)
    _zero_grads(domain)
    domain.setAGrad(g_A.clone())
    domain.CGrad.setValue(g_C.clone())
    SetupAdvectionMatrixGrad(domain, dt)
    collect(domain)
    # leave the domain at the step's output state
    for b, u in zip(blocks, u_out):
        b.setVelocity(u)
    return g_u, g_src, g_visc


# ---------------------------------------------------------------------------
# per-cell vector / matrix operations (matrix_vector_ops.cu:7-303, matrix_vector_ops_grads.cu:8-310), FP64 numpy
# ---------------------------------------------------------------------------

def _cells(t):
    """[1, C, (Z), Y, X] -> numpy [cells, C] (float64)."""
    a = t.detach().cpu().double().numpy()
    return a.reshape(a.shape[1], -1).T.copy()


def _from_cells(arr, like, channels):
    """numpy [cells, channels] -> torch [1, channels, (Z), Y, X] in the dtype of `like`."""
    out = torch.from_numpy(np.ascontiguousarray(arr.T)).reshape([1, channels] + list(like.shape[2:]))
    return out.to(like.dtype)


def _inverse(m):
    """Per-cell adjugate / determinant of [cells, D, D] (transformations.h:418-482)."""
    D = m.shape[-1]
    if D == 1:
        return 1.0 / m
    if D == 2:
        det = m[:, 0, 0] * m[:, 1, 1] - m[:, 1, 0] * m[:, 0, 1]
        adj = np.stack([np.stack([m[:, 1, 1], -m[:, 0, 1]], -1), np.stack([-m[:, 1, 0], m[:, 0, 0]], -1)], -2)
        return adj * (1.0 / det)[:, None, None]
    # D == 3: cofactor C_ij = (-1)^(i+j) minor_ij; adjugate = C^T
    cof = np.empty_like(m)
    for i in range(3):
        for j in range(3):
            r = [k for k in range(3) if k != i]
            c = [k for k in range(3) if k != j]
            minor = m[:, r[0], c[0]] * m[:, r[1], c[1]] - m[:, r[0], c[1]] * m[:, r[1], c[0]]
            cof[:, i, j] = (-1) ** (i + j) * minor
    det = (m[:, 0, :] * cof[:, 0, :]).sum(-1)
    return np.transpose(cof, (0, 2, 1)) * (1.0 / det)[:, None, None]


def _operand(x, D, transpose, invert):
    m = x.reshape(-1, D, D)
    if transpose:
        m = np.transpose(m, (0, 2, 1))
    return _inverse(m) if invert else m


def matmul(vectorMatrixA, vectorMatrixB, transposeA=False, invertA=False, transposeB=False, invertB=False,
           transposeOutput=False, invertOutput=False):
    """PISOtorch.matmul: per cell A B; flags act on matrices only, the output flags on matrix x matrix only
    (matrix_vector_ops.cu:7-200)."""
    D = vectorMatrixA.dim() - 2
    A, B = _cells(vectorMatrixA), _cells(vectorMatrixB)
    am, bm = A.shape[1] == D * D, B.shape[1] == D * D
    if am and bm:
        out = np.einsum("nij,njk->nik", _operand(A, D, transposeA, invertA), _operand(B, D, transposeB, invertB))
        if invertOutput:
            out = _inverse(out)
        if transposeOutput:
            out = np.transpose(out, (0, 2, 1))
        return _from_cells(out.reshape(-1, D * D), vectorMatrixA, D * D)
    if am:
        out = np.einsum("nij,nj->ni", _operand(A, D, transposeA, invertA), B)
        return _from_cells(out, vectorMatrixA, D)
    if bm:
        out = np.einsum("ni,nik->nk", A, _operand(B, D, transposeB, invertB))
        return _from_cells(out, vectorMatrixA, D)
    return _from_cells((A * B).sum(-1, keepdims=True), vectorMatrixA, 1)


def matmulGrad(vectorMatrixA, vectorMatrixB, outputGrad, transposeA=False, invertA=False, transposeB=False,
               invertB=False, transposeOutput=False, invertOutput=False):
    """PISOtorch.matmulGrad (matrix_vector_ops_grads.cu:8-207): zero for an inverted operand; matrix x matrix gives
    zeros with an inverted output or with both operands inverted; a transposed operand's gradient is transposed."""
    D = vectorMatrixA.dim() - 2
    A, B, G = _cells(vectorMatrixA), _cells(vectorMatrixB), _cells(outputGrad)
    am, bm = A.shape[1] == D * D, B.shape[1] == D * D
    gA, gB = np.zeros_like(A), np.zeros_like(B)

    def put(m, transpose):
        return (np.transpose(m, (0, 2, 1)) if transpose else m).reshape(-1, D * D)

    if am and bm:
        if not (invertOutput or (invertA and invertB)):
            g = G.reshape(-1, D, D)
            if transposeOutput:
                g = np.transpose(g, (0, 2, 1))
            if not invertA:
                gA = put(np.einsum("nrk,nck->nrc", g, _operand(B, D, transposeB, invertB)), transposeA)
            if not invertB:
                gB = put(np.einsum("nkr,nkc->nrc", _operand(A, D, transposeA, invertA), g), transposeB)
    elif am:
        if not invertA:
            gA = put(np.einsum("nr,nc->nrc", G, B), transposeA)
        gB = np.einsum("nkc,nk->nc", _operand(A, D, transposeA, invertA), G)
    elif bm:
        gA = np.einsum("nrk,nk->nr", _operand(B, D, transposeB, invertB), G)
        if not invertB:
            gB = put(np.einsum("nr,nc->nrc", A, G), transposeB)
    else:
        gA, gB = B * G, A * G
    return [_from_cells(gA, vectorMatrixA, A.shape[1]), _from_cells(gB, vectorMatrixB, B.shape[1])]


def VectorToDiagMatrix(vectors):
    """matrix_vector_ops.cu:203-262."""
    D = vectors.dim() - 2
    V = _cells(vectors)
    out = np.zeros((V.shape[0], D, D))
    for k in range(D):
        out[:, k, k] = V[:, k]
    return _from_cells(out.reshape(-1, D * D), vectors, D * D)


def InvertMatrix(matrices):
    """matrix_vector_ops.cu:265-303."""
    D = matrices.dim() - 2
    return _from_cells(_inverse(_cells(matrices).reshape(-1, D, D)).reshape(-1, D * D), matrices, D * D)
