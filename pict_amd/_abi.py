"""ctypes mirror of include/pict_hip.h and the loader of libpict_hip.so.

The structures here must match the C header field for field; tests check the
sizes and offsets against a C probe compiled from the header.
"""
import ctypes
import os
from ctypes import POINTER, Structure, c_double, c_int32, c_void_p

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libpict_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(HERE), "include", "pict_hip.h")

PICT_OK = 0
PICT_ERR_INVALID = 1
PICT_ERR_HIP = 2
PICT_ERR_UNSUPPORTED = 3
PICT_ERR_DEVICE_SYNC = 4

PICT_DTYPE_F32 = 0
PICT_DTYPE_F64 = 1

PICT_BOUNDARY_CONNECTED = 20
PICT_BOUNDARY_PERIODIC = 21
PICT_BOUNDARY_FIXED = 30

PICT_MATMUL_TRANSPOSE_A = 1
PICT_MATMUL_INVERT_A = 2
PICT_MATMUL_TRANSPOSE_B = 4
PICT_MATMUL_INVERT_B = 8
PICT_MATMUL_TRANSPOSE_OUT = 16
PICT_MATMUL_INVERT_OUT = 32


class BoundaryDesc(Structure):
    _fields_ = [
        ("type", c_int32),
        ("connected_block", c_int32),
        ("axes", c_int32 * 4),
        ("velocity_static", c_int32),
        ("velocity_bc", c_int32),
        ("has_transform", c_int32),
        ("reserved", c_int32),
        ("size", c_int32 * 4),
        ("stride", c_int32 * 4),
        ("velocity", c_void_p),
        ("velocity_grad", c_void_p),
        ("transform", c_void_p),
        ("scalar_static", c_int32),
        ("scalar_bc_mask", c_int32),
        ("scalar", c_void_p),
        ("scalar_grad", c_void_p),
    ]


class BlockDesc(Structure):
    _fields_ = [
        ("global_offset", c_int32),
        ("csr_offset", c_int32),
        ("size", c_int32 * 4),
        ("stride", c_int32 * 4),
        ("viscosity_static", c_int32),
        ("velocity_source_static", c_int32),
        ("has_transform", c_int32),
        ("has_face_transform", c_int32),
        ("velocity", c_void_p),
        ("pressure", c_void_p),
        ("viscosity", c_void_p),
        ("velocity_source", c_void_p),
        ("transform", c_void_p),
        ("face_transform", c_void_p),
        ("velocity_grad", c_void_p),
        ("pressure_grad", c_void_p),
        ("viscosity_grad", c_void_p),
        ("velocity_source_grad", c_void_p),
        ("bounds", BoundaryDesc * 6),
        ("scalar", c_void_p),
        ("scalar_grad", c_void_p),
    ]


class DomainDesc(Structure):
    _fields_ = [
        ("num_dims", c_int32),
        ("num_blocks", c_int32),
        ("num_cells", c_int32),
        ("nnz", c_int32),
        ("dtype", c_int32),
        ("reserved", c_int32),
        ("viscosity", c_double),
        ("blocks", POINTER(BlockDesc)),
        ("csr_row", c_void_p),
        ("csr_index", c_void_p),
        ("C_value", c_void_p),
        ("P_value", c_void_p),
        ("A", c_void_p),
        ("velocity_rhs", c_void_p),
        ("velocity_result", c_void_p),
        ("pressure_rhs", c_void_p),
        ("pressure_rhs_div", c_void_p),
        ("pressure_result", c_void_p),
        ("viscosity_grad", c_void_p),
        ("C_grad", c_void_p),
        ("P_grad", c_void_p),
        ("A_grad", c_void_p),
        ("velocity_rhs_grad", c_void_p),
        ("velocity_result_grad", c_void_p),
        ("pressure_rhs_grad", c_void_p),
        ("pressure_rhs_div_grad", c_void_p),
        ("pressure_result_grad", c_void_p),
        ("scalar_channels", c_int32),
        ("scalar_viscosity_static", c_int32),
        ("scalar_viscosity", c_void_p),
        ("scalar_viscosity_grad", c_void_p),
        ("scalar_rhs", c_void_p),
        ("scalar_result", c_void_p),
        ("scalar_rhs_grad", c_void_p),
        ("scalar_result_grad", c_void_p),
    ]


class CsrDesc(Structure):
    _fields_ = [
        ("n", c_int32),
        ("nnz", c_int32),
        ("dtype", c_int32),
        ("reserved", c_int32),
        ("row", c_void_p),
        ("index", c_void_p),
        ("value", c_void_p),
        ("transpose_perm", c_void_p),
    ]


class SolveOpts(Structure):
    _fields_ = [
        ("max_iterations", c_int32),
        ("convergence_criterion", c_int32),
        ("use_bicg", c_int32),
        ("matrix_rank_deficient", c_int32),
        ("residual_reset_steps", c_int32),
        ("transpose_a", c_int32),
        ("print_residual", c_int32),
        ("return_best_result", c_int32),
        ("bicg_with_preconditioner", c_int32),
        ("zero_rhs_shortcut", c_int32),
        ("tolerance", c_double),
    ]


class SolveInfo(Structure):
    _fields_ = [
        ("final_residual", c_double),
        ("used_iterations", c_int32),
        ("converged", c_int32),
        ("is_finite_residual", c_int32),
        ("reserved", c_int32),
    ]


class StepOpts(Structure):
    _fields_ = [
        ("time_step", c_double),
        ("corrector_steps", c_int32),
        ("velocity_corrector", c_int32),
        ("time_step_norm", c_int32),
        ("normalize_pressure", c_int32),
        ("advection", SolveOpts),
        ("pressure", SolveOpts),
    ]


class StepTape(Structure):
    _fields_ = [
        ("C_value", c_void_p),
        ("A", c_void_p),
        ("P_value", c_void_p),
        ("velocity_star", c_void_p),
        ("pressure_rhs", c_void_p),
        ("pressure_raw", c_void_p),
        ("velocity_result", c_void_p),
    ]


class BatchConfig(Structure):
    _fields_ = [
        ("tiles_per_workgroup", c_int32),
        ("pressure_rows_per_thread", c_int32),
        ("advection_rows_per_workgroup", c_int32),
        ("adaptive_schedule", c_int32),
        ("advection_rows_per_thread", c_int32),
        ("reserved", c_int32),
    ]


class BatchInfo(Structure):
    _fields_ = [
        ("count", c_int32),
        ("tiles", c_int32),
        ("tiles_per_workgroup", c_int32),
        ("advection_grid", c_int32),
        ("advection_jobs", c_int32),
        ("advection_rows_per_workgroup", c_int32),
        ("advection_rows_per_thread", c_int32),
        ("pressure_grid", c_int32),
        ("pressure_jobs", c_int32),
        ("pressure_rows_per_thread", c_int32),
        ("graph", c_int32),
        ("advection_kernel", ctypes.c_char * 64),
        ("pressure_kernel", ctypes.c_char * 64),
    ]


# layout version of the descriptor structs (PICT_ABI_VERSION)
ABI_VERSION = 2

# exported symbol -> (restype, argtypes)
_P = c_void_p
SIGNATURES = {
    "pict_last_error": (ctypes.c_char_p, []),
    "pict_abi_version": (c_int32, []),
    "pict_atlas_create": (c_int32, [POINTER(DomainDesc), POINTER(c_void_p), _P]),
    "pict_atlas_update": (c_int32, [_P, POINTER(DomainDesc), _P]),
    "pict_atlas_destroy": (c_int32, [_P]),
    "pict_build_csr_pattern": (c_int32, [_P, _P]),
    "pict_setup_advection_matrix": (c_int32, [_P, c_double, c_int32, _P]),
    "pict_setup_advection_velocity": (c_int32, [_P, c_double, c_int32, c_int32, _P]),
    "pict_setup_pressure_matrix": (c_int32, [_P, c_double, c_int32, c_int32, _P]),
    "pict_setup_pressure_rhs": (c_int32, [_P, c_double, c_int32, c_int32, c_int32, _P]),
    "pict_setup_pressure_rhs_div": (c_int32, [_P, c_double, c_int32, c_int32, c_int32, _P]),
    "pict_setup_pressure_correction": (c_int32, [_P, c_double, c_int32, c_int32, c_int32, _P]),
    "pict_correct_velocity": (c_int32, [_P, c_double, c_int32, c_int32, _P]),
    "pict_compute_velocity_divergence": (c_int32, [_P, _P, _P]),
    "pict_compute_spatial_velocity_gradients": (c_int32, [_P, _P, _P]),
    "pict_sgs_viscosity_smagorinsky": (c_int32, [_P, c_double, _P, _P]),
    "pict_copy_velocity_result_to_blocks": (c_int32, [_P, _P]),
    "pict_setup_advection_matrix_scalar": (c_int32, [_P, c_double, c_int32, c_int32, _P]),
    "pict_setup_advection_scalar": (c_int32, [_P, c_double, c_int32, _P]),
    "pict_copy_scalar_result_to_blocks": (c_int32, [_P, _P]),
    "pict_copy_scalar_result_from_blocks": (c_int32, [_P, _P]),
    "pict_setup_advection_matrix_scalar_grad": (c_int32, [_P, c_double, c_int32, c_int32, _P]),
    "pict_setup_advection_scalar_grad": (c_int32, [_P, c_double, c_int32, _P]),
    "pict_copy_scalar_result_grad_from_blocks": (c_int32, [_P, _P]),
    "pict_copy_scalar_result_grad_to_blocks": (c_int32, [_P, _P]),
    "pict_copy_velocity_result_from_blocks": (c_int32, [_P, _P]),
    "pict_copy_pressure_result_to_blocks": (c_int32, [_P, _P]),
    "pict_copy_pressure_result_from_blocks": (c_int32, [_P, _P]),
    "pict_setup_advection_matrix_grad": (c_int32, [_P, c_double, c_int32, _P]),
    "pict_setup_advection_velocity_grad": (c_int32, [_P, c_double, c_int32, c_int32, _P]),
    "pict_setup_pressure_matrix_grad": (c_int32, [_P, c_double, c_int32, c_int32, _P]),
    "pict_setup_pressure_rhs_div_grad": (c_int32, [_P, c_double, c_int32, c_int32, c_int32, _P]),
    "pict_setup_pressure_rhs_grad": (c_int32, [_P, c_double, c_int32, c_int32, c_int32, _P]),
    "pict_setup_pressure_correction_grad": (c_int32, [_P, c_double, c_int32, c_int32, c_int32, _P]),
    "pict_correct_velocity_grad": (c_int32, [_P, c_double, c_int32, _P]),
    "pict_copy_velocity_result_grad_to_blocks": (c_int32, [_P, _P]),
    "pict_copy_velocity_result_grad_from_blocks": (c_int32, [_P, _P]),
    "pict_copy_pressure_result_grad_from_blocks": (c_int32, [_P, _P]),
    "pict_solve_linear": (c_int32, [POINTER(CsrDesc), _P, _P, c_int32, POINTER(SolveOpts), POINTER(SolveInfo), _P]),
    "pict_solve_linear_noalloc": (c_int32, [POINTER(CsrDesc), _P, _P, c_int32, POINTER(SolveOpts), POINTER(SolveInfo),
                                            _P]),
    "pict_solve_refined": (c_int32, [POINTER(CsrDesc), _P, _P, POINTER(SolveOpts), c_double, c_int32, c_int32,
                                     POINTER(SolveInfo), _P]),
    "pict_solver_kernel_name": (c_int32, [POINTER(CsrDesc), c_int32, POINTER(SolveOpts), ctypes.c_char_p, c_int32]),
    "pict_sparse_outer_product": (c_int32, [_P, _P, POINTER(CsrDesc), _P, _P]),
    "pict_csr_transpose_permutation": (c_int32, [POINTER(CsrDesc), _P, _P]),
    "pict_piso_step": (c_int32, [_P, POINTER(StepOpts), _P, POINTER(SolveInfo), _P]),
    "pict_piso_step_vjp": (c_int32, [_P, POINTER(StepOpts), POINTER(StepTape), _P, POINTER(SolveInfo), _P]),
    "pict_piso_step_async": (c_int32, [_P, POINTER(StepOpts), _P, POINTER(c_int32), _P]),
    "pict_piso_step_vjp_async": (c_int32, [_P, POINTER(StepOpts), POINTER(StepTape), _P, POINTER(c_int32), _P]),
    "pict_step_wait": (c_int32, [c_int32, POINTER(SolveInfo)]),
    "pict_subtract_mean": (c_int32, [_P, c_int32, c_int32, _P]),
    "pict_batch_create": (c_int32, [POINTER(c_void_p), c_int32, POINTER(c_void_p)]),
    "pict_batch_destroy": (c_int32, [_P]),
    "pict_batch_set_config": (c_int32, [_P, POINTER(BatchConfig)]),
    "pict_batch_get_config": (c_int32, [_P, POINTER(BatchConfig)]),
    "pict_batch_step": (c_int32, [_P, POINTER(StepOpts), POINTER(c_double), POINTER(SolveInfo), _P]),
    "pict_batch_tune": (c_int32, [_P, POINTER(StepOpts), POINTER(c_double), c_int32, POINTER(BatchConfig), _P]),
    "pict_batch_describe": (c_int32, [_P, POINTER(BatchInfo)]),
    "pict_batch_set_timing": (c_int32, [_P, c_int32]),
    "pict_batch_last_solve_ms": (c_int32, [_P, POINTER(c_double), c_int32]),
    "pict_set_step_timing": (c_int32, [c_int32]),
    "pict_last_step_solve_ms": (c_int32, [POINTER(c_double), c_int32]),
    "pict_set_kernel_timing": (c_int32, [c_int32]),
    "pict_last_solve_kernel_ms": (c_double, []),
    "pict_matmul": (c_int32, [_P, c_int32, _P, c_int32, _P, c_int32, ctypes.c_int64, c_int32, c_int32, _P]),
    "pict_matmul_grad": (c_int32, [_P, c_int32, _P, c_int32, _P, _P, _P, c_int32, ctypes.c_int64, c_int32, c_int32,
                                   _P]),
    "pict_vector_to_diag_matrix": (c_int32, [_P, _P, c_int32, ctypes.c_int64, c_int32, _P]),
    "pict_invert_matrix": (c_int32, [_P, _P, c_int32, ctypes.c_int64, c_int32, _P]),
    "pict_coords_to_transforms": (c_int32, [_P, c_int32, POINTER(c_int32), c_int32, _P, _P]),
    "pict_coords_to_face_transforms": (c_int32, [_P, c_int32, POINTER(c_int32), c_int32, _P, _P]),
}


class PictError(RuntimeError):
    """Error reported by libpict_hip (mirrors the reference's TORCH_CHECK RuntimeError)."""


class PictUnsupportedError(PictError, NotImplementedError):
    """Option of the reference API that libpict_hip does not implement."""


_lib = None


def load_library(path=None):
    """Load libpict_hip.so. Raises if it is missing: there is no fallback path."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise ImportError(
            f"libpict_hip.so not found at {path}; build it with `python -m pict_amd.build` "
            "(there is no CPU fallback for the PISO step)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.pict_abi_version() != ABI_VERSION:
        raise ImportError(f"{path} has descriptor layout {lib.pict_abi_version()}, this module expects "
                          f"{ABI_VERSION}; rebuild it with `python -m pict_amd.build`")
    if path == LIB_PATH:
        _lib = lib
    return lib


def check(status, what=""):
    if status == PICT_OK:
        return
    msg = load_library().pict_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if status == PICT_ERR_UNSUPPORTED:
        raise PictUnsupportedError(text)
    raise PictError(text)
