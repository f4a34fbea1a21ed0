"""ctypes binding of libowtk_hip.so (the C ABI declared in include/owtk_hip.h).

The shared library is built in-tree by owt-krylov_amd/build.py and loaded from
this directory.  Importing this module never touches the GPU; loading fails
loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import pathlib

LIB_PATH = pathlib.Path(__file__).resolve().parent / "libowtk_hip.so"

OK = 0
ERROR_INVALID_ARGUMENT = -1
ERROR_ZERO_PIVOT = -2
ERROR_HIP = -3
ERROR_NO_DEVICE = -4
ERROR_OUT_OF_MEMORY = -5
ERROR_UNSUPPORTED = -6
ERROR_RUNTIME = -7

F64, F32, BF16 = 0, 1, 2
MATRIX_CSR, MATRIX_SPLIT, MATRIX_BSR = 0, 1, 2
PC_IDENTITY, PC_JACOBI, PC_SSOR, PC_ILU, PC_LEARNED = 0, 1, 2, 3, 4
ORDER_IDENTITY, ORDER_RCM, ORDER_HILBERT = 0, 1, 2
LAW_UNIFORM, LAW_LOGNORMAL_FIELD, LAW_EDGE_CONTRAST = 0, 1, 2


class OwtkError(RuntimeError):
    """Failure reported by libowtk_hip (non-zero status)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"owtk error {code}: {message}")
        self.code = code


class ZeroPivotError(OwtkError):
    """std::runtime_error("zero pivot in ILU(0)") of the reference."""


class NoDeviceError(OwtkError):
    pass


class UnsupportedError(OwtkError):
    pass


class InvalidArgumentError(OwtkError, ValueError):
    """std::invalid_argument of the reference."""


class MatrixInfo(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("block_size", C.c_int32),
        ("owned_nodes", C.c_int64), ("ghost_nodes", C.c_int64), ("entries", C.c_int64),
        ("device_bytes", C.c_int64), ("spmv_bytes_f64", C.c_int64), ("spmv_bytes_f32", C.c_int64),
        ("spmv_bytes_bf16", C.c_int64), ("has_f32", C.c_int32), ("has_bf16", C.c_int32),
    ]


class PrecondParams(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("storage", C.c_int32), ("omega", C.c_double),
        ("minimum_diagonal", C.c_double), ("block_nodes", C.c_int64), ("fill_level", C.c_int32),
        ("schedule", C.c_int32), ("learned_model", C.c_void_p),
    ]


class PrecondInfo(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("storage", C.c_int32), ("blocks", C.c_int64),
        ("forward_levels", C.c_int64), ("backward_levels", C.c_int64),
        ("factor_entries", C.c_int64), ("apply_bytes", C.c_int64), ("rows", C.c_int64),
        ("block_size", C.c_int64),
    ]


# int (*)(void* user, uint64_t iteration, const double* device_solution, size_t count)
CONVERGENCE_TEST = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint64, C.c_void_p, C.c_size_t)
REDUCTION_SERIAL, REDUCTION_COMPENSATED, REDUCTION_MIXED_PRECISION = 0, 1, 2


class SolverOptions(C.Structure):
    _fields_ = [
        ("relative_tolerance", C.c_double), ("absolute_tolerance", C.c_double),
        ("breakdown_tolerance", C.c_double), ("maximum_iterations", C.c_uint64),
        ("convergence_check_interval", C.c_uint64), ("restart", C.c_uint64),
        ("residual_replacement_interval", C.c_uint64), ("gmres_orthogonalization", C.c_int32),
        ("collect_timings", C.c_int32), ("idr_shadow_space", C.c_uint64),
        ("relaxation", C.c_double), ("verify_true_residual", C.c_int32), ("reduction", C.c_int32),
        ("convergence_test", CONVERGENCE_TEST), ("convergence_test_data", C.c_void_p),
    ]


class MultigridOptions(C.Structure):
    _fields_ = [
        ("pre_smoothing_steps", C.c_uint64), ("post_smoothing_steps", C.c_uint64),
        ("coarse_smoothing_steps", C.c_uint64), ("target_aggregate_size", C.c_uint64),
        ("jacobi_weight", C.c_double), ("smoother", C.c_int32), ("chebyshev_levels", C.c_uint64),
    ]


class SolverResult(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("breakdown_reason", C.c_int32),
        ("iterations", C.c_uint64), ("operator_applications", C.c_uint64),
        ("preconditioner_applications", C.c_uint64), ("global_reductions", C.c_uint64),
        ("reorthogonalizations", C.c_uint64), ("arnoldi_norm_verifications", C.c_uint64),
        ("initial_residual_norm", C.c_double), ("recursive_residual_norm", C.c_double),
        ("true_residual_norm", C.c_double), ("relative_residual_norm", C.c_double),
        ("solve_seconds", C.c_double), ("converged_by_application", C.c_int32), ("refinement_steps", C.c_uint64),
        ("inner_iterations", C.c_uint64), ("fallback_refinements", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class RefinementOptions(C.Structure):
    _fields_ = [
        ("outer", SolverOptions), ("inner_storage", C.c_int32), ("inner_restart", C.c_int32),
        ("inner_relative_tolerance", C.c_double), ("inner_maximum_iterations", C.c_uint64),
        ("maximum_refinements", C.c_uint64), ("fallback_ratio", C.c_double),
    ]


class BatchResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("breakdown_reason", C.c_int32), ("iterations", C.c_uint64),
                ("true_residual_norm", C.c_double), ("rhs_norm", C.c_double),
                ("relative_residual_norm", C.c_double), ("stop_ms", C.c_double)]


class GridSpec(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("block_size", C.c_int32), ("ordering", C.c_int32),
        ("seed", C.c_uint64), ("permute", C.c_int32), ("coefficient_law", C.c_int32),
        ("jitter", C.c_double), ("diagonal_shift", C.c_double), ("field_sigma", C.c_double),
        ("field_correlation", C.c_double),
    ]


_P = C.c_void_p
_I32P = C.POINTER(C.c_int32)
_I64P = C.POINTER(C.c_int64)
_DP = C.POINTER(C.c_double)

# name -> (restype, argtypes).  Every function of include/owtk_hip.h.
PROTOTYPES: dict[str, tuple] = {
    "owtk_last_error": (C.c_char_p, []),
    "owtk_version": (C.c_char_p, []),
    "owtk_policy_name": (C.c_char_p, []),
    "owtk_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "owtk_profile_enable": (C.c_int, [_P, C.c_int]),
    "owtk_profile_reset": (C.c_int, [_P]),
    "owtk_profile_region": (C.c_int, [_P, C.c_int, C.POINTER(C.c_double)]),
    "owtk_profile_read": (C.c_int, [_P, C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_double),
                                    C.POINTER(C.c_int64)]),
    "owtk_context_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "owtk_context_destroy": (C.c_int, [_P]),
    "owtk_context_synchronize": (C.c_int, [_P]),
    "owtk_malloc": (C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    "owtk_free": (C.c_int, [_P, _P]),
    "owtk_memcpy_htod": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "owtk_memcpy_dtoh": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "owtk_memcpy_dtod": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "owtk_fill_zero": (C.c_int, [_P, _P, C.c_size_t]),
    "owtk_convert": (C.c_int, [_P, C.c_size_t, C.c_int, _P, C.c_int, _P]),
    "owtk_matrix_create_csr": (C.c_int, [_P, C.c_int64, C.c_int64, C.c_int64, _P, _P, _P, C.POINTER(_P)]),
    "owtk_matrix_create_split": (C.c_int, [_P, C.c_int64, C.c_int64, C.c_int64, _P, _P, _P, _P, C.POINTER(_P)]),
    "owtk_matrix_create_bsr": (C.c_int, [_P, C.c_int64, C.c_int64, C.c_int64, _P, _P, _P, C.POINTER(_P)]),
    "owtk_matrix_update_values": (C.c_int, [_P, _P, _P]),
    "owtk_matrix_add_storage": (C.c_int, [_P, C.c_int]),
    "owtk_matrix_info_get": (C.c_int, [_P, C.POINTER(MatrixInfo)]),
    "owtk_matrix_destroy": (C.c_int, [_P]),
    "owtk_spmv": (C.c_int, [_P, C.c_int, _P, _P]),
    "owtk_residual": (C.c_int, [_P, _P, _P, _P]),
    "owtk_precond_params_default": (None, [C.POINTER(PrecondParams), C.c_int32]),
    "owtk_precond_create": (C.c_int, [_P, C.POINTER(PrecondParams), C.POINTER(_P)]),
    "owtk_precond_update_values": (C.c_int, [_P]),
    "owtk_precond_apply": (C.c_int, [_P, _P, _P]),
    "owtk_precond_numeric_updates": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "owtk_precond_info_get": (C.c_int, [_P, C.POINTER(PrecondInfo)]),
    "owtk_precond_ilu_factors": (C.c_int, [_P, _I64P, _P, _P, _P, _P]),
    "owtk_precond_destroy": (C.c_int, [_P]),
    "owtk_solver_options_default": (None, [C.POINTER(SolverOptions)]),
    "owtk_bicgstab": (C.c_int, [_P, _P, _P, _P, C.POINTER(SolverOptions), C.POINTER(SolverResult)]),
    "owtk_pipelined_bicgstab": (C.c_int, [_P, _P, _P, _P, C.POINTER(SolverOptions), C.POINTER(SolverResult)]),
    "owtk_idrs": (C.c_int, [_P, _P, _P, _P, C.POINTER(SolverOptions), C.POINTER(SolverResult)]),
    "owtk_communication_hiding_bicgstab": (C.c_int, [_P, _P, _P, _P, C.POINTER(SolverOptions),
                                                     C.POINTER(SolverResult)]),
    "owtk_fgmres": (C.c_int, [_P, _P, _P, _P, C.POINTER(SolverOptions), C.POINTER(SolverResult)]),
    "owtk_fgmres_with_snapshot": (C.c_int, [_P, _P, _P, _P, C.POINTER(SolverOptions), _P,
                                            C.POINTER(SolverResult)]),
    "owtk_arnoldi_snapshot_create": (C.c_int, [_P, C.POINTER(_P)]),
    "owtk_arnoldi_snapshot_destroy": (C.c_int, [_P]),
    "owtk_arnoldi_snapshot_clear": (C.c_int, [_P]),
    "owtk_arnoldi_snapshot_columns": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "owtk_arnoldi_snapshot_h": (C.c_int, [_P, C.c_uint64, C.c_uint64, _DP]),
    "owtk_arnoldi_snapshot_basis": (C.c_int, [_P, C.c_uint64, C.POINTER(_P)]),
    "owtk_arnoldi_snapshot_begin_cycle": (C.c_int, [_P, _P, C.c_uint64]),
    "owtk_arnoldi_snapshot_record_column": (C.c_int, [_P, C.c_uint64, _DP, C.c_size_t]),
    "owtk_arnoldi_snapshot_finish_cycle": (C.c_int, [_P, C.POINTER(_P), C.c_uint64]),
    "owtk_harmonic_ggev": (C.c_int, [C.c_int, _DP, _DP, _DP, _DP, _DP, _DP]),
    "owtk_extract_harmonic_ritz": (C.c_int, [_P, _P, _P, C.POINTER(SolverResult), C.POINTER(C.c_uint64)]),
    "owtk_gcrodr": (C.c_int, [_P, _P, _P, _P, _P, C.POINTER(SolverOptions), _P, C.POINTER(SolverResult)]),
    "owtk_refinement_options_default": (None, [C.POINTER(RefinementOptions)]),
    "owtk_fgmres_refined": (C.c_int, [_P, _P, _P, _P, C.POINTER(RefinementOptions), C.POINTER(SolverResult)]),
    "owtk_fgmres_refined_fallback": (C.c_int, [_P, _P, _P, _P, _P, C.POINTER(RefinementOptions),
                                               C.POINTER(SolverResult)]),
    "owtk_dot": (C.c_int, [_P, C.c_int, C.c_size_t, _P, _P, _DP]),
    "owtk_norm": (C.c_int, [_P, C.c_int, C.c_size_t, _P, _DP]),
    "owtk_axpy": (C.c_int, [_P, C.c_size_t, C.c_double, _P, _P]),
    "owtk_axpy_dot": (C.c_int, [_P, C.c_size_t, C.c_double, _P, _P, _P, _DP]),
    "owtk_multidot": (C.c_int, [_P, C.c_size_t, C.c_size_t, C.POINTER(_P), _P, _DP]),
    "owtk_multidot_noalloc": (C.c_int, [_P, C.c_size_t, C.c_size_t, C.POINTER(_P), _P, _P]),
    "owtk_reduction_dot": (C.c_int, [_P, C.c_int32, C.c_int, C.c_size_t, _P, _P, _DP]),
    "owtk_reduction_norm": (C.c_int, [_P, C.c_int32, C.c_int, C.c_size_t, _P, _DP]),
    "owtk_batch_create": (C.c_int, [_P, C.c_int32, C.c_int32, _P, _P, _P, _P, C.POINTER(_P)]),
    "owtk_batch_destroy": (C.c_int, [_P]),
    "owtk_batch_set_tile": (C.c_int, [_P, C.c_int32]),
    "owtk_batch_autotune": (C.c_int, [_P, _P, C.c_int32, C.POINTER(C.c_int32), _DP]),
    "owtk_batch_set_group_limit": (C.c_int, [_P, C.c_int64, C.c_int32]),
    "owtk_batch_autotune_groups": (C.c_int, [_P, _P, _P, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                             _DP]),
    "owtk_batch_bicgstab": (C.c_int, [_P, _P, _P, C.POINTER(SolverOptions), C.c_int32, C.POINTER(BatchResult)]),
    "owtk_read_block_csr": (C.c_int, [C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                      C.POINTER(C.c_int64), _P, _P, _P]),
    "owtk_write_block_csr": (C.c_int, [C.c_char_p, C.c_int64, C.c_int64, _P, _P, _P]),
    "owtk_write_ilu_factors": (C.c_int, [_P, C.c_char_p]),
    "owtk_recycle_space_create": (C.c_int, [_P, C.c_uint64, C.POINTER(_P)]),
    "owtk_recycle_space_destroy": (C.c_int, [_P]),
    "owtk_recycle_space_size": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "owtk_recycle_space_clear": (C.c_int, [_P]),
    "owtk_recycle_space_get": (C.c_int, [_P, C.c_int64, _P, _P]),
    "owtk_recycle_space_add_candidate": (C.c_int, [_P, _P, _P, C.c_double, C.POINTER(C.c_int32),
                                                   C.POINTER(SolverResult)]),
    "owtk_recycle_space_refresh": (C.c_int, [_P, _P, C.POINTER(SolverResult)]),
    "owtk_recycle_space_project": (C.c_int, [_P, _P, _P, C.POINTER(SolverResult)]),
    "owtk_recycled_fgmres": (C.c_int, [_P, _P, _P, _P, _P, C.POINTER(SolverOptions), C.c_int32,
                                       C.POINTER(SolverResult)]),
    "owtk_recycled_fgmres_with_snapshot": (C.c_int, [_P, _P, _P, _P, _P, C.POINTER(SolverOptions), C.c_int32, _P,
                                                     C.POINTER(SolverResult)]),
    "owtk_multigrid_options_default": (None, [C.POINTER(MultigridOptions)]),
    "owtk_multigrid_create": (C.c_int, [_P, C.POINTER(MultigridOptions), C.POINTER(_P)]),
    "owtk_multigrid_destroy": (C.c_int, [_P]),
    "owtk_multigrid_coarse": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), _P, _P, _P, _P]),
    "owtk_multigrid_solve": (C.c_int, [_P, _P, _P, C.POINTER(SolverOptions), C.c_int32, C.POINTER(SolverResult)]),
    "owtk_read_state": (C.c_int, [C.POINTER(C.c_char_p), C.c_int32, C.c_int64, C.c_int64, C.POINTER(C.c_int64),
                                  _P]),
    "owtk_gnn_create": (C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(_P)]),
    "owtk_gnn_destroy": (C.c_int, [_P]),
    "owtk_gnn_info": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
    "owtk_gnn_set_parameters": (C.c_int, [_P, _P]),
    "owtk_gnn_get_parameters": (C.c_int, [_P, _P]),
    "owtk_gnn_coefficients": (C.c_int, [_P, _P, _P]),
    "owtk_gnn_problem_create": (C.c_int, [_P, _P, C.c_int32, _P, _P, C.POINTER(_P)]),
    "owtk_gnn_problem_destroy": (C.c_int, [_P]),
    "owtk_gnn_zero_grad": (C.c_int, [_P]),
    "owtk_gnn_accumulate": (C.c_int, [_P, _P, C.c_double]),
    "owtk_gnn_read": (C.c_int, [_P, _DP, _P]),
    "owtk_gnn_adam_step": (C.c_int, [_P, C.c_double, C.c_double, C.c_double, C.c_double, _DP]),
    "owtk_grid_spec_default": (None, [C.POINTER(GridSpec), C.c_int32, C.c_int32, C.c_int32, C.c_uint64]),
    "owtk_grid_sizes": (C.c_int, [C.POINTER(GridSpec), _I64P, _I64P]),
    "owtk_grid_generate_split": (C.c_int, [C.POINTER(GridSpec), _P, _P, _P, _P, _P, _P, _P]),
    "owtk_grid_generate_bsr": (C.c_int, [C.POINTER(GridSpec), _P, _P, _P, _P]),
    "owtk_split_to_csr": (C.c_int, [C.c_int64, C.c_int64, _P, _P, _P, _P, _P, _P, _P]),
    "owtk_order_rcm": (C.c_int, [C.c_int64, _P, _P, _P]),
    "owtk_order_hilbert": (C.c_int, [C.c_int64, _P, _P]),
}

_lib = None


def lib() -> C.CDLL:
    """Load libowtk_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise OSError(f"{LIB_PATH} is missing: run `python owt-krylov_amd/build.py` first")
        handle = C.CDLL(str(LIB_PATH))
        for name, (restype, argtypes) in PROTOTYPES.items():
            fn = getattr(handle, name)
            fn.restype = restype
            fn.argtypes = argtypes
        _lib = handle
    return _lib


def check(code: int) -> None:
    if code == OK:
        return
    message = lib().owtk_last_error().decode(errors="replace")
    if code == ERROR_INVALID_ARGUMENT:
        raise InvalidArgumentError(code, message)
    if code == ERROR_ZERO_PIVOT:
        raise ZeroPivotError(code, message)
    if code == ERROR_NO_DEVICE:
        raise NoDeviceError(code, message)
    if code == ERROR_UNSUPPORTED:
        raise UnsupportedError(code, message)
    raise OwtkError(code, message)
