"""ctypes binding of ``include/ferrix_hip.h`` (``libferrix_hip.so``).

The library is loaded from this package directory only.  There is no CPU
fallback: if the library is missing or no gfx950 device is present, calls
raise.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libferrix_hip.so")

FX_OK = 0
FX_ERR_INVALID = 1
FX_ERR_OOM = 2
FX_ERR_HIP = 3
FX_ERR_SINGULAR_JACOBIAN = 4
FX_ERR_NOT_CONVERGED = 5
FX_ERR_BREAKDOWN = 6
FX_ERR_NO_DEVICE = 7

FX_PREC_FP64, FX_PREC_MIXED, FX_PREC_MIXED_POLY = 0, 1, 2
FX_STOP_ABS_SQRT_RZ, FX_STOP_REL_R2 = 0, 1
FX_LUT_MAX = 16

P = C.c_void_p
i32, i64, dbl = C.c_int32, C.c_int64, C.c_double


class fx_material_table(C.Structure):
    _fields_ = [("n_mat", i32), ("n_e", P), ("t_e", P), ("v_e", P), ("n_nu", P), ("t_nu", P), ("v_nu", P)]


class fx_solve_opts(C.Structure):
    _fields_ = [("precision", i32), ("stop_mode", i32), ("tol", dbl), ("max_iter", i64), ("inner_tol", dbl),
                ("max_refine", i32), ("spmv_group", i32), ("poly_degree", i32)]


class fx_solve_report(C.Structure):
    _fields_ = [("iterations", i64), ("outer_spmv", i64), ("refinements", i32), ("spmv_group", i32),
                ("final_criterion", dbl), ("seconds", dbl)]


class fx_amplitudes(C.Structure):
    _fields_ = [("n", i32), ("total_time", P), ("shift_x", P), ("shift_y", P), ("has_data", P), ("ptr", P),
                ("times", P), ("values", P)]


class fx_step_def(C.Structure):
    _fields_ = [("step_id", i32), ("nlgeom", i32), ("solver", i32),
                ("initial_time_increment", dbl), ("time_period", dbl), ("min_time_increment", dbl),
                ("max_time_increment", dbl), ("max_increments", i64),
                ("n_loads", i64), ("load_dof", P), ("load_val", P), ("load_amp", P), ("load_origin", P),
                ("n_bcs", i64), ("bc_dof", P), ("bc_val", P), ("bc_amp", P), ("bc_origin", P),
                ("amps", fx_amplitudes), ("solve_opts", fx_solve_opts), ("direct_opts", fx_solve_opts),
                ("on_increment", P), ("on_increment_user", P), ("solve", P), ("solve_user", P)]


# fx_increment_fn: int32_t (*)(void *user, int32_t step_id, int32_t inc_id, double global_time)
fx_increment_fn = C.CFUNCTYPE(C.c_int32, C.c_void_p, C.c_int32, C.c_int32, C.c_double)
# fx_solve_fn: int32_t (*)(void *user, fx_mesh *mesh, int64_t n, const double *rhs, double *x, int64_t *iterations)
fx_solve_fn = C.CFUNCTYPE(C.c_int32, C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_double),
                          C.POINTER(C.c_double), C.POINTER(C.c_int64))


class fx_timer(C.Structure):
    _fields_ = [("global_time", dbl), ("local_time", dbl), ("local_max", dbl)]


class fx_step_report(C.Structure):
    _fields_ = [("increments", i32), ("cutbacks", i32), ("nr_iterations", i32), ("linear_solves", i32),
                ("pcg_iterations", i64), ("outer_spmv", i64), ("last_rel_residual", dbl), ("seconds", dbl),
                ("seconds_assembly", dbl), ("seconds_solve", dbl), ("host_syncs", i64)]


# every symbol the header declares, with its signature
SIGNATURES = {
    "fx_ctx_create": (i32, [i32, P]),
    "fx_ctx_destroy": (None, [P]),
    "fx_ctx_memory": (i32, [P, P, P, P]),
    "fx_last_error": (C.c_char_p, [P]),
    "fx_version": (C.c_char_p, []),
    "fx_mesh_upload": (i32, [P, i64, P, P, i64, P, P, P, P, P]),
    "fx_mesh_destroy": (None, [P]),
    "fx_mesh_info": (i32, [P, P, P, P]),
    "fx_state_set_u": (i32, [P, P]),
    "fx_state_get_u": (i32, [P, P]),
    "fx_assemble": (i32, [P, i32, P, P]),
    "fx_internal_force": (i32, [P, i32, P]),
    "fx_matrix_get": (i32, [P, P, P, P]),
    "fx_stress_strain": (i32, [P, P, P]),
    "fx_apply_penalty": (i32, [P, i64, P, P, P, P]),
    "fx_solve_opts_default": (None, [P]),
    "fx_solve": (i32, [P, P, P, P, P]),
    "fx_solve_csr": (i32, [P, i64, P, P, P, P, P, P, P]),
    "fx_static_step": (i32, [P, P, P, P]),
    "fx_profile_spmv": (i32, [P, i32, i32, i32, P, P]),
    "fx_measure_fp64_peak": (i32, [P, i32, P, P]),
    "fx_profile_element": (i32, [P, i32, i32, P, P]),
    "fx_mesh_autotune_spmv": (i32, [P, i32, i32, P, P]),
    "fx_mesh_autotune_elements": (i32, [P, i32, i32, P, P, P]),
    "fx_batch_create": (i32, [P, P, P]),
    "fx_batch_destroy": (None, [P]),
    "fx_batch_add": (i32, [P, P, i32, P, P]),
    "fx_batch_add_system": (i32, [P, P, i32, i64, P, P, P, i64, P, P, P]),
    "fx_batch_run": (i32, [P, i64, P]),
    "fx_batch_poll": (i32, [P, i32, P, P]),
    "fx_batch_get": (i32, [P, i32, P]),
    "fx_batch_remove": (i32, [P, i32]),
    "fx_batch_time_spmv": (i32, [P, i32]),
    "fx_batch_finished": (i32, [P, P, i32, P]),
    "fx_batch_autotune": (i32, [P, i32, P, P, P]),
    "fx_batch_set_inner_shape": (i32, [P, i32]),
    "fx_batch_set_overlap": (i32, [P, i32]),
    "fx_batch_spmv_stats": (i32, [P, P, P, P, P]),
    "fx_solve_csr_noalloc": (i32, [P, i64, i64, P, P, P, P, P, P, P]),
}

_LIB = None


def load():
    """Loads the in-tree library (raises if it was not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m ferrix_amd.build`")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _LIB = lib
    return _LIB
