"""ctypes binding of the engine's C-ABI (include/lblg.h).

The shared library is nasoq_amd/liblblg.so, built in-tree by
nasoq_amd/build.py.  Loading fails loudly when it is missing: there is no
fallback implementation.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblblg.so")

_i32p = ctypes.POINTER(ctypes.c_int)
_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)
_f32p = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p


class Settings(ctypes.Structure):
    """Mirror of lblg_settings (field order must match include/lblg.h)."""
    _fields_ = [
        ("ldl_variant", ctypes.c_int),
        ("req_ref_iter", ctypes.c_int),
        ("max_inner_iter", ctypes.c_int),
        ("tol_abs", ctypes.c_double),
        ("tol_rel", ctypes.c_double),
        ("reg_diag", ctypes.c_double),
        ("regularization", ctypes.c_int),
        ("cost_param", ctypes.c_int),
        ("level_param", ctypes.c_int),
        ("final_seq_node", ctypes.c_int),
        ("n_relax", ctypes.c_int * 3),
        ("z_relax", ctypes.c_double * 3),
        ("ordering", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("target_bwd_err", ctypes.c_double),
        ("max_refine_rounds", ctypes.c_int),
    ]


class QPSettings(ctypes.Structure):
    """Mirror of lblg_qp_settings."""
    _fields_ = [
        ("diag_perturb", ctypes.c_double),
        ("eps_abs", ctypes.c_double),
        ("stop_tol", ctypes.c_double),
        ("max_iter", ctypes.c_int),
        ("max_iter_nas", ctypes.c_int),
        ("variant", ctypes.c_int),
        ("ordering", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("separate_rows", ctypes.c_int),
    ]


class QPInfo(ctypes.Structure):
    """Mirror of lblg_qp_info."""
    _fields_ = [
        ("status", ctypes.c_int),
        ("num_iter", ctypes.c_int),
        ("n_active", ctypes.c_int),
        ("inner_iter_ref", ctypes.c_int),
        ("outer_iter_ref", ctypes.c_int),
        ("diag_perturb", ctypes.c_double),
        ("stop_tol", ctypes.c_double),
        ("primal_obj", ctypes.c_double),
        ("dual_obj", ctypes.c_double),
        ("objective", ctypes.c_double),
        ("lag_res", ctypes.c_double),
        ("cons_sat_norm", ctypes.c_double),
        ("non_negativity_infn", ctypes.c_double),
        ("complementarity_infn", ctypes.c_double),
        ("factorizations", ctypes.c_int64),
        ("kkt_solves", ctypes.c_int64),
        ("factor_seconds", ctypes.c_double),
        ("solve_seconds", ctypes.c_double),
        ("refactored_supernodes", ctypes.c_int64),
        ("supernodes", ctypes.c_int64),
    ]


# (name, restype, argtypes) for every exported entry point of include/lblg.h
class Tuning(ctypes.Structure):
    _fields_ = [("group_w", ctypes.c_int), ("chain_w", ctypes.c_int), ("small_solve_nnz", ctypes.c_int),
                ("schur_big_m", ctypes.c_int), ("wide_sub", ctypes.c_int), ("factor_ms", ctypes.c_double),
                ("solve_ms", ctypes.c_double),
                ("schedules_tried", ctypes.c_int)]


SIGNATURES = [
    ("lblg_default_settings", None, [ctypes.POINTER(Settings)]),
    ("lblg_create", ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, _i32p, _i32p, _f64p]),
    ("lblg_destroy", ctypes.c_int, [_vp]),
    ("lblg_analyze", ctypes.c_int, [_vp, ctypes.POINTER(Settings), _i32p, _i32p]),
    ("lblg_set_settings", ctypes.c_int, [_vp, ctypes.POINTER(Settings)]),
    ("lblg_order_nd", ctypes.c_int, [ctypes.c_int, _i32p, _i32p, _i32p]),
    ("lblg_symbolic_scalar", ctypes.c_int64, [_vp, ctypes.c_int]),
    ("lblg_symbolic_ints", ctypes.c_int64, [_vp, ctypes.c_int, _i32p]),
    ("lblg_symbolic_longs", ctypes.c_int64, [_vp, ctypes.c_int, _i64p]),
    ("lblg_symbolic_flops", ctypes.c_int, [_vp, _f64p, _f64p]),
    ("lblg_device_count", ctypes.c_int, [_i32p]),
    ("lblg_last_error", ctypes.c_char_p, []),
    ("lblg_factor", ctypes.c_int, [_vp]),
    ("lblg_refactor_resident", ctypes.c_int, [_vp]),
    ("lblg_set_values", ctypes.c_int, [_vp, _f64p]),
    ("lblg_update_values", ctypes.c_int, [_vp, ctypes.c_int64, _i64p, _f64p]),
    ("lblg_refactored_supernodes", ctypes.c_int64, [_vp]),
    ("lblg_solve", ctypes.c_int, [_vp, ctypes.c_int, _f64p, ctypes.c_int64, _f64p, ctypes.c_int64, _f64p]),
    ("lblg_solve_device", ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int64, _vp, ctypes.c_int64, _f64p]),
    ("lblg_num_ref_iter", ctypes.c_int, [_vp]),
    ("lblg_residual_norms", ctypes.c_int, [_vp, ctypes.c_int, _f64p, ctypes.c_int64, _f64p, ctypes.c_int64,
                                            _f64p]),
    ("lblg_backward_error_device", ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int64, _vp, ctypes.c_int64,
                                                   _vp]),
    ("lblg_kkt_handle", _vp, [_vp]),
    ("lblg_backward_error_device_noalloc", ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int64, _vp,
                                                           ctypes.c_int64, _vp]),
    ("lblg_synchronize", ctypes.c_int, [_vp]),
    ("lblg_stream", _vp, [_vp]),
    ("lblg_get_factor", ctypes.c_int, [_vp, _f64p, _f64p, _i32p, _i32p]),
    ("lblg_inertia", ctypes.c_int, [_vp, _i64p, _i64p, _i64p]),
    ("lblg_get_scaling", ctypes.c_int, [_vp, _f64p]),
    ("lblg_factor_info", ctypes.c_int, [_vp, _i32p, _f64p, _f64p]),
    ("lblg_pivot_stats", ctypes.c_int, [_vp, _i64p, _i64p]),
    ("lblg_panel_stats", ctypes.c_int, [_vp, _i64p, _i64p]),
    ("lblg_batch_create", ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, _i32p, _vp, _vp, _vp]),
    ("lblg_batch_analyze", ctypes.c_int, [_vp, _vp]),
    ("lblg_batch_info", ctypes.c_int, [_vp, _i32p, _i64p]),
    ("lblg_batch_backward_errors", ctypes.c_int, [_vp, ctypes.c_int, _f64p, ctypes.c_int64, _f64p, ctypes.c_int64,
                                                   _f64p]),
    ("lblg_selftest_mfma", ctypes.c_int, [_f64p, _f64p, _f64p, ctypes.c_int]),
    ("lblg_set_kernel_timing", ctypes.c_int, [_vp, ctypes.c_int]),
    ("lblg_set_kernel_timing_only", ctypes.c_int, [_vp, ctypes.c_char_p]),
    ("lblg_graph_stats", ctypes.c_int, [_vp, _i64p, _i64p]),
    ("lblg_get_tuning", ctypes.c_int, [_vp, ctypes.POINTER(Tuning)]),
    ("lblg_set_tuning", ctypes.c_int, [_vp, ctypes.POINTER(Tuning)]),
    ("lblg_autotune", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Tuning)]),
    ("lblg_kernel_time", ctypes.c_int, [_vp, ctypes.c_char_p, _f64p, _i64p]),
    ("lblg_qp_default_settings", None, [ctypes.POINTER(QPSettings)]),
    ("lblg_qp_create", ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, _i32p, _i32p, _f64p, _f64p, ctypes.c_int,
                                      _i32p, _i32p, _f64p, _f64p, ctypes.c_int, _i32p, _i32p, _f64p, _f64p]),
    ("lblg_qp_destroy", ctypes.c_int, [_vp]),
    ("lblg_qp_set_settings", ctypes.c_int, [_vp, ctypes.POINTER(QPSettings)]),
    ("lblg_qp_analyze", ctypes.c_int, [_vp]),
    ("lblg_qp_autotune", ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(Tuning)]),
    ("lblg_qp_solve", ctypes.c_int, [_vp, _i32p]),
    ("lblg_qp_solve_warm", ctypes.c_int, [_vp, _f64p, _f64p, _i32p, ctypes.c_int, _f64p, _i32p]),
    ("lblg_qp_solve_predicted", ctypes.c_int, [_vp, _i32p, ctypes.c_int, _i32p, _i32p]),
    ("lblg_qp_set_linear", ctypes.c_int, [_vp, _f64p, _f64p, _f64p]),
    ("lblg_mlp_create", ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_uint64, ctypes.c_int]),
    ("lblg_mlp_destroy", ctypes.c_int, [_vp]),
    ("lblg_mlp_num_params", ctypes.c_int64, [_vp]),
    ("lblg_mlp_get_params", ctypes.c_int, [_vp, _f32p]),
    ("lblg_mlp_set_params", ctypes.c_int, [_vp, _f32p]),
    ("lblg_mlp_gradients", ctypes.c_int, [_vp, _f32p, _f32p, ctypes.c_int, _f32p, _f32p]),
    ("lblg_mlp_train", ctypes.c_int, [_vp, _f32p, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                      ctypes.c_uint64, _f32p]),
    ("lblg_mlp_predict", ctypes.c_int, [_vp, _f32p, ctypes.c_int, _f32p]),
    ("lblg_selftest_mfma_bf16", ctypes.c_int, [_f32p, _f32p, _f32p]),
    ("lblg_qp_get", ctypes.c_int, [_vp, _f64p, _f64p, _f64p, _i32p]),
    ("lblg_qp_get_info", ctypes.c_int, [_vp, ctypes.POINTER(QPInfo)]),
    ("lblg_qp_super_kkt", ctypes.c_int64, [_vp, _i32p, _i32p, _f64p]),
    ("lblg_kkt_analyze", ctypes.c_int, [_vp, ctypes.POINTER(Settings)]),
    ("lblg_kkt_size", ctypes.c_int64, [_vp]),
    ("lblg_kkt_edit", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _i32p]),
    ("lblg_kkt_factor", ctypes.c_int, [_vp]),
    ("lblg_kkt_solve", ctypes.c_int, [_vp, _f64p, _f64p, _f64p]),
]

_lib = None


def load():
    """Load liblblg.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "nasoq_amd/liblblg.so is missing: build it with `python nasoq_amd/build.py` "
            "(there is no non-native fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def iptr(a):
    return a.ctypes.data_as(_i32p) if a is not None else None


def lptr(a):
    return a.ctypes.data_as(_i64p) if a is not None else None


def dptr(a):
    return a.ctypes.data_as(_f64p) if a is not None else None


def fptr(a):
    return a.ctypes.data_as(_f32p) if a is not None else None


def check(rc, what):
    if rc != 0:
        msg = ""
        if _lib is not None:
            m = _lib.lblg_last_error()
            msg = m.decode() if m else ""
        raise RuntimeError(f"{what} failed with status {rc} {msg}")
    return rc
