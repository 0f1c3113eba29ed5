"""ctypes binding of libspral_ssids_hip.so (the C-ABI in include/).

The library is loaded from the package directory only; there is no fallback:
if it is missing, every call raises.
"""
import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libspral_ssids_hip.so"

i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class Options(C.Structure):
    """struct spral_ssids_hip_options."""
    _fields_ = [
        ("print_level", C.c_int),
        ("action", C.c_int),
        ("small", C.c_double),
        ("u", C.c_double),
        ("multiplier", C.c_double),
        ("pivot_method", C.c_int),
        ("failed_pivot_method", C.c_int),
        ("schur_precision", C.c_int),
        ("refine_max", C.c_int),
        ("refine_tol", C.c_double),
        ("timing", C.c_int),
        ("level_schedule", C.c_int),
    ]


class Stats(C.Structure):
    """struct spral_ssids_hip_stats."""
    _fields_ = [
        ("flag", C.c_int),
        ("num_delay", C.c_int),
        ("num_factor", C.c_int64),
        ("num_flops", C.c_int64),
        ("num_neg", C.c_int),
        ("num_two", C.c_int),
        ("num_zero", C.c_int),
        ("maxfront", C.c_int),
        ("maxsupernode", C.c_int),
        ("not_first_pass", C.c_int),
        ("not_second_pass", C.c_int),
        ("hip_error", C.c_int),
        ("num_levels", C.c_int),
        ("factor_ms", C.c_double),
        ("kernel_ms", C.c_double),
        ("kernel_flops", C.c_double),
        ("kernel_bytes", C.c_double),
        ("kernel_launches", C.c_int),
        ("pool_peak_bytes", C.c_int64),
        ("static_plan", C.c_int),
    ]


# name -> (restype, argtypes)
SIGNATURES = {
    "spral_ssids_hip_default_options": (None, [C.POINTER(Options)]),
    "spral_ssids_hip_analyse": (C.c_void_p, [C.c_int, i64p, i32p, i32p, C.c_int,
                                             C.POINTER(C.c_int)]),
    "spral_ssids_hip_analyse_info": (None, [C.c_void_p, C.POINTER(C.c_int),
                                            C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                            C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                            C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "spral_ssids_hip_analyse_arrays": (None, [C.c_void_p, i32p, i32p, i64p, i32p,
                                              i64p, i64p]),
    "spral_ssids_hip_analyse_free": (None, [C.c_void_p]),
    "spral_ssids_hip_order_grid_nd": (None, [C.c_int, C.c_int, C.c_int, i32p]),
    "spral_ssids_hip_order_min_degree": (C.c_int, [C.c_int, i64p, i32p, i32p]),
    "spral_ssids_hip_order_amd": (C.c_int, [C.c_int, i64p, i32p, i32p]),
    "spral_ssids_hip_create_symbolic": (C.c_void_p, [
        C.c_int, C.c_int, C.c_int, C.c_int, i32p, i32p, i64p, i32p, i64p, i64p,
        C.POINTER(Options), C.POINTER(C.c_int)]),
    "spral_ssids_hip_destroy_symbolic": (None, [C.c_void_p]),
    "spral_ssids_hip_factor": (C.c_void_p, [C.c_bool, C.c_void_p, f64p, C.c_int64,
                                            C.c_void_p, C.POINTER(Options),
                                            C.POINTER(Stats)]),
    "spral_ssids_hip_factor_device": (C.c_void_p, [C.c_bool, C.c_void_p, C.c_void_p,
                                                   C.c_int64, C.c_void_p,
                                                   C.POINTER(Options),
                                                   C.POINTER(Stats)]),
    "spral_ssids_hip_refactor_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p,
                                                  C.POINTER(Options),
                                                  C.POINTER(Stats)]),
    "spral_ssids_hip_destroy_numeric": (None, [C.c_void_p]),
    "spral_ssids_hip_solve_fwd": (C.c_int, [C.c_void_p, C.c_int, f64p, C.c_int]),
    "spral_ssids_hip_solve_diag": (C.c_int, [C.c_void_p, C.c_int, f64p, C.c_int]),
    "spral_ssids_hip_solve_diag_bwd": (C.c_int, [C.c_void_p, C.c_int, f64p, C.c_int]),
    "spral_ssids_hip_solve_bwd": (C.c_int, [C.c_void_p, C.c_int, f64p, C.c_int]),
    "spral_ssids_hip_solve_device": (C.c_int, [C.c_void_p, C.c_int, C.c_int,
                                               C.c_void_p, C.c_int, C.c_void_p]),
    "spral_hungarian_scale_sym": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_void_p, C.c_int,
                                            C.POINTER(C.c_int)]),  # raw pointers
    "spral_auction_scale_sym": (C.c_int, [C.c_int] + [C.c_void_p] * 5 +
                                [C.c_int, C.c_void_p, C.c_void_p, C.c_float,
                                 C.POINTER(C.c_int), C.POINTER(C.c_int),
                                 C.POINTER(C.c_int)]),
    "spral_match_order": (C.c_int, [C.c_int, i64p, i32p, f64p, i32p, f64p]),
    "spral_ssids_hip_save_numeric": (C.c_int, [C.c_void_p, C.c_char_p]),
    "spral_ssids_hip_load_numeric": (C.c_void_p, [C.c_void_p, C.c_char_p, C.c_void_p,
                                                  C.POINTER(C.c_int)]),
    "spral_ssids_hip_numeric_scaling": (C.c_int, [C.c_void_p, f64p]),
    "spral_ssids_hip_solve_device_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_int,
                                                       C.c_void_p, C.c_int, C.c_void_p]),
    "spral_ssids_hip_solve_workspace_doubles": (C.c_int64, [C.c_void_p, C.c_int]),
    "spral_ssids_hip_solve_refine": (C.c_int, [C.c_void_p, C.c_int, f64p, f64p, C.c_int,
                                               C.c_int, C.c_double, f64p,
                                               C.POINTER(C.c_int)]),
    "spral_ssids_hip_solve_refine_device": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p,
                                                      C.c_void_p, C.c_int, C.c_int,
                                                      C.c_double, f64p, C.POINTER(C.c_int),
                                                      C.c_void_p]),
    "spral_ssids_hip_enquire_indef": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "spral_ssids_hip_alter": (C.c_int, [C.c_void_p, f64p]),
    "spral_ssids_hip_enquire_posdef": (C.c_int, [C.c_void_p, f64p]),
    "spral_ssids_hip_get_contrib": (C.c_int, [C.c_void_p] + [C.c_void_p] * 8),
    "spral_ssids_hip_free_contrib": (None, [C.c_void_p]),
    "spral_ssids_hip_device_count": (C.c_int, []),
    "spral_ssids_hip_node_pivots": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "spral_ssids_hip_tree_levels": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                              C.c_void_p]),
    "spral_ssids_hip_node_factors": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                               C.c_void_p, C.c_void_p]),
    "spral_rb_peek": (C.c_int, [C.c_char_p] + [C.c_void_p] * 9),
    "spral_rb_write": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                 C.c_void_p, C.c_void_p, C.c_char_p, C.c_char_p,
                                 C.c_char_p]),
    "spral_rb_read": (C.c_void_p, [C.c_char_p, C.c_int, C.c_int, C.c_int]
                      + [C.c_void_p] * 6),
    "spral_rb_fetch": (None, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "spral_ssids_hip_equilib_scale_sym": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p,
                                                    C.c_void_p, C.c_void_p, C.c_int,
                                                    C.c_float, C.POINTER(C.c_int)]),
    "spral_ssids_hip_equilib_scale_sym_device": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p,
                                                           C.c_void_p, C.c_void_p, C.c_int,
                                                           C.c_float, C.POINTER(C.c_int)]),
}

_lib = None


def lib():
    """Load the library (once) and attach signatures.  Raises if absent."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is not built; run spral_amd.build.build() first")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            f = getattr(handle, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = args
        _lib = handle
    return _lib


def default_options() -> Options:
    o = Options()
    lib().spral_ssids_hip_default_options(C.byref(o))
    return o
