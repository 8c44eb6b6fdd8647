"""ctypes binding of libifea_amd.so (include/ifea_spd.h, include/ifea_step.h, include/ifea_remesh.h).

The library is built in-tree by ``impressionfea_amd.build``. There is no CPU fallback: ``lib()`` raises
``EngineUnavailableError`` when the shared object is missing, and the entry points return IFEA_ERR_NO_DEVICE when no
HIP device is present.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .errors import EngineUnavailableError

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libifea_amd.so")

# exported symbols, as declared in include/*.h
SPD_SYMBOLS = ["ifea_spd_create", "ifea_spd_name", "ifea_spd_analyze", "ifea_spd_analyze_perm", "ifea_spd_factorize",
               "ifea_spd_solve", "ifea_spd_factor_nnz", "ifea_spd_set_option", "ifea_spd_symbolic", "ifea_spd_stats",
               "ifea_spd_destroy", "ifea_spd_last_error", "ifea_symbolic_analyze", "ifea_spd_set_coordinates"]
MLP_SYMBOLS = ["ifea_mlp_create", "ifea_mlp_num_params", "ifea_mlp_set_params", "ifea_mlp_get_params",
               "ifea_mlp_forward", "ifea_mlp_loss_grad", "ifea_mlp_set_dataset", "ifea_mlp_train_epoch",
               "ifea_mlp_destroy"]
STEP_SYMBOLS = ["ifea_step_create", "ifea_step_set_params", "ifea_step_set_option", "ifea_step_set_perm",
                "ifea_step_prepare", "ifea_step_set_geometry", "ifea_step_prepare_device", "ifea_step_load_guess",
                "ifea_step_get_frames", "ifea_step_set_element_state", "ifea_step_assemble", "ifea_step_solve_velocity", "ifea_step_final",
                "ifea_step_advance", "ifea_step_element_results", "ifea_step_set_damage", "ifea_step_get_v",
                "ifea_step_set_v", "ifea_step_quality", "ifea_step_quality_noalloc", "ifea_step_info",
                "ifea_step_destroy"]

REMESH_SYMBOLS = ["ifea_locator_create", "ifea_locator_find", "ifea_locator_find_nearest", "ifea_locator_interpolate",
                  "ifea_locator_destroy", "ifea_stuff", "ifea_stuffing_sizes", "ifea_stuffing_get",
                  "ifea_stuffing_destroy", "ifea_smooth_interior", "ifea_remesh_last_error"]

IFEA_OK = 0
IFEA_ERR_ARG = -1
IFEA_ERR_PERM = -2
IFEA_ERR_SYMBOLIC = -3
IFEA_ERR_NOT_SPD = -4
IFEA_ERR_ORDER = -5
IFEA_ERR_DEVICE = -6
IFEA_ERR_NO_DEVICE = -7

IFEA_OPT_PRECISION = 1
IFEA_OPT_REFINE_TOL = 2
IFEA_OPT_REFINE_MAXIT = 3
IFEA_OPT_LEAF_BLOCKS = 4

IFEA_STEP_OPT_FIRST_SOLVE = 1


class StepParams(C.Structure):
    _fields_ = [
        ("C", C.c_double), ("n", C.c_double), ("m", C.c_double), ("eps0", C.c_double),
        ("friction", C.c_double), ("ram", C.c_double), ("kpen", C.c_double), ("rate_cut", C.c_double),
        ("u0", C.c_double), ("newton_blend", C.c_double), ("tolerance", C.c_double), ("pcg_tol", C.c_double),
        ("max_iterations", C.c_int), ("pcg_budget", C.c_int), ("quadratic_ls", C.c_int),
        ("adaptive_newton", C.c_int), ("precision", C.c_int), ("newton_limit", C.c_int),
        ("profile", C.c_int),
    ]


class StepStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int), ("converged", C.c_int), ("newton_iterations", C.c_int),
        ("factorizations", C.c_int), ("fp64_fallbacks", C.c_int), ("pcg_calls", C.c_int),
        ("pcg_iterations", C.c_int), ("solves", C.c_int), ("assemblies", C.c_int),
        ("last_rel", C.c_double), ("last_rnorm", C.c_double),
        ("ms_assembly", C.c_double), ("ms_factor", C.c_double), ("ms_solve", C.c_double), ("ms_total", C.c_double),
        ("first_solve_skips", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """Load the HIP library; raises EngineUnavailableError if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise EngineUnavailableError(
            f"{LIB_PATH} is missing: build it with `python -m impressionfea_amd.build` (hipcc, gfx950)")
    L = C.CDLL(LIB_PATH)
    vp, ip, dp, u8p = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_ubyte)
    L.ifea_spd_create.argtypes = [C.POINTER(vp)]
    L.ifea_spd_name.argtypes = [vp, C.c_char_p, C.c_int]
    L.ifea_spd_analyze.argtypes = [vp, C.c_int, ip, ip]
    L.ifea_spd_analyze_perm.argtypes = [vp, C.c_int, ip, ip, ip]
    L.ifea_spd_factorize.argtypes = [vp, dp]
    L.ifea_spd_solve.argtypes = [vp, dp, dp]
    L.ifea_spd_factor_nnz.restype = C.c_longlong
    L.ifea_spd_factor_nnz.argtypes = [vp]
    L.ifea_spd_set_option.argtypes = [vp, C.c_int, C.c_double]
    L.ifea_spd_symbolic.argtypes = [vp, ip, ip, ip]
    L.ifea_spd_stats.argtypes = [vp, dp]
    L.ifea_spd_destroy.argtypes = [vp]
    L.ifea_spd_destroy.restype = None
    L.ifea_spd_last_error.restype = C.c_char_p
    L.ifea_symbolic_analyze.argtypes = [C.c_int, ip, ip, ip, dp, ip, ip, ip, C.POINTER(C.c_longlong), dp]
    L.ifea_spd_set_coordinates.argtypes = [vp, C.c_int, dp]
    L.ifea_step_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int, ip, ip, ip, C.c_int, ip, ip]
    L.ifea_step_set_params.argtypes = [vp, C.POINTER(StepParams)]
    L.ifea_step_set_option.argtypes = [vp, C.c_int, C.c_double]
    L.ifea_step_set_perm.argtypes = [vp, ip]
    L.ifea_step_prepare.argtypes = [vp, dp, dp, ip, dp, ip, dp]
    L.ifea_step_set_geometry.argtypes = [vp, u8p, u8p, C.c_int, dp, C.c_int, dp]
    L.ifea_step_prepare_device.argtypes = [vp, dp, ip, dp]
    L.ifea_step_load_guess.argtypes = [vp, dp]
    L.ifea_step_get_frames.argtypes = [vp, dp, ip, ip, dp]
    L.ifea_step_set_element_state.argtypes = [vp, dp, dp]
    L.ifea_step_assemble.argtypes = [vp, C.c_int, C.c_int, dp, dp, dp, dp, dp]
    L.ifea_step_solve_velocity.argtypes = [vp, C.c_int, C.POINTER(StepStats)]
    L.ifea_step_final.argtypes = [vp, ip, dp, dp, dp, dp, u8p]
    L.ifea_step_advance.argtypes = [vp, C.c_double, dp, dp, dp]
    L.ifea_step_element_results.argtypes = [vp, dp, dp, dp]
    L.ifea_step_set_damage.argtypes = [vp, dp]
    L.ifea_step_get_v.argtypes = [vp, dp]
    L.ifea_step_set_v.argtypes = [vp, dp]
    L.ifea_step_quality.argtypes = [vp, dp, C.c_double, ip, dp]
    L.ifea_step_quality_noalloc.argtypes = [C.c_int, C.c_int, ip, dp, C.c_double, ip, dp]
    L.ifea_step_info.argtypes = [vp, C.c_int, dp]
    L.ifea_step_destroy.argtypes = [vp]
    L.ifea_step_destroy.restype = None
    L.ifea_locator_create.argtypes = [C.c_int, dp, C.c_int, ip, C.c_double, C.POINTER(vp)]
    L.ifea_locator_find.argtypes = [vp, C.c_int, dp, ip, dp]
    L.ifea_locator_find_nearest.argtypes = [vp, C.c_int, dp, ip]
    L.ifea_locator_interpolate.argtypes = [vp, C.c_int, dp, dp, dp]
    L.ifea_locator_destroy.argtypes = [vp]
    L.ifea_locator_destroy.restype = None
    L.ifea_stuff.argtypes = [vp, dp, dp, C.c_double, C.c_int, C.c_int, C.POINTER(vp)]
    L.ifea_stuffing_sizes.argtypes = [vp, ip, ip]
    L.ifea_stuffing_get.argtypes = [vp, dp, ip]
    L.ifea_stuffing_destroy.argtypes = [vp]
    L.ifea_stuffing_destroy.restype = None
    L.ifea_smooth_interior.argtypes = [C.c_int, C.c_int, ip, ip, ip, u8p, u8p, u8p, dp, C.c_double, C.c_int, ip]
    L.ifea_remesh_last_error.restype = C.c_char_p
    _lib = L
    return L


def last_error() -> str:
    return lib().ifea_spd_last_error().decode(errors="replace")


def ptr(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)
