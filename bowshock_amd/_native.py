"""ctypes binding of libbowshock.so (include/bowshock.h).

The library is required: importing this module raises if it has not been
built (``python -m bowshock_amd.build``). There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libbowshock.so")

BS_OK, BS_ERR_RANGE, BS_ERR_NONPHYSICAL, BS_ERR_MEMORY, BS_ERR_EXECUTION, BS_ERR_PRECISION, BS_ERR_BUSY, \
    BS_ERR_UNSUPPORTED = range(8)
BS_PRECISION_FP32, BS_PRECISION_FP64 = 0, 1
BS_RELAX_FLUX_DIFFERENCE, BS_RELAX_LINEARIZED = 0, 1

# Every function include/bowshock.h declares, with its ctypes signature.
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_vp = C.c_void_p


class Boundary(C.Structure):
    _fields_ = [("kind", C.c_int32), ("rho", C.c_double), ("u", C.c_double), ("v", C.c_double), ("p", C.c_double)]


class Config(C.Structure):
    _fields_ = [("ni", C.c_int32), ("nj", C.c_int32), ("x", _dp), ("y", _dp), ("geometry", C.c_int32),
                ("gamma", C.c_double), ("flux", C.c_int32), ("limiter", C.c_int32),
                ("reconstruction", C.c_int32), ("integrator", C.c_int32), ("cfl", C.c_double),
                ("time_stepping", C.c_int32), ("boundaries", Boundary * 4)]


class Report(C.Structure):
    _fields_ = [("steps", C.c_int64), ("status", C.c_uint32), ("time", C.c_double), ("dt", C.c_double)]


SIGNATURES = {
    "bs_last_error": (C.c_char_p, []),
    "bs_version": (C.c_char_p, []),
    "bs_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "bs_ctx_destroy": (C.c_int, [_vp]),
    "bs_solver_create": (C.c_int, [_vp, C.POINTER(Config), C.POINTER(_vp)]),
    "bs_solver_create_fp64": (C.c_int, [_vp, C.POINTER(Config), C.POINTER(_vp)]),
    "bs_solver_precision": (C.c_int, [_vp, C.POINTER(C.c_int32)]),
    "bs_solver_destroy": (C.c_int, [_vp]),
    "bs_compute_metrics": (C.c_int, [C.c_int32, C.c_int32, _dp, _dp, C.c_int32] + [_dp] * 10),
    "bs_set_primitives": (C.c_int, [_vp, _dp]),
    "bs_set_state": (C.c_int, [_vp, _fp, _fp, _fp, _fp, C.c_int32]),
    "bs_read_state": (C.c_int, [_vp, _fp, _fp, _fp, _fp, C.POINTER(Report)]),
    "bs_set_state_f64": (C.c_int, [_vp, _dp, _dp, _dp, _dp, C.c_int32]),
    "bs_check_precision": (C.c_int, [_vp, _dp]),
    "bs_read_state_f64": (C.c_int, [_vp, _dp, _dp, _dp, _dp, C.POINTER(Report)]),
    "bs_set_face_states": (C.c_int, [_vp, C.c_int32, _dp, C.c_int32]),
    "bs_set_face_fluxes": (C.c_int, [_vp, C.c_int32, _dp, C.c_int32]),
    "bs_iterate": (C.c_int, [_vp, C.c_int32, C.POINTER(Report)]),
    "bs_advance_to": (C.c_int, [_vp, C.c_double, C.c_int32, C.POINTER(Report)]),
    "bs_residual_norm": (C.c_int, [_vp, _dp]),
    "bs_residual_vector": (C.c_int, [_vp, _dp]),
    "bs_wall_fluxes": (C.c_int, [_vp, C.c_int32, _dp, C.c_int32]),
    "bs_forebody_drag": (C.c_int, [_vp, C.c_int32, C.c_double, C.c_double, C.c_double, _dp]),
    "bs_set_implicit_cfl": (C.c_int, [_vp, C.c_double]),
    "bs_iterate_implicit": (C.c_int, [_vp, C.c_int32, C.c_double, C.c_int32, C.c_double, C.c_double,
                                      C.POINTER(Report)]),
    "bs_enable_graphs": (C.c_int, [_vp, C.c_int32]),
    "bs_set_relaxation": (C.c_int, [_vp, C.c_int32]),
    "bs_set_fitted_shock": (C.c_int, [_vp, _dp]),
    "bs_read_shock": (C.c_int, [_vp, _dp, _dp, _dp, C.c_int32]),
    "bs_set_line_grid": (C.c_int, [_vp, _dp, _dp, C.c_int32]),
    "bs_set_heights": (C.c_int, [_vp, _dp, C.c_int32]),
    "bs_move_grid": (C.c_int, [_vp, _dp, C.c_int32]),
    "bs_iterate_fitted": (C.c_int, [_vp, C.c_int32, C.c_int32, C.POINTER(Report)]),
    "bs_read_heights": (C.c_int, [_vp, _dp, C.c_int32]),
    "bs_anderson_create": (C.c_int, [_vp, C.c_int64, C.c_int32, C.POINTER(_vp)]),
    "bs_anderson_destroy": (C.c_int, [_vp]),
    "bs_anderson_reset": (C.c_int, [_vp]),
    "bs_anderson_next": (C.c_int, [_vp, _dp, _dp, _dp]),
    "bs_anderson_state": (C.c_int, [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32), _dp]),
    "bs_iterate_anderson": (C.c_int, [_vp, _vp, C.c_int32, C.c_double, C.c_int32, C.c_double, C.c_double,
                                      C.POINTER(Report)]),
    "bs_enable_timing": (C.c_int, [_vp, C.c_int32]),
    "bs_kernel_timing": (C.c_int, [_vp, C.c_int32, _dp, C.POINTER(C.c_int64)]),
    "bs_batch_create": (C.c_int, [_vp, C.POINTER(Config), C.c_int32, C.POINTER(_vp)]),
    "bs_batch_create_fp64": (C.c_int, [_vp, C.POINTER(Config), C.c_int32, C.POINTER(_vp)]),
    "bs_batch_precision": (C.c_int, [_vp, C.POINTER(C.c_int32)]),
    "bs_batch_destroy": (C.c_int, [_vp]),
    "bs_batch_set_uniform": (C.c_int, [_vp, _dp]),
    "bs_batch_set_primitives": (C.c_int, [_vp, C.c_int32, _dp]),
    "bs_batch_set_coarse": (C.c_int, [_vp, _dp, C.c_int32, C.c_int32, _dp, C.c_double]),
    "bs_batch_set_coarse_noalloc": (C.c_int, [_vp, _dp, C.c_int32, C.c_int32, _dp, C.c_double]),
    "bs_batch_read_state": (C.c_int, [_vp, C.c_int32, _fp, _fp, _fp, _fp]),
    "bs_batch_read_state_f64": (C.c_int, [_vp, C.c_int32, _dp, _dp, _dp, _dp]),
    "bs_batch_set_from_coarser": (C.c_int, [_vp, _vp]),
    "bs_batch_set_active": (C.c_int, [_vp, C.POINTER(C.c_uint8), C.c_int32]),
    "bs_batch_set_tiling": (C.c_int, [_vp, C.c_int32]),
    "bs_batch_set_persistent": (C.c_int, [_vp, C.c_int32]),
    "bs_batch_set_relax_fusion": (C.c_int, [_vp, C.c_int32]),
    "bs_batch_persistent_info": (C.c_int, [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "bs_batch_set_relaxation": (C.c_int, [_vp, C.c_int32]),
    "bs_batch_iterate": (C.c_int, [_vp, C.c_int32]),
    "bs_batch_set_implicit_cfl": (C.c_int, [_vp, C.c_double]),
    "bs_batch_set_implicit_cfls": (C.c_int, [_vp, _dp]),
    "bs_batch_iterate_implicit": (C.c_int, [_vp, C.c_int32, C.c_double, C.c_int32, C.c_double, C.c_double]),
    "bs_batch_residual_norms": (C.c_int, [_vp, _dp]),
    "bs_batch_drags": (C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "bs_batch_status": (C.c_int, [_vp, C.POINTER(C.c_uint32), C.POINTER(C.c_int64)]),
    "bs_batch_enable_timing": (C.c_int, [_vp, C.c_int32]),
    "bs_batch_kernel_timing": (C.c_int, [_vp, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "bs_mlp_create": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(_vp)]),
    "bs_mlp_destroy": (C.c_int, [_vp]),
    "bs_mlp_param_count": (C.c_int64, [_vp]),
    "bs_mlp_set_params": (C.c_int, [_vp, _fp, C.c_int64]),
    "bs_mlp_get_params": (C.c_int, [_vp, _fp, C.c_int64]),
    "bs_mlp_load_data": (C.c_int, [_vp, _fp, _fp, C.c_int32]),
    "bs_mlp_gradients": (C.c_int, [_vp, C.POINTER(C.c_int32), C.c_int32, _fp, _dp]),
    "bs_mlp_train_batch": (C.c_int, [_vp, C.POINTER(C.c_int32), C.c_int32, C.c_double, _dp]),
    "bs_mlp_predict": (C.c_int, [_vp, _fp, C.c_int32, _fp]),
    "bs_mlp_time_gemm": (C.c_int, [_vp, C.c_int32, C.c_int32, _dp, _dp]),
}

_lib = None


class LibraryMissing(ImportError):
    pass


def lib():
    """The loaded library; raises LibraryMissing if it was never built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(f"{LIB_PATH} is missing: build it with `python -m bowshock_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return lib().bs_last_error().decode()


def dptr(a):
    return a.ctypes.data_as(_dp)


def fptr(a):
    return a.ctypes.data_as(_fp)
