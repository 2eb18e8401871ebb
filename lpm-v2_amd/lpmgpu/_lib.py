"""ctypes binding of liblpm_direct.so (the C ABI in include/lpm_direct.h).

The library is the product: every call here goes to the gfx950 HIP code.
There is no CPU fallback; a missing library raises LpmError.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_ROOT, "liblpm_direct.so")
HEADER_PATH = os.path.join(os.path.dirname(PKG_ROOT), "include", "lpm_direct.h")

LPM_OK = 0
LPM_ERR_ARG = 1
LPM_ERR_HIP = 2
LPM_ERR_NOMEM = 3
LPM_ERR_STATE = 4
LPM_ERR_UNSUPPORTED = 5

PREC_FP64 = 0
PREC_FP32_KAHAN = 1
PRECISIONS = {"fp64": PREC_FP64, "fp32": PREC_FP32_KAHAN, "fp32k": PREC_FP32_KAHAN}

(K_SPHERE_BS, K_PLANE_BS, K_PSE8, K_PSE2, K_SPHERE_GREEN, K_PLANE_GREEN, K_FAR_FIELD, K_PSE_SPHERE8,
 K_PSE_DERIV_PLANE, K_PSE_DERIV_SPHERE, K_PSE_INTERP_PLANE, K_PSE_INTERP_SPHERE, K_SWE_PLANE) = range(13)
K_COUNT = 13
KIND_NAMES = ["sphere_bs", "plane_bs", "pse8", "pse2", "sphere_green", "plane_green", "far_field", "pse_sphere8",
              "pse_deriv_plane", "pse_deriv_sphere", "pse_interp_plane", "pse_interp_sphere", "swe_plane"]

SOLVER_STREAM_EACH_STEP = 1
SOLVER_USE_GRAPH = 2
SOLVER_PROJECT_STAGES = 4


class LpmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[lpm status {code}] {msg}")
        self.code = code


_lib = None

_d = ctypes.c_double
_i = ctypes.c_int
_pd = ctypes.POINTER(ctypes.c_double)
_pi = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p

_SIGS = {
    "lpm_version": ([], _i),
    "lpm_create": ([ctypes.POINTER(_vp), _i], _i),
    "lpm_destroy": ([_vp], _i),
    "lpm_last_error": ([_vp], ctypes.c_char_p),
    "lpm_synchronize": ([_vp], _i),
    "lpm_bve_sphere_velocity": ([_vp, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _d, _d, _pi, _i, _i, _i, _i], _i),
    # device-memory arrays: passed as addresses
    "lpm_bve_sphere_velocity_noalloc": ([_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _d, _vp, _i, _i, _i, _i],
                                        _i),
    "lpm_plane_velocity": ([_vp, _pd, _pd, _pd, _pd, _pd, _pd, _pi, _i, _i, _i, _i], _i),
    "lpm_pse_plane_laplacian": ([_vp, _pd, _pd, _pd, _pd, _pd, _pi, _i, _i, _i, _d, _i, _i], _i),
    "lpm_pse_sphere_laplacian": ([_vp, _pd, _pd, _pd, _pd, _pd, _pd, _pi, _i, _i, _i, _d, _d, _i], _i),
    "lpm_pse_sphere_laplacian_at_points": ([_vp, _pd, _pd, _pd, _pd, _pd, _i, _pd, _pd, _pd, _pd, _pd, _pi, _i, _d,
                                            _d, _i], _i),
    "lpm_pse_plane_gradient": ([_vp, _pd, _pd, _pd, _pd, _pd, _pd, _pi, _i, _i, _i, _d, _i], _i),
    "lpm_pse_sphere_gradient": ([_vp, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pi, _i, _i, _i, _d, _d, _i], _i),
    "lpm_pse_plane_second_partials": ([_vp, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pi, _i, _i, _i, _d, _i], _i),
    "lpm_pse_plane_double_dot": ([_vp, _pd, _pd, _pd, _pd, _pd, _pd, _pi, _i, _i, _i, _d, _i], _i),
    "lpm_pse_sphere_double_dot": ([_vp, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pi, _i, _i, _i, _d, _d, _i], _i),
    "lpm_pse_sphere_divergence": ([_vp, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pi, _i, _i, _i, _d, _d, _i], _i),
    "lpm_pse_plane_interpolate": ([_vp, _pd, _pd, _pd, _i, _pd, _pd, _pd, _pd, _pi, _i, _d, _i], _i),
    "lpm_pse_sphere_interpolate": ([_vp, _pd, _pd, _pd, _pd, _i, _pd, _pd, _pd, _pd, _pd, _pi, _i, _d, _d, _i], _i),
    "lpm_swe_plane_rhs": ([_vp, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pi, _i, _i, _i, _d, _i],
                          _i),
    "lpm_bve_sphere_stream": ([_vp, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _d, _pi, _i, _i, _i, _i], _i),
    "lpm_plane_stream": ([_vp, _pd, _pd, _pd, _pd, _pd, _pi, _i, _i, _i, _i], _i),
    "lpm_pse_eps": ([_d, _d, _pd], _i),
    "lpm_load_balance": ([_i, _i, _pi, _pi, _pi], _i),
    "lpm_diagnostics": ([_vp, _pd, _pd, _pd, _pd, _pd, _pi, _i, _pd, _pd, _pd], _i),
    "lpm_bve_solver_new": (
        [_vp, ctypes.POINTER(_vp), _i, _pd, _pd, _pd, _pd, _pd, _pd, _pi, _pd, _pd, _pd, _d, _d, _i, _i],
        _i,
    ),
    "lpm_bve_solver_timestep": ([_vp, _d], _i),
    "lpm_bve_solver_steps": ([_vp, _d, _i], _i),
    "lpm_bve_solver_get_state": ([_vp, _pd, _pd, _pd, _pd, _pd, _pd, _pd], _i),
    "lpm_bve_solver_stream": ([_vp, _pd, _pd], _i),
    "lpm_bve_solver_diagnostics": ([_vp, _pd, _pd, _pd], _i),
    "lpm_bve_solver_delete": ([_vp], _i),
    "lpm_bve_solver_set_particles": ([_vp, _i, _pd, _pd, _pd, _pd, _pd, _pd, _pi, _pd, _pd, _pd], _i),
    "lpm_plane_solver_new": ([_vp, ctypes.POINTER(_vp), _i, _pd, _pd, _pd, _pd, _pi, _pd, _pd, _i, _i], _i),
    "lpm_plane_solver_set_particles": ([_vp, _i, _pd, _pd, _pd, _pd, _pi, _pd, _pd], _i),
    "lpm_plane_solver_timestep": ([_vp, _d], _i),
    "lpm_plane_solver_steps": ([_vp, _d, _i], _i),
    "lpm_plane_solver_get_state": ([_vp, _pd, _pd, _pd, _pd], _i),
    "lpm_plane_solver_stream": ([_vp, _pd], _i),
    "lpm_plane_solver_diagnostics": ([_vp, _pd, _pd, _pd], _i),
    "lpm_plane_solver_delete": ([_vp], _i),
    "lpm_bve_batch_new": (
        [_vp, ctypes.POINTER(_vp), _i, ctypes.POINTER(ctypes.c_int64), _pd, _pd, _pd, _pd, _pd, _pi, _pd, _pd,
         _i, _i],
        _i,
    ),
    "lpm_bve_batch_steps": ([_vp, _d, _i], _i),
    "lpm_bve_batch_get_state": ([_vp, _pd, _pd, _pd, _pd, _pd, _pd, _pd], _i),
    "lpm_bve_batch_set_tuning": ([_vp, _i, _i, _i], _i),
    "lpm_bve_batch_get_tuning": ([_vp, _pi, _pi, _pi, _pi], _i),
    "lpm_bve_batch_autotune": ([_vp, _pi, _pi, _pi, _pd], _i),
    "lpm_bve_batch_pairs": ([_vp, ctypes.POINTER(ctypes.c_int64)], _i),
    "lpm_bve_batch_delete": ([_vp], _i),
    "lpm_mlp_new": ([_vp, ctypes.POINTER(_vp), _i, _i, _i, _i, _i, ctypes.c_uint64], _i),
    "lpm_mlp_param_count": ([_vp, ctypes.POINTER(ctypes.c_int64)], _i),
    "lpm_mlp_set_params": ([_vp, ctypes.POINTER(ctypes.c_float), ctypes.c_int64], _i),
    "lpm_mlp_get_params": ([_vp, ctypes.POINTER(ctypes.c_float), ctypes.c_int64], _i),
    "lpm_mlp_forward": ([_vp, _i, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), _i], _i),
    "lpm_mlp_loss_grad": ([_vp, _i, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                           ctypes.POINTER(ctypes.c_float), _pd], _i),
    "lpm_mlp_train": ([_vp, _i, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), _i, _i, _d,
                       ctypes.c_uint64, _pd], _i),
    "lpm_mlp_delete": ([_vp], _i),
    "lpm_set_timing": ([_vp, _i], _i),
    "lpm_kernel_stats": ([_vp, _i, ctypes.POINTER(ctypes.c_int64), _pd], _i),
    "lpm_reset_kernel_stats": ([_vp], _i),
    "lpm_set_tuning": ([_vp, _i, _i, _i, _i, _i], _i),
    "lpm_get_tuning": ([_vp, _i, _i, _pi, _pi, _pi], _i),
}


def extra_signatures(sigs: dict):
    """Register signatures of further entry points (used by other modules)."""
    _SIGS.update(sigs)
    if _lib is not None:
        _bind(_lib, sigs)


def _bind(lib, sigs):
    for name, (args, res) in sigs.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.argtypes = args
        fn.restype = res


def lib():
    """Load the HIP library (raises LpmError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LpmError(LPM_ERR_STATE, f"HIP library not built: {LIB_PATH} missing "
                           "(run `python lpm-v2_amd/build.py`)")
        _lib = ctypes.CDLL(LIB_PATH)
        _bind(_lib, _SIGS)
    return _lib


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def pd(a):
    return None if a is None else a.ctypes.data_as(_pd)


def pi(a):
    return None if a is None else a.ctypes.data_as(_pi)


def prec_code(p) -> int:
    if isinstance(p, str):
        return PRECISIONS[p.lower()]
    return int(p)
