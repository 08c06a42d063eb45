"""ctypes binding of ``libhqt.so`` (the C ABI declared in ``include/hqt.h``).

The library is built in-tree (``csrc/libhqt.so``).  There is no fallback: if the shared
library is missing or a HIP call fails, the call raises :class:`HQTError`.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HQT_OK, HQT_ERR_ARG, HQT_ERR_NAN, HQT_ERR_HIP, HQT_ERR_OOM, HQT_ERR_STATE, HQT_ERR_NOCONV, HQT_ERR_TIMEOUT = range(8)
HQT_FP64, HQT_MIXED = 0, 1
HQT_NO_GRAPH = 1
HQT_LINES16 = 2
HQT_LINES32 = 4
HQT_RADIAL_W2 = 8
HQT_RADIAL_W3 = 16
HQT_DENSE_FILTER = 32
HQT_DENSE_LINES = 64
HQT_ORDER_XCD = 128
HQT_FOLDED_STAGES = 256
HQT_FACTORED_STAGES = 512
HQT_FOLDED_FILTER = 1024
HQT_FP32_FILTER = 2048
HQT_SERVER_NCOUNTERS = 10

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "csrc", "libhqt.so")

EXPORTS = (
    "hqt_ctx_create", "hqt_ctx_destroy", "hqt_set_state", "hqt_get_state", "hqt_step_n", "hqt_sync",
    "hqt_last_winding", "hqt_rhs", "hqt_order_parameter", "hqt_ff", "hqt_ff_mask", "hqt_fourdif", "hqt_chebdif",
    "hqt_newton_vy", "hqt_newton_vy_batch", "hqt_newton_vy_batch_noalloc", "hqt_newton_vy_continue_batch", "hqt_vortex_census", "hqt_ensemble_create", "hqt_ensemble_step_n", "hqt_ensemble_sync",
    "hqt_ensemble_launches", "hqt_ensemble_destroy", "hqt_server_create", "hqt_server_destroy",
    "hqt_server_submit_evolution", "hqt_server_submit_newton", "hqt_server_submit_newton_continued", "hqt_server_wait", "hqt_server_evolution_result",
    "hqt_server_newton_result", "hqt_server_stats", "hqt_server_counters", "hqt_ctx_footprint", "hqt_ctx_device_bytes",
    "hqt_checkpoint_name", "hqt_checkpoint_write_host", "hqt_checkpoint_save", "hqt_checkpoint_save_async",
    "hqt_checkpoint_wait", "hqt_init_state", "hqt_spline", "hqt_debug_field", "hqt_profile", "hqt_kernel_stats", "hqt_calibration_copy", "hqt_calibration_radial",
    "hqt_last_error", "hqt_version",
)


class HQTError(RuntimeError):
    """A non-zero status from the C ABI; ``code`` is the HQT_ERR_* value."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"hqt error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    """Mirror of ``hqt_params`` (include/hqt.h)."""
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("Lx", C.c_double), ("Ly", C.c_double), ("mu", C.c_double), ("dt", C.c_double),
        ("winding", C.c_int32), ("filter", C.c_int32), ("precision", C.c_int32), ("flags", C.c_int32),
        ("ax_b", C.POINTER(C.c_double)), ("ay_b", C.POINTER(C.c_double)),
        ("ax_b_const", C.c_double), ("ay_b_const", C.c_double),
    ]


class Init(C.Structure):
    """Mirror of ``hqt_init`` (include/hqt.h)."""
    _fields_ = [
        ("nz0", C.c_int32),
        ("Z1", C.POINTER(C.c_double)), ("Z2", C.POINTER(C.c_double)), ("Vy1", C.POINTER(C.c_double)),
        ("rhoi", C.POINTER(C.c_double)),
        ("amp_noise", C.c_double), ("phase_noise", C.c_double), ("seed", C.c_uint64),
        ("nv", C.c_int32),
        ("vx", C.POINTER(C.c_double)), ("vy", C.POINTER(C.c_double)), ("q", C.POINTER(C.c_double)),
        ("sigma", C.c_double),
    ]


_lib = None


def lib():
    """Load ``libhqt.so`` once; raise if it is absent (no CPU fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HQTError(HQT_ERR_STATE, f"{LIB_PATH} not built; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    i32 = C.c_int32
    L.hqt_last_error.restype = C.c_char_p
    L.hqt_version.restype = C.c_char_p
    L.hqt_ctx_create.argtypes = [C.POINTER(Params), C.POINTER(vp)]
    L.hqt_ctx_destroy.argtypes = [vp]
    L.hqt_set_state.argtypes = [vp, dp, dp, dp, dp, dp]
    L.hqt_get_state.argtypes = [vp, dp, dp, dp, dp, dp]
    L.hqt_step_n.argtypes = [vp, i32]
    L.hqt_sync.argtypes = [vp, C.POINTER(i32)]
    L.hqt_last_winding.argtypes = [vp, dp]
    L.hqt_rhs.argtypes = [vp, dp, dp, dp, dp, dp, dp, dp, dp]
    L.hqt_order_parameter.argtypes = [vp, dp]
    L.hqt_ensemble_create.argtypes = [C.POINTER(vp), i32, i32, C.POINTER(vp)]
    L.hqt_ensemble_step_n.argtypes = [vp, i32]
    L.hqt_ensemble_sync.argtypes = [vp, C.POINTER(i32)]
    L.hqt_ensemble_launches.argtypes = [vp, C.POINTER(i32)]
    L.hqt_ensemble_destroy.argtypes = [vp]
    L.hqt_server_create.argtypes = [C.c_int64, i32, C.POINTER(vp)]
    L.hqt_server_destroy.argtypes = [vp]
    L.hqt_server_submit_evolution.argtypes = [vp, C.POINTER(Params), dp, dp, dp, dp, dp, i32, C.POINTER(C.c_int64)]
    L.hqt_server_submit_newton.argtypes = [vp, i32, C.c_double, C.c_double, C.c_double, dp, dp, dp, C.c_double,
                                           C.POINTER(C.c_int64)]
    L.hqt_server_submit_newton_continued.argtypes = [vp, i32, C.c_double, C.c_double, C.c_double, dp, dp, dp,
                                                     C.c_double, C.c_double, C.c_double, C.POINTER(C.c_int64)]
    L.hqt_server_wait.argtypes = [vp, C.c_int64, C.c_double, dp]
    L.hqt_server_evolution_result.argtypes = [vp, C.c_int64, dp, dp, dp, dp, dp, C.POINTER(i32)]
    L.hqt_server_newton_result.argtypes = [vp, C.c_int64, dp, dp, dp, C.POINTER(i32), dp]
    L.hqt_server_stats.argtypes = [vp, i32, dp, dp, C.POINTER(C.c_int64)]
    L.hqt_server_counters.argtypes = [vp, C.POINTER(C.c_int64)]
    L.hqt_ctx_footprint.argtypes = [C.POINTER(Params), C.POINTER(C.c_int64)]
    L.hqt_ctx_device_bytes.argtypes = [vp, C.POINTER(C.c_int64)]
    L.hqt_checkpoint_name.argtypes = [C.c_double, C.c_char_p, i32]
    L.hqt_checkpoint_write_host.argtypes = [C.c_char_p, i32, i32, i32, dp, dp, dp, dp, dp]
    L.hqt_checkpoint_save.argtypes = [vp, C.c_char_p]
    L.hqt_checkpoint_save_async.argtypes = [vp, C.c_char_p]
    L.hqt_checkpoint_wait.argtypes = [vp]
    L.hqt_init_state.argtypes = [vp, C.POINTER(Init)]
    L.hqt_spline.argtypes = [i32, dp, dp, i32, dp, dp]
    L.hqt_vortex_census.argtypes = [vp, C.POINTER(i32), C.POINTER(C.c_int8), dp]
    L.hqt_ff.argtypes = [i32, i32, i32, dp, dp, C.POINTER(C.c_uint8), C.POINTER(C.c_uint8), dp]
    L.hqt_ff_mask.argtypes = [i32, i32, i32, dp, dp, C.POINTER(C.c_uint8), i32, dp]
    L.hqt_fourdif.argtypes = [i32, i32, dp, dp]
    L.hqt_chebdif.argtypes = [i32, i32, dp, dp]
    L.hqt_newton_vy.argtypes = [i32, C.c_double, C.c_double, C.c_double, dp, dp, dp, C.c_double, dp, dp, dp,
                                C.POINTER(i32), dp]
    L.hqt_newton_vy_batch.argtypes = [i32, C.c_double, C.c_double, C.c_double, i32, dp, dp, dp, dp, i32,
                                      C.c_double, dp, dp, dp, C.POINTER(i32), dp, C.POINTER(C.c_uint8)]
    L.hqt_newton_vy_batch_noalloc.argtypes = L.hqt_newton_vy_batch.argtypes
    L.hqt_newton_vy_continue_batch.argtypes = [i32, C.c_double, C.c_double, C.c_double, i32, dp, dp, dp, dp, dp,
                                               C.c_double, i32, C.c_double, dp, dp, dp, C.POINTER(i32), dp,
                                               C.POINTER(C.c_uint8), C.POINTER(i32)]
    L.hqt_debug_field.argtypes = [vp, C.c_char_p, dp]
    L.hqt_profile.argtypes = [vp, i32]
    L.hqt_calibration_copy.argtypes = [C.c_int64, dp]
    L.hqt_calibration_radial.argtypes = [C.c_int64, C.c_int32, C.c_int32, C.c_int32, dp, dp]
    L.hqt_kernel_stats.argtypes = [vp, i32, C.c_char_p, i32, C.POINTER(C.c_int64), dp]
    for name in EXPORTS:
        if name not in ("hqt_last_error", "hqt_version"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check(rc: int):
    if rc != HQT_OK:
        raise HQTError(rc, lib().hqt_last_error().decode(errors="replace"))


def dptr(a: np.ndarray | None):
    """``double*`` to the data of a contiguous float64 or complex128 array (None -> NULL).
    A complex128 array in Fortran order is the interleaved (re, im) column-major layout
    the C ABI expects."""
    if a is None:
        return None
    if a.dtype not in (np.float64, np.complex128) or not (a.flags.f_contiguous or a.flags.c_contiguous):
        raise ValueError("dptr: need a contiguous float64/complex128 array")
    return a.ctypes.data_as(C.POINTER(C.c_double))
