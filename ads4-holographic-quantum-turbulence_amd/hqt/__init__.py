"""hqt: MI355X-native AdS4 holographic-superfluid RK4 stepper (gfx950 HIP kernels).

The compute path is ``csrc/libhqt.so`` (C ABI in ``include/hqt.h``); this package is its
Python mirror of the reference's MATLAB interface (``fourdif``, ``chebdif``, ``ff``,
``NewtonRaphson_vy``, the ``F1..F4`` closures and the RK4 loop body as :class:`Stepper`).
"""

from ._lib import HQT_ERR_ARG, HQT_ERR_HIP, HQT_ERR_NAN, HQT_ERR_NOCONV, HQT_ERR_OOM, HQT_ERR_TIMEOUT, HQT_ERR_STATE, HQTError, LIB_PATH, lib
from .api import Ensemble, NewtonRaphson_vy, NewtonRaphson_vy_batch, Server, Stepper, checkpoint_name, footprint, spline, write_checkpoint, chebdif, ff, fourdif, mask_factors

__all__ = [
    "Stepper", "Ensemble", "Server", "footprint", "checkpoint_name", "write_checkpoint", "spline", "fourdif", "chebdif", "ff", "NewtonRaphson_vy", "NewtonRaphson_vy_batch", "mask_factors", "HQTError", "lib", "LIB_PATH",
    "HQT_ERR_ARG", "HQT_ERR_NAN", "HQT_ERR_HIP", "HQT_ERR_OOM", "HQT_ERR_STATE", "HQT_ERR_NOCONV",
]
