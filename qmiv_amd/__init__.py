"""MI355X-native (gfx950) implementation of QMIV's metric hot path.

The product is ``libivqm_hip.so`` (HIP kernels + C ABI, include/ivqm_hip.h).
This package holds its build script, the ctypes bindings used by the tests,
the bench and the smoke check, and the seeded synthetic-frame generator.
"""
from ._lib import (IVQM_M_ALL, IVQM_M_IVMSSSIM, IVQM_M_IVPSNR, IVQM_M_IVSSIM, IVQM_M_MSE, IVQM_M_MSSSIM,
                   IVQM_M_PSNR, IVQM_M_SSIM, IVQM_M_WSPSNR, IvqmError)
from .context import IvqmBatch, IvqmContext, IvqmServer, default_config

__all__ = [
    "IvqmBatch", "IvqmContext", "IvqmServer", "default_config", "IvqmError", "IVQM_M_ALL", "IVQM_M_IVMSSSIM", "IVQM_M_IVPSNR", "IVQM_M_IVSSIM",
    "IVQM_M_MSE", "IVQM_M_MSSSIM", "IVQM_M_PSNR", "IVQM_M_SSIM", "IVQM_M_WSPSNR",
]
