"""MI355X-native Simultaneous Graphical DLM engine.

``SGDLM`` mirrors the reference's Rcpp module class (gpuSGDLM/SGDLM); the
compute runs in HIP kernels for gfx950 behind the C-ABI of include/sgdlm.h.
"""
from ._native import PREC_BF16X3, PREC_FP32, PREC_FP64, load  # noqa: F401
from .sgdlm import NA_INDEX, SGDLM, SGDLMError, batch_vb, encode_sp  # noqa: F401

__all__ = ["SGDLM", "SGDLMError", "batch_vb", "encode_sp", "NA_INDEX", "load",
           "PREC_FP64", "PREC_FP32", "PREC_BF16X3"]
