"""MI355X-native sparse CP-ALS engine (drop-in for the `--decompose cpals` path).

Python surface (the C++ drop-in lives in csrc/host, the C-ABI in include/):
    GpuContext            one device context (tensor, per-mode copies, factors)
    cpals, CpOptions, CpResult
    Tensor, load_coo_tensor, load_tns, load_dnb, save_dnb, load_dyfm, save_dyfm
"""
from .tensor_io import (Tensor, TensorFormatError, load_coo_tensor, load_dnb, load_dyfm,
                        load_tns, save_dnb, save_dyfm, save_tns)
from .engine import CpOptions, CpResult, GpuContext, cpals, init_factors, padded_rank
from ._lib import DynGpuError, load_library

__all__ = [
    "Tensor", "TensorFormatError", "load_coo_tensor", "load_dnb", "load_dyfm", "load_tns",
    "save_dnb", "save_dyfm", "save_tns", "CpOptions", "CpResult", "GpuContext", "cpals",
    "init_factors", "padded_rank", "DynGpuError", "load_library",
]
