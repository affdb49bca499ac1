"""ctypes binding of the engine's C-ABI (include/dynasor_gpu.h).

The shared library is built in-tree by ``dynasor_plus_amd.build``; there is no
CPU fallback: importing a symbol from a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_DIR = PKG / "lib"
GPU_LIB = LIB_DIR / "libdynasor_gpu.so"

DYN_OK = 0
DYN_ERR_INVALID = 1
DYN_ERR_STATE = 2
DYN_ERR_OOM = 3
DYN_ERR_HIP = 4
DYN_ERR_UNSUPPORTED = 5

DYN_PREC_FP32 = 0
DYN_PREC_BF16 = 1
DYN_CP_GAUSS_SEIDEL = 0
DYN_CP_JACOBI = 1
DYN_LAYOUT_COPIES = 0
DYN_LAYOUT_LINEAR = 1
DYN_LOSS_LS = 0
DYN_LOSS_BERNOULLI = 1
MAX_MODES = 8
MAX_RANK = 128

_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_f32p = C.POINTER(C.c_float)
_f64p = C.POINTER(C.c_double)
_vp = C.c_void_p


class CpOpts(C.Structure):
    _fields_ = [
        ("max_iters", C.c_int),
        ("tol", C.c_double),
        ("ridge", C.c_double),
        ("kernel", C.c_int),
        ("precision", C.c_int),
        ("verbose", C.c_int),
        ("reserved", C.c_int * 5),
    ]


class CpResultC(C.Structure):
    _fields_ = [
        ("iters_done", C.c_int),
        ("final_fit", C.c_double),
        ("final_residual", C.c_double),
        ("tensor_norm", C.c_double),
        ("total_time_s", C.c_double),
        ("mttkrp_time_s", C.c_double),
        ("linalg_time_s", C.c_double),
        ("remap_time_s", C.c_double),
        ("fit_history", _f64p),
        ("iter_time_s", _f64p),
    ]


# name -> (restype, argtypes)
SIGNATURES = {
    "dyn_gpu_abi_version": (C.c_int, []),
    "dyn_gpu_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "dyn_gpu_destroy": (C.c_int, [_vp]),
    "dyn_gpu_last_error": (C.c_char_p, [_vp]),
    "dyn_gpu_set_stream": (C.c_int, [_vp, _vp]),
    "dyn_gpu_reset_stream": (C.c_int, [_vp]),
    "dyn_gpu_synchronize": (C.c_int, [_vp]),
    "dyn_gpu_mem_used": (C.c_int, [_vp, _u64p]),
    "dyn_gpu_device_memory": (C.c_int, [C.c_int, _u64p, _u64p]),
    "dyn_gpu_upload_coo": (C.c_int, [_vp, C.c_int, C.c_uint64, _u32p, _f32p, C.POINTER(_u32p)]),
    "dyn_gpu_remap": (C.c_int, [_vp, C.c_int, C.c_int, _u64p, _u32p]),
    "dyn_gpu_remap_all": (C.c_int, [_vp]),
    "dyn_gpu_get_copy": (C.c_int, [_vp, C.c_int, _f32p, C.POINTER(_u32p)]),
    "dyn_gpu_attach_host_coo": (C.c_int, [_vp, C.c_int, C.c_uint64, _u32p, _f32p,
                                          C.POINTER(_u32p), C.c_uint64]),
    "dyn_gpu_attach_dnb": (C.c_int, [_vp, C.c_char_p, C.c_uint64]),
    "dyn_gpu_save_processed": (C.c_int, [_vp, C.c_char_p]),
    "dyn_gpu_load_processed": (C.c_int, [_vp, C.c_char_p]),
    "dyn_gpu_processed_info": (C.c_int, [C.c_char_p, C.POINTER(C.c_int), _u64p, _u32p, _u32p]),
    "dyn_gpu_set_factors": (C.c_int, [_vp, C.c_int, C.POINTER(_f32p), C.c_int]),
    "dyn_gpu_get_factors": (C.c_int, [_vp, C.c_int, _f32p, C.c_int]),
    "dyn_gpu_get_lambda": (C.c_int, [_vp, _f64p]),
    "dyn_gpu_mttkrp": (C.c_int, [_vp, C.c_int, C.c_int, _f32p, C.c_int]),
    "dyn_gpu_mttkrp_noalloc": (C.c_int, [_vp, C.c_int, C.c_int, _vp, C.c_int]),
    "dyn_gpu_mttkrp_all": (C.c_int, [_vp, C.c_int, C.POINTER(_f32p), C.c_int]),
    "dyn_gpu_gram": (C.c_int, [_vp, C.c_int, _f64p]),
    "dyn_linear_plan": (C.c_int, [C.c_int, _u32p, _u64p, C.POINTER(C.c_uint8),
                                  C.POINTER(C.c_uint8)]),
    "dyn_gpu_build_linear": (C.c_int, [_vp, _u32p]),
    "dyn_gpu_set_layout": (C.c_int, [_vp, C.c_int]),
    "dyn_gpu_cpals": (C.c_int, [_vp, C.POINTER(CpOpts), C.POINTER(CpResultC)]),
    "dyn_gpu_forward": (C.c_int, [_vp, C.c_int, _f32p]),
    "dyn_gpu_adjoint": (C.c_int, [_vp, _f32p, C.c_int, _f32p, C.c_int]),
    "dyn_gpu_forward_dev": (C.c_int, [_vp, C.POINTER(_vp), C.c_int, _vp]),
    "dyn_gpu_adjoint_dev": (C.c_int, [_vp, _vp, C.c_int, C.POINTER(_vp), C.c_int, _vp]),
    "dyn_gpu_completion_grad_dev": (C.c_int, [_vp, C.c_int, C.POINTER(_vp), C.c_int,
                                              C.POINTER(_vp), _f64p]),
    "dyn_gpu_enable_timing": (C.c_int, [_vp, C.c_int]),
    "dyn_gpu_kernel_stats": (C.c_int, [_vp, _u64p, _f64p]),
}

_lib = None


class DynGpuError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[dyn_gpu error {code}] {msg}")
        self.code = code


def load_library(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load libdynasor_gpu.so and declare every C-ABI signature.

    Raises FileNotFoundError if the library has not been built: the engine has
    no non-GPU implementation to fall back to.
    """
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else GPU_LIB
    if not p.exists():
        raise FileNotFoundError(
            f"{p} is missing: build it with `python -m dynasor_plus_amd.build` "
            "(the engine has no CPU fallback)")
    lib = C.CDLL(str(p), mode=C.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib
