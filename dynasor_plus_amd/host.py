"""ctypes binding of the host helper library (include/dynasor_host.h):
reference-stream factor initialisation and the synthetic tensor generator."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from .tensor_io import Tensor

HOST_LIB = Path(__file__).resolve().parent / "lib" / "libdynasor_host.so"
_lib = None


class GenSpec(C.Structure):
    _fields_ = [
        ("num_modes", C.c_int),
        ("dims", C.c_uint32 * 8),
        ("nnz", C.c_uint64),
        ("seed", C.c_uint64),
        ("coord_dist", C.c_int),
        ("alpha", C.c_double),
        ("value_kind", C.c_int),
        ("mu", C.c_double),
        ("sigma", C.c_double),
        ("planted_rank", C.c_int),
        ("planted_factor_dist", C.c_int),
        ("planted_var", C.c_double),
        ("noise_frac", C.c_double),
    ]


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not HOST_LIB.exists():
        raise FileNotFoundError(f"{HOST_LIB} is missing: run `python -m dynasor_plus_amd.build`")
    from . import _lib as gpu
    gpu.load_library()          # libdynasor_host links against libdynasor_gpu
    lib = C.CDLL(str(HOST_LIB), mode=C.RTLD_GLOBAL)
    fp = C.POINTER(C.POINTER(C.c_float))
    for name in ("dyn_host_init_normal", "dyn_host_init_uniform"):
        fn = getattr(lib, name)
        fn.restype = None
        fn.argtypes = [C.c_int, C.POINTER(C.c_uint32), C.c_int, C.c_int, C.c_uint64, fp]
    lib.dyn_gen_tensor.restype = C.c_int
    lib.dyn_gen_tensor.argtypes = [C.POINTER(GenSpec), C.POINTER(C.POINTER(C.c_uint32)),
                                   C.POINTER(C.c_float), C.POINTER(C.c_double),
                                   C.POINTER(C.c_double), C.c_int]
    _lib = lib
    return lib


def init_factors(dims, rank: int, seed: int, kind: str = "normal") -> list:
    lib = load()
    dims_a = np.array(dims, dtype=np.uint32)
    Y = [np.zeros((int(d), rank), dtype=np.float32) for d in dims]
    arr = (C.POINTER(C.c_float) * len(Y))(*[y.ctypes.data_as(C.POINTER(C.c_float)) for y in Y])
    fn = lib.dyn_host_init_normal if kind == "normal" else lib.dyn_host_init_uniform
    fn(len(dims), dims_a.ctypes.data_as(C.POINTER(C.c_uint32)), rank, rank, seed, arr)
    return Y


@dataclass
class GenStats:
    rounds: int
    draws: int
    max_row_nnz_mode0: int
    planted_rms: float


def generate(dims, nnz: int, seed: int, coords: str = "uniform", alpha: float = 1.0,
             values: str = "lognormal", mu: float = 0.0, sigma: float = 1.0,
             planted_rank: int = 10, planted_factors: str = "uniform", planted_var: float = 1.0,
             noise_frac: float = 0.0, threads: int = 0, want_f64: bool = False):
    """Distinct-coordinate synthetic tensor.  Returns (Tensor, vals64 or None, GenStats)."""
    lib = load()
    s = GenSpec()
    s.num_modes = len(dims)
    for n, d in enumerate(dims):
        s.dims[n] = int(d)
    s.nnz = int(nnz)
    s.seed = int(seed)
    s.coord_dist = {"uniform": 0, "zipf": 1}[coords]
    s.alpha = float(alpha)
    s.value_kind = {"lognormal": 0, "planted": 1, "bernoulli": 2, "uniform": 3}[values]
    s.mu, s.sigma = float(mu), float(sigma)
    s.planted_rank = int(planted_rank)
    s.planted_factor_dist = {"uniform": 0, "normal": 1}[planted_factors]
    s.planted_var = float(planted_var)
    s.noise_frac = float(noise_frac)
    idx = [np.empty(int(nnz), dtype=np.uint32) for _ in dims]
    vals = np.empty(int(nnz), dtype=np.float32)
    v64 = np.empty(int(nnz), dtype=np.float64) if want_f64 else None
    stats = np.zeros(4, dtype=np.float64)
    arr = (C.POINTER(C.c_uint32) * len(idx))(*[i.ctypes.data_as(C.POINTER(C.c_uint32)) for i in idx])
    rc = lib.dyn_gen_tensor(C.byref(s), arr, vals.ctypes.data_as(C.POINTER(C.c_float)),
                            v64.ctypes.data_as(C.POINTER(C.c_double)) if want_f64 else None,
                            stats.ctypes.data_as(C.POINTER(C.c_double)), int(threads))
    if rc != 0:
        raise ValueError(f"dyn_gen_tensor failed (rc={rc}) for dims={dims} nnz={nnz}")
    st = GenStats(int(stats[0]), int(stats[1]), int(stats[2]), float(stats[3]))
    return Tensor(tuple(int(d) for d in dims), vals, idx), v64, st
