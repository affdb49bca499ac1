"""ctypes access to the CPU oracle (oracle/liboracle.so) for the tests."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "oracle" / "liboracle.so"

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)


class CpOpts(C.Structure):
    _fields_ = [("max_iters", C.c_int), ("tol", C.c_double), ("ridge", C.c_double),
                ("jacobi", C.c_int), ("fp64", C.c_int), ("stale_jacobi_fit", C.c_int),
                ("bf16_gather", C.c_int)]


class Oracle:
    def __init__(self, lib):
        self.lib = lib

    @staticmethod
    def _cols(T):
        cols = [np.ascontiguousarray(c, dtype=np.uint32) for c in T.idx]
        return cols, (u32p * len(cols))(*[c.ctypes.data_as(u32p) for c in cols])

    def mttkrp(self, T, Y, mode):
        R = Y[0].shape[1]
        mats = [np.ascontiguousarray(y, dtype=np.float32) for y in Y]
        out = np.zeros((T.dims[mode], R), dtype=np.float32)
        cols, carr = self._cols(T)
        yarr = (f32p * len(mats))(*[m.ctypes.data_as(f32p) for m in mats])
        dims = np.array(T.dims, dtype=np.uint32)
        rc = self.lib.orc_mttkrp(T.num_modes, T.nnz, dims.ctypes.data_as(u32p),
                                 T.vals.ctypes.data_as(f32p), carr, R, R, yarr, mode,
                                 out.ctypes.data_as(f32p))
        assert rc == 0
        return out

    def mttkrp_f64(self, T, Y, mode):
        R = Y[0].shape[1]
        mats = [np.ascontiguousarray(y, dtype=np.float64) for y in Y]
        out = np.zeros((T.dims[mode], R), dtype=np.float64)
        cols, carr = self._cols(T)
        yarr = (f64p * len(mats))(*[m.ctypes.data_as(f64p) for m in mats])
        dims = np.array(T.dims, dtype=np.uint32)
        rc = self.lib.orc_mttkrp_f64(T.num_modes, T.nnz, dims.ctypes.data_as(u32p),
                                     T.vals.ctypes.data_as(f32p), carr, R, R, yarr, mode,
                                     out.ctypes.data_as(f64p))
        assert rc == 0
        return out

    def init(self, dims, R, seed, kind="normal"):
        Y = [np.zeros((d, R), dtype=np.float32) for d in dims]
        arr = (f32p * len(Y))(*[y.ctypes.data_as(f32p) for y in Y])
        d = np.array(dims, dtype=np.uint32)
        fn = self.lib.orc_init_normal if kind == "normal" else self.lib.orc_init_uniform
        fn(len(dims), d.ctypes.data_as(u32p), R, R, seed, arr)
        return Y

    def remap(self, T, mode, secondary=-1):
        cols, carr = self._cols(T)
        dims = np.array(T.dims, dtype=np.uint32)
        perm = np.zeros(max(T.nnz, 1), dtype=np.uint32)
        rowptr = np.zeros(T.dims[mode] + 1, dtype=np.uint64)
        rc = self.lib.orc_remap(T.num_modes, T.nnz, dims.ctypes.data_as(u32p), carr, mode,
                                secondary, perm.ctypes.data_as(u32p), rowptr.ctypes.data_as(u64p))
        assert rc == 0
        return perm[: T.nnz], rowptr

    def cpals(self, T, Y0, max_iters, tol, ridge=1e-10, jacobi=False, fp64=False,
              stale_jacobi_fit=False, bf16_gather=False):
        R = Y0[0].shape[1]
        Y = [np.array(y, dtype=np.float32, copy=True) for y in Y0]
        yarr = (f32p * len(Y))(*[y.ctypes.data_as(f32p) for y in Y])
        cols, carr = self._cols(T)
        dims = np.array(T.dims, dtype=np.uint32)
        o = CpOpts(max_iters, tol, ridge, int(jacobi), int(fp64), int(stale_jacobi_fit),
                   int(bf16_gather))
        hist = np.zeros(max(max_iters, 1), dtype=np.float64)
        lam = np.zeros(R, dtype=np.float64)
        fit = C.c_double(0)
        k = self.lib.orc_cpals(T.num_modes, T.nnz, dims.ctypes.data_as(u32p),
                               T.vals.ctypes.data_as(f32p), carr, R, yarr, C.byref(o),
                               hist.ctypes.data_as(f64p), lam.ctypes.data_as(f64p), C.byref(fit))
        return hist[:k], Y, lam

    def forward(self, T, A, lam=None):
        R = A[0].shape[1]
        mats = [np.ascontiguousarray(a, dtype=np.float64) for a in A]
        arr = (f64p * len(mats))(*[m.ctypes.data_as(f64p) for m in mats])
        cols, carr = self._cols(T)
        m = np.zeros(max(T.nnz, 1), dtype=np.float64)
        lp = None
        if lam is not None:
            lam = np.ascontiguousarray(lam, dtype=np.float64)
            lp = lam.ctypes.data_as(f64p)
        self.lib.orc_forward(T.num_modes, T.nnz, carr, R, R, arr, lp, m.ctypes.data_as(f64p))
        return m[: T.nnz]

    def completion_loss(self, T, A, loss):
        R = A[0].shape[1]
        mats = [np.ascontiguousarray(a, dtype=np.float64) for a in A]
        arr = (f64p * len(mats))(*[m.ctypes.data_as(f64p) for m in mats])
        cols, carr = self._cols(T)
        return self.lib.orc_completion_loss(T.num_modes, T.nnz, carr,
                                            T.vals.ctypes.data_as(f32p), R, R, arr, int(loss))


def load() -> Oracle:
    if not LIB.exists():
        raise FileNotFoundError(f"{LIB} missing: run `make -C oracle`")
    lib = C.CDLL(str(LIB))
    i, u64 = C.c_int, C.c_uint64
    pp32 = C.POINTER(u32p)
    lib.orc_mttkrp.restype = i
    lib.orc_mttkrp.argtypes = [i, u64, u32p, f32p, pp32, i, i, C.POINTER(f32p), i, f32p]
    lib.orc_mttkrp_f64.restype = i
    lib.orc_mttkrp_f64.argtypes = [i, u64, u32p, f32p, pp32, i, i, C.POINTER(f64p), i, f64p]
    for n in ("orc_init_normal", "orc_init_uniform"):
        getattr(lib, n).restype = None
        getattr(lib, n).argtypes = [i, u32p, i, i, u64, C.POINTER(f32p)]
    lib.orc_remap.restype = i
    lib.orc_remap.argtypes = [i, u64, u32p, pp32, i, i, u32p, u64p]
    lib.orc_cpals.restype = i
    lib.orc_cpals.argtypes = [i, u64, u32p, f32p, pp32, i, C.POINTER(f32p), C.POINTER(CpOpts),
                              f64p, f64p, f64p]
    lib.orc_forward.restype = None
    lib.orc_forward.argtypes = [i, u64, pp32, i, i, C.POINTER(f64p), f64p, f64p]
    lib.orc_completion_loss.restype = C.c_double
    lib.orc_completion_loss.argtypes = [i, u64, pp32, f32p, i, i, C.POINTER(f64p), i]
    return Oracle(lib)


def tensor_from_npz(z):
    from dynasor_plus_amd.tensor_io import Tensor
    dims = tuple(int(d) for d in z["dims"])
    return Tensor(dims, z["vals"], [z[f"idx{n}"] for n in range(len(dims))])
