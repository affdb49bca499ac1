"""Host-side wrapper of the device engine (C-ABI in include/dynasor_gpu.h).

`GpuContext` owns one device context: the uploaded COO tensor, its per-mode
sorted copies and the factor matrices.  `cpals()` mirrors the reference entry
point `CpResult cpals(Tensor&, FactorMatrices&, const CpOptions&)`
(include/dynasor_cpals.h:82-121): same option names and meanings, same result
fields, the same N(0,1) initialisation stream per seed
(src/dynasor_cpals.cpp:42-69), and errors as exceptions carrying the status
code instead of an empty result.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .tensor_io import Tensor


def padded_rank(R: int) -> int:
    p = 16
    while p < R:
        p <<= 1
    return p


def _f32ptr(a: np.ndarray):
    return a.ctypes.data_as(L._f32p)


def _u32ptr(a: np.ndarray):
    return a.ctypes.data_as(L._u32p)


class GpuContext:
    """One device context (one GPU, one HIP stream)."""

    def __init__(self, device: int = 0):
        self.lib = L.load_library()
        h = L._vp()
        rc = self.lib.dyn_gpu_create(int(device), C.byref(h))
        if rc != L.DYN_OK:
            raise L.DynGpuError(rc, f"dyn_gpu_create(device={device}) failed (is a GPU visible?)")
        self.h = h
        self.N = 0
        self.dims: tuple = ()
        self.nnz = 0
        self.R = 0

    # -- plumbing --------------------------------------------------------
    def _check(self, rc: int, what: str) -> None:
        if rc != L.DYN_OK:
            msg = self.lib.dyn_gpu_last_error(self.h)
            raise L.DynGpuError(rc, f"{what}: {msg.decode() if msg else ''}")

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.dyn_gpu_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_ptr: int) -> None:
        """Run on hipStream_t `stream_ptr` (0 = the default stream, torch's default)."""
        self._check(self.lib.dyn_gpu_set_stream(self.h, L._vp(int(stream_ptr))), "set_stream")

    def reset_stream(self) -> None:
        self._check(self.lib.dyn_gpu_reset_stream(self.h), "reset_stream")

    def on_stream(self, stream_ptr: int):
        """Context manager: run on `stream_ptr`, then return to the own stream."""
        eng = self

        class _On:
            def __enter__(self):
                eng.set_stream(stream_ptr)
                return eng

            def __exit__(self, *exc):
                eng.reset_stream()

        return _On()

    def synchronize(self) -> None:
        self._check(self.lib.dyn_gpu_synchronize(self.h), "synchronize")

    def mem_used(self) -> int:
        v = C.c_uint64(0)
        self._check(self.lib.dyn_gpu_mem_used(self.h, C.byref(v)), "mem_used")
        return int(v.value)

    # -- tensor ------------------------------------------------------------
    def upload(self, T: Tensor) -> None:
        dims = np.array(T.dims, dtype=np.uint32)
        cols = [np.ascontiguousarray(c, dtype=np.uint32) for c in T.idx]
        arr = (L._u32p * len(cols))(*[_u32ptr(c) for c in cols])
        vals = np.ascontiguousarray(T.vals, dtype=np.float32)
        self._check(self.lib.dyn_gpu_upload_coo(self.h, T.num_modes, T.nnz, _u32ptr(dims),
                                                _f32ptr(vals), arr), "upload_coo")
        self.N, self.dims, self.nnz, self.R = T.num_modes, tuple(T.dims), T.nnz, 0

    def attach_host(self, T: Tensor, chunk_nnz: int = 0) -> None:
        """Stream T from host memory instead of uploading it (out-of-HBM
        tensors); T's arrays are kept referenced while attached."""
        dims = np.array(T.dims, dtype=np.uint32)
        cols = [np.ascontiguousarray(c, dtype=np.uint32) for c in T.idx]
        vals = np.ascontiguousarray(T.vals, dtype=np.float32)
        arr = (L._u32p * len(cols))(*[_u32ptr(c) for c in cols])
        self._host_arrays = (vals, cols)
        self._check(self.lib.dyn_gpu_attach_host_coo(self.h, T.num_modes, T.nnz, _u32ptr(dims),
                                                     _f32ptr(vals), arr, int(chunk_nnz)),
                    "attach_host_coo")
        self.N, self.dims, self.nnz, self.R = T.num_modes, tuple(T.dims), T.nnz, 0

    def attach_dnb(self, path, chunk_nnz: int = 0) -> None:
        """Stream the tensor of a .dnb file from disk, chunk by chunk, once
        per sweep (only the header is read here, then one checking pass)."""
        self._host_arrays = None
        self._check(self.lib.dyn_gpu_attach_dnb(self.h, str(path).encode(), int(chunk_nnz)),
                    "attach_dnb")
        with open(path, "rb") as f:
            head = f.read(88)
        N = int.from_bytes(head[16:20], "little")
        nnz = int.from_bytes(head[20:28], "little")
        dims = tuple(int.from_bytes(head[28 + 4 * n:32 + 4 * n], "little") for n in range(N))
        self.N, self.dims, self.nnz, self.R = N, dims, nnz, 0

    def remap(self, mode: int, secondary: int = -1, want_outputs: bool = True):
        rowptr = np.zeros(self.dims[mode] + 1, dtype=np.uint64) if want_outputs else None
        perm = np.zeros(max(self.nnz, 1), dtype=np.uint32) if want_outputs else None
        self._check(self.lib.dyn_gpu_remap(
            self.h, mode, secondary,
            rowptr.ctypes.data_as(L._u64p) if want_outputs else None,
            _u32ptr(perm) if want_outputs else None), f"remap({mode})")
        if want_outputs:
            return rowptr, perm[: self.nnz]
        return None

    def remap_all(self) -> None:
        self._check(self.lib.dyn_gpu_remap_all(self.h), "remap_all")

    def get_copy(self, mode: int):
        vals = np.zeros(max(self.nnz, 1), dtype=np.float32)
        cols = [np.zeros(max(self.nnz, 1), dtype=np.uint32) for _ in range(self.N)]
        arr = (L._u32p * self.N)(*[None if w == mode else _u32ptr(cols[w]) for w in range(self.N)])
        self._check(self.lib.dyn_gpu_get_copy(self.h, mode, _f32ptr(vals), arr), "get_copy")
        return vals[: self.nnz], {w: cols[w][: self.nnz] for w in range(self.N) if w != mode}

    # -- processed state ---------------------------------------------------
    def save_processed(self, path) -> None:
        """Write the tensor and every built copy's permutation (dyn_gpu_save_processed)."""
        self._check(self.lib.dyn_gpu_save_processed(self.h, str(path).encode()),
                    "save_processed")

    def load_processed(self, path) -> None:
        """Replace the tensor with a processed-state file's; its copies are
        gathered through the stored permutations (no sort)."""
        self._check(self.lib.dyn_gpu_load_processed(self.h, str(path).encode()),
                    "load_processed")
        N, nnz, dims, _ = processed_info(path)
        self.N, self.dims, self.nnz, self.R = N, dims, nnz, 0

    # -- factors -----------------------------------------------------------
    def set_factors(self, Y: list | None, rank: int) -> None:
        if Y is None:
            self._check(self.lib.dyn_gpu_set_factors(self.h, rank, None, rank), "set_factors")
        else:
            mats = [np.ascontiguousarray(y, dtype=np.float32) for y in Y]
            for n, m in enumerate(mats):
                if m.shape != (self.dims[n], rank):
                    raise ValueError(f"factor {n} has shape {m.shape}, expected "
                                     f"{(self.dims[n], rank)}")
            arr = (L._f32p * self.N)(*[_f32ptr(m) for m in mats])
            self._check(self.lib.dyn_gpu_set_factors(self.h, rank, arr, rank), "set_factors")
        self.R = rank

    def get_factors(self) -> list:
        out = []
        for n in range(self.N):
            y = np.zeros((self.dims[n], self.R), dtype=np.float32)
            self._check(self.lib.dyn_gpu_get_factors(self.h, n, _f32ptr(y), self.R), "get_factors")
            out.append(y)
        return out

    def get_lambda(self) -> np.ndarray:
        lam = np.zeros(self.R, dtype=np.float64)
        self._check(self.lib.dyn_gpu_get_lambda(self.h, lam.ctypes.data_as(L._f64p)), "get_lambda")
        return lam

    # -- kernels -----------------------------------------------------------
    def mttkrp(self, mode: int, precision: int = L.DYN_PREC_FP32) -> np.ndarray:
        out = np.zeros((self.dims[mode], self.R), dtype=np.float32)
        self._check(self.lib.dyn_gpu_mttkrp(self.h, mode, precision, _f32ptr(out), self.R),
                    f"mttkrp({mode})")
        return out

    def mttkrp_all(self, precision: int = L.DYN_PREC_FP32) -> list:
        outs = [np.zeros((self.dims[n], self.R), dtype=np.float32) for n in range(self.N)]
        arr = (L._f32p * self.N)(*[_f32ptr(o) for o in outs])
        self._check(self.lib.dyn_gpu_mttkrp_all(self.h, precision, arr, self.R), "mttkrp_all")
        return outs

    def build_linear(self, want_perm: bool = True):
        """Build the Z-order copy; returns its permutation (Z-order position ->
        input position) when want_perm."""
        perm = np.zeros(max(self.nnz, 1), dtype=np.uint32) if want_perm else None
        self._check(self.lib.dyn_gpu_build_linear(self.h, _u32ptr(perm) if want_perm else None),
                    "build_linear")
        return perm[: self.nnz] if want_perm else None

    def set_layout(self, layout: int) -> None:
        """DYN_LAYOUT_COPIES (per-mode copies) or DYN_LAYOUT_LINEAR (one
        Z-order copy, single-pass all-modes spMTTKRP)."""
        self._check(self.lib.dyn_gpu_set_layout(self.h, int(layout)), "set_layout")

    def gram(self, mode: int) -> np.ndarray:
        G = np.zeros((self.R, self.R), dtype=np.float64)
        self._check(self.lib.dyn_gpu_gram(self.h, mode, G.ctypes.data_as(L._f64p)), "gram")
        return G

    def forward(self, use_lambda: bool = False) -> np.ndarray:
        m = np.zeros(max(self.nnz, 1), dtype=np.float32)
        self._check(self.lib.dyn_gpu_forward(self.h, int(use_lambda), _f32ptr(m)), "forward")
        return m[: self.nnz]

    def adjoint(self, g: np.ndarray, mode: int) -> np.ndarray:
        g = np.ascontiguousarray(g, dtype=np.float32)
        out = np.zeros((self.dims[mode], self.R), dtype=np.float32)
        self._check(self.lib.dyn_gpu_adjoint(self.h, _f32ptr(g), mode, _f32ptr(out), self.R),
                    "adjoint")
        return out

    def enable_timing(self, on: bool = True) -> None:
        self._check(self.lib.dyn_gpu_enable_timing(self.h, int(on)), "enable_timing")

    def kernel_stats(self):
        n = C.c_uint64(0)
        ms = C.c_double(0)
        self._check(self.lib.dyn_gpu_kernel_stats(self.h, C.byref(n), C.byref(ms)), "kernel_stats")
        return int(n.value), float(ms.value)

    def run_cpals(self, max_iters: int, tol: float, ridge: float = 1e-10, kernel: int = 0,
                  precision: int = 0, verbose: bool = False) -> "CpResult":
        o = L.CpOpts()
        o.max_iters, o.tol, o.ridge = int(max_iters), float(tol), float(ridge)
        o.kernel, o.precision, o.verbose = int(kernel), int(precision), int(verbose)
        hist = np.zeros(max(max_iters, 1), dtype=np.float64)
        itim = np.zeros(max(max_iters, 1), dtype=np.float64)
        r = L.CpResultC()
        r.fit_history = hist.ctypes.data_as(L._f64p)
        r.iter_time_s = itim.ctypes.data_as(L._f64p)
        self._check(self.lib.dyn_gpu_cpals(self.h, C.byref(o), C.byref(r)), "cpals")
        k = r.iters_done
        return CpResult(iters_done=k, final_fit=r.final_fit, final_residual=r.final_residual,
                        tensor_norm=r.tensor_norm, total_time_s=r.total_time_s,
                        mttkrp_time_s=r.mttkrp_time_s, linalg_time_s=r.linalg_time_s,
                        remap_time_s=r.remap_time_s, fit_history=list(hist[:k]),
                        iter_time_s=list(itim[:k]), lambda_=self.get_lambda())


def linear_plan(dims):
    """(mask, bits, shift, key_bits) of the Z-order key for mode sizes `dims`
    (needs no device)."""
    lib = L.load_library()
    N = len(dims)
    d = np.asarray(dims, dtype=np.uint32)
    mask = (C.c_uint64 * N)()
    bits = (C.c_uint8 * N)()
    shift = (C.c_uint8 * N)()
    total = lib.dyn_linear_plan(N, _u32ptr(d), mask, bits, shift)
    if total < 0:
        raise ValueError(f"invalid mode sizes {dims}")
    return list(mask), list(bits), list(shift), int(total)


def processed_info(path):
    """(num_modes, nnz, dims, copies_mask) of a processed-state file; raises
    ValueError if `path` is not one (needs no device)."""
    lib = L.load_library()
    n = C.c_int(0)
    nnz = C.c_uint64(0)
    dims = (C.c_uint32 * L.MAX_MODES)()
    mask = C.c_uint32(0)
    rc = lib.dyn_gpu_processed_info(str(path).encode(), C.byref(n), C.byref(nnz), dims,
                                    C.byref(mask))
    if rc != L.DYN_OK:
        raise ValueError(f"{path} is not a processed-state file (status {rc})")
    return n.value, int(nnz.value), tuple(int(dims[i]) for i in range(n.value)), int(mask.value)


# ---------------------------------------------------------------------------
# reference-shaped API
# ---------------------------------------------------------------------------
GaussSeidel = "gauss-seidel"
Jacobi = "jacobi"


@dataclass
class CpOptions:
    """Mirror of CpOptions (include/dynasor_cpals.h:82-101) plus device fields."""
    max_iters: int = 50
    tol: float = 1e-4
    seed: int = 42
    num_threads: int = 0          # accepted for source compatibility; unused on the device
    ridge: float = 1e-10
    verbose: bool = True
    kernel: str = GaussSeidel     # "gauss-seidel" | "jacobi"
    init_dump_prefix: str = ""    # writes <prefix>.cp_init_fm<n>.bin (DYFM)
    precision: str = "fp32"       # "fp32" | "bf16" (BF16 gathered operand, FP32 accumulate)
    init: str = "normal"          # "normal" (reference stream) | "uniform" (U[0,1), same seed)


@dataclass
class CpResult:
    """Mirror of CpResult (include/dynasor_cpals.h:103-114)."""
    iters_done: int = 0
    final_fit: float = 0.0
    final_residual: float = 0.0
    tensor_norm: float = 0.0
    total_time_s: float = 0.0
    mttkrp_time_s: float = 0.0
    linalg_time_s: float = 0.0
    remap_time_s: float = 0.0
    fit_history: list = field(default_factory=list)
    iter_time_s: list = field(default_factory=list)
    lambda_: np.ndarray | None = None


def _kernel_code(k: str) -> int:
    k = k.lower()
    if k in ("gauss-seidel", "gs", "classical", "per-mode", "per_mode"):
        return L.DYN_CP_GAUSS_SEIDEL
    if k in ("jacobi", "js", "all-modes", "all_modes"):
        return L.DYN_CP_JACOBI
    raise ValueError(f"unknown CP kernel {k!r}")


def _prec_code(p: str) -> int:
    p = p.lower()
    if p in ("fp32", "plain", "f32"):
        return L.DYN_PREC_FP32
    if p in ("bf16", "mixed"):
        return L.DYN_PREC_BF16
    raise ValueError(f"unknown precision {p!r}")


def init_factors(dims, rank: int, seed: int, kind: str = "normal") -> list:
    """Initial factors.  "normal" is the reference's CP-ALS stream (std::mt19937_64
    seeded per mode by a splitmix64 step of (seed, mode), std::normal_distribution
    <float>), produced by the host library so it is the same libstdc++ stream;
    "uniform" is U[0,1) from one std::mt19937_64(seed) over the modes in order."""
    from . import host
    return host.init_factors(dims, rank, seed, kind)


def cpals(T: Tensor, rank: int, opt: CpOptions | None = None, ctx: GpuContext | None = None,
          factors: list | None = None):
    """Run CP-ALS on the device.  Returns (CpResult, factors, lambda)."""
    opt = opt or CpOptions()
    if T.num_modes < 2:
        raise ValueError(f"need num_modes >= 2, got {T.num_modes}")
    if rank <= 0:
        raise ValueError(f"need rank >= 1, got {rank}")
    own = ctx is None
    ctx = ctx or GpuContext()
    try:
        if ctx.dims != tuple(T.dims) or ctx.nnz != T.nnz:
            ctx.upload(T)
        Y0 = factors if factors is not None else init_factors(T.dims, rank, opt.seed, opt.init)
        if opt.init_dump_prefix:
            from .tensor_io import save_dyfm
            for n, y in enumerate(Y0):
                save_dyfm(f"{opt.init_dump_prefix}.cp_init_fm{n}.bin", y)
        ctx.set_factors(Y0, rank)
        res = ctx.run_cpals(opt.max_iters, opt.tol, opt.ridge, _kernel_code(opt.kernel),
                            _prec_code(opt.precision), opt.verbose)
        return res, ctx.get_factors(), res.lambda_
    finally:
        if own:
            ctx.close()
