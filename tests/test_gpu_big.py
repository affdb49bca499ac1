"""A tensor with more than 2^30 nonzeros resident on one device.

Its value and index columns exceed 4 GiB, so the spMTTKRP reads them through
per-piece buffer ranges instead of one launch-wide range.  Checked through
linearity (the spMTTKRP of a tensor is the sum of those of two halves of its
nonzeros, each below 2^30 and so read the ordinary way), in FP32 and BF16,
on the per-mode copies and on the linearised copy, and one CP-ALS sweep runs
to a finite fit.  Coordinates are drawn with repeats, which the MTTKRP sums.
"""
import sys

import numpy as np
import pytest

from dynasor_plus_amd import GpuContext, Tensor
from dynasor_plus_amd import _lib as L

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_more_than_2_30_nonzeros():
    dims = (100_000, 50_000, 8_000_000)       # mode-2 rows of ~140 nonzeros: light segments
    nnz = (1 << 30) + (1 << 26)
    say("big: generate", nnz)
    rng = np.random.default_rng(31)
    idx = [np.empty(nnz, np.uint32) for _ in dims]
    vals = np.empty(nnz, np.float32)
    step = 1 << 27
    for b in range(0, nnz, step):
        e = min(nnz, b + step)
        for n, d in enumerate(dims):
            idx[n][b:e] = rng.integers(0, d, size=e - b, dtype=np.uint32)
        vals[b:e] = rng.random(e - b, dtype=np.float32)
    T = Tensor(dims, vals, idx)
    half = nnz // 2
    A = Tensor(dims, vals[:half], [c[:half] for c in idx])
    B = Tensor(dims, vals[half:], [c[half:] for c in idx])
    R = 16
    Y = [rng.random((d, R), dtype=np.float32) for d in dims]
    with GpuContext(0) as ctx:
        parts = {}
        for name, S in (("A", A), ("B", B)):
            say("big: half", name)
            ctx.upload(S)
            ctx.set_factors(Y, R)
            parts[name] = {p: [ctx.mttkrp(n, p) for n in range(3)]
                           for p in (L.DYN_PREC_FP32, L.DYN_PREC_BF16)}
        say("big: whole tensor, per-mode copies")
        ctx.upload(T)
        ctx.set_factors(Y, R)
        whole = {p: [ctx.mttkrp(n, p) for n in range(3)]
                 for p in (L.DYN_PREC_FP32, L.DYN_PREC_BF16)}
        say("big: whole tensor, linearised copy")
        ctx.set_layout(L.DYN_LAYOUT_LINEAR)
        lin = ctx.mttkrp_all()
        ctx.set_layout(L.DYN_LAYOUT_COPIES)
        say("big: one CP-ALS sweep")
        ctx.set_factors(Y, R)
        res = ctx.run_cpals(1, 0.0)
    for p in (L.DYN_PREC_FP32, L.DYN_PREC_BF16):
        for n in range(3):
            want = parts["A"][p][n].astype(np.float64) + parts["B"][p][n]
            err = rel(whole[p][n], want)
            say("prec", p, "mode", n, "rel", err)
            assert err < 1e-5, (p, n, err)
    for n in range(3):
        assert rel(lin[n], whole[L.DYN_PREC_FP32][n]) < 1e-5, n
    assert res.iters_done == 1 and np.isfinite(res.final_fit)
