"""dyn_gpu_mttkrp_noalloc (the entry point that allocates nothing) against
dyn_gpu_mttkrp on the same context: the same kernels in the same order, with
the heavy-row partials in the fixed per-GPU array instead of context scratch,
so the outputs must be bitwise equal.  Two inputs in a row, the second
smaller, both with heavy rows (Zipf coordinates) so the partials are used;
the context's device memory must not grow across the calls."""
import ctypes as C
import sys

import numpy as np
import pytest

from dynasor_plus_amd import GpuContext, host
from dynasor_plus_amd import _lib as L
from dynasor_plus_amd.engine import padded_rank

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def noalloc(ctx, mode, prec, buf, ld):
    return ctx.lib.dyn_gpu_mttkrp_noalloc(ctx.h, mode, prec, C.c_void_p(buf.data_ptr()), ld)


def test_noalloc_matches_mttkrp_on_two_inputs():
    torch = pytest.importorskip("torch")
    cases = [((4000, 3000, 5000), 400_000, 32, 1.1),
             ((900, 700, 300, 50), 60_000, 16, 1.2)]
    with GpuContext(0) as ctx:
        for dims, nnz, R, alpha in cases:
            say("noalloc", dims, nnz, R)
            T, _, _ = host.generate(dims, nnz, seed=11, coords="zipf", alpha=alpha,
                                    values="lognormal")
            ctx.upload(T)
            rng = np.random.default_rng(3)
            ctx.set_factors([rng.random((d, R), dtype=np.float32) for d in dims], R)
            Rp = padded_rank(R)
            buf = torch.empty((max(dims), Rp), dtype=torch.float32, device="cuda")
            # nothing prepared yet: the call must refuse rather than allocate
            torch.cuda.synchronize()
            assert noalloc(ctx, 0, L.DYN_PREC_FP32, buf, Rp) == L.DYN_ERR_STATE
            for prec in (L.DYN_PREC_FP32, L.DYN_PREC_BF16):
                for mode in range(len(dims)):
                    want = ctx.mttkrp(mode, prec)          # builds / orders the copy
                    used = ctx.mem_used()
                    buf.fill_(float("nan"))
                    torch.cuda.synchronize()
                    ctx._check(noalloc(ctx, mode, prec, buf, Rp), "mttkrp_noalloc")
                    ctx.synchronize()
                    got = buf[: dims[mode], :R].cpu().numpy()
                    assert ctx.mem_used() == used
                    np.testing.assert_array_equal(got, want)
            assert noalloc(ctx, 0, L.DYN_PREC_FP32, buf, Rp * 2) == L.DYN_ERR_INVALID
