"""The linearised (Z-order) layout and its single-pass all-modes spMTTKRP.

* The Z-order copy is an index computation, so it must be bit-exact: its
  permutation equals a stable argsort of the key that numpy builds here
  independently (frequency ranks with ties by index, the key plan's bits
  deposited LSB-first), and the copy's columns are the input's through it.
* The pass accumulates in FP32 with atomics outside the hottest rows, so its
  outputs are checked like the per-mode kernel's: within 1e-5 Frobenius-
  relative of the FP64 oracle MTTKRP (src/dynasor_reference.cpp:18-50), on
  2- to 8-mode tensors, padded ranks 16 to 128, Zipf-skewed rows (private LDS
  rows in use) and a mode small enough to be private whole.
* CP-ALS on the linear copy runs Jacobi (the reference's Morton path does the
  same, src/dynasor_cpals.cpp:232-240) and must track the FP64 oracle's
  Jacobi fits like the per-mode path does.
"""
import sys

import numpy as np
import pytest

from dynasor_plus_amd import host
from dynasor_plus_amd import _lib as L
from dynasor_plus_amd.engine import linear_plan
from tests.test_gpu_parity import frob_rel, skewed_tensor

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def ranks_of(idx, dim):
    """Frequency rank of every index: rank 0 = most nonzeros, ties by index."""
    cnt = np.bincount(idx, minlength=dim)
    order = np.argsort(-cnt.astype(np.int64), kind="stable")
    rank = np.empty(dim, dtype=np.uint64)
    rank[order] = np.arange(dim, dtype=np.uint64)
    return rank


def zorder_perm(T):
    mask, bits, shift, total = linear_plan(T.dims)
    key = np.zeros(T.nnz, dtype=np.uint64)
    for n in range(T.num_modes):
        r = ranks_of(T.idx[n], T.dims[n])[T.idx[n]] >> np.uint64(shift[n])
        j = 0
        for b in range(64):
            if (mask[n] >> b) & 1:
                key |= ((r >> np.uint64(j)) & np.uint64(1)) << np.uint64(b)
                j += 1
    return np.argsort(key, kind="stable").astype(np.uint32), key


@pytest.mark.parametrize("dims,nnz,gen", [
    ((300, 200, 100), 60000, "zipf"),
    ((60, 50, 40, 30), 40000, "uniform"),
    ((500_000, 100_000, 50_000, 5_000, 100), 200_000, "zipf"),   # 72 index bits
    ((7, 5), 30, "uniform")])
def test_zorder_permutation_bit_exact(gpu_ctx, dims, nnz, gen):
    say("zorder", dims, nnz)
    T, _, _ = host.generate(dims, nnz, seed=21, coords=gen, alpha=1.1, values="lognormal")
    gpu_ctx.upload(T)
    perm = gpu_ctx.build_linear()
    want, key = zorder_perm(T)
    np.testing.assert_array_equal(perm, want)
    ks = key[perm]
    assert np.all(ks[1:] >= ks[:-1])


@pytest.mark.parametrize("dims,nnz,R", [
    ((300, 200, 100), 60000, 32), ((60, 50, 40, 30), 30000, 64), ((30, 24, 20, 16, 12), 20000, 16),
    ((200, 150), 9000, 20), ((100, 80, 60), 40000, 100), ((20, 16, 12, 10, 8, 6), 15000, 8),
    ((8, 6, 5, 4, 4, 3, 3), 12000, 16), ((6, 5, 5, 4, 4, 3, 3, 3), 20000, 16)])
def test_linear_mttkrp_all_vs_oracle(gpu_ctx, oracle, dims, nnz, R):
    say("linear all-modes", dims, nnz, R)
    rng = np.random.default_rng(nnz + R + 1)
    T = skewed_tensor(rng, dims, nnz)
    Y = [rng.standard_normal((d, R)).astype(np.float32) for d in dims]
    gpu_ctx.upload(T)
    gpu_ctx.set_factors(Y, R)
    try:
        gpu_ctx.set_layout(L.DYN_LAYOUT_LINEAR)
        got = gpu_ctx.mttkrp_all()
        for n in range(len(dims)):
            err = frob_rel(got[n], oracle.mttkrp_f64(T, Y, n))
            assert err <= 1e-5, (n, err)
        # one mode at a time (the pass restricted to that mode), any precision
        for n in range(len(dims)):
            one = gpu_ctx.mttkrp(n, L.DYN_PREC_BF16)
            assert frob_rel(one, oracle.mttkrp_f64(T, Y, n)) <= 1e-5
    finally:
        gpu_ctx.set_layout(L.DYN_LAYOUT_COPIES)
    # back on the per-mode copies (rebuilt on demand)
    for n in range(len(dims)):
        assert frob_rel(gpu_ctx.mttkrp(n), oracle.mttkrp_f64(T, Y, n)) <= 1e-5


@pytest.mark.parametrize("dims,nnz,alpha,R", [
    ((5000, 4000, 100), 2_000_000, 1.1, 32),      # a 100-row mode: private whole
    ((20000, 3000, 2000, 50), 1_500_000, 1.2, 16),
    ((3000, 2500, 2000), 1_000_000, 1.0, 64)])
def test_linear_hot_rows_zipf(gpu_ctx, oracle, dims, nnz, alpha, R):
    """Zipf marginals put most updates on ranks held in the waves' private LDS
    tiles; the rest go to global atomics."""
    say("linear zipf", dims, nnz, alpha, R)
    T, _, _ = host.generate(dims, nnz, seed=23, coords="zipf", alpha=alpha, values="lognormal")
    rng = np.random.default_rng(23)
    Y = [rng.random((d, R), dtype=np.float32) for d in dims]
    gpu_ctx.upload(T)
    gpu_ctx.set_factors(Y, R)
    try:
        gpu_ctx.set_layout(L.DYN_LAYOUT_LINEAR)
        got = gpu_ctx.mttkrp_all()
    finally:
        gpu_ctx.set_layout(L.DYN_LAYOUT_COPIES)
    for n in range(len(dims)):
        err = frob_rel(got[n], oracle.mttkrp_f64(T, Y, n))
        assert err <= 1e-5, (n, err)
    # the column invariant: sum_i Yhat_n[i, r] Y_n[i, r] = <T, [[Y]]>_r for every n
    inv = [np.einsum("ir,ir->r", got[n].astype(np.float64), Y[n].astype(np.float64))
           for n in range(len(dims))]
    for n in range(1, len(dims)):
        np.testing.assert_allclose(inv[n], inv[0], rtol=2e-5)


@pytest.mark.parametrize("dims,nnz,R", [((60, 50, 40, 30), 30000, 8), ((80, 70, 60), 40000, 16)])
def test_linear_cpals_jacobi_vs_fp64_oracle(gpu_ctx, oracle, dims, nnz, R):
    say("linear cpals", dims, nnz, R)
    rng = np.random.default_rng(R + 31)
    T = skewed_tensor(rng, dims, nnz, heavy_frac=0.2)
    Y0 = [rng.standard_normal((d, R)).astype(np.float32) for d in T.dims]
    fits, _, lam = oracle.cpals(T, Y0, 8, 0.0, jacobi=True, fp64=True)
    gpu_ctx.upload(T)
    gpu_ctx.set_factors(Y0, R)
    try:
        gpu_ctx.set_layout(L.DYN_LAYOUT_LINEAR)
        res = gpu_ctx.run_cpals(8, 0.0, kernel=L.DYN_CP_JACOBI)
        np.testing.assert_allclose(res.fit_history, fits, rtol=0, atol=1e-5)
        np.testing.assert_allclose(res.lambda_, lam, rtol=1e-3)
        # a Gauss-Seidel request on the linear copy runs Jacobi (with a warning)
        gpu_ctx.set_factors(Y0, R)
        gs = gpu_ctx.run_cpals(8, 0.0, kernel=L.DYN_CP_GAUSS_SEIDEL, precision=L.DYN_PREC_BF16)
        np.testing.assert_allclose(gs.fit_history, fits, rtol=0, atol=1e-5)
    finally:
        gpu_ctx.set_layout(L.DYN_LAYOUT_COPIES)
    # the per-mode copies give the same Jacobi trajectory
    gpu_ctx.set_factors(Y0, R)
    ref = gpu_ctx.run_cpals(8, 0.0, kernel=L.DYN_CP_JACOBI)
    np.testing.assert_allclose(res.fit_history, ref.fit_history, rtol=0, atol=1e-5)


def test_linear_edge_cases(gpu_ctx, oracle):
    say("linear edges")
    from dynasor_plus_amd import Tensor
    rng = np.random.default_rng(37)
    # duplicates (summed), a size-1 mode, empty rows, fewer nonzeros than a wave batch
    idx = [np.array([0, 0, 0, 5, 5], dtype=np.uint32), np.array([0, 0, 0, 0, 0], dtype=np.uint32),
           np.array([1, 1, 2, 0, 3], dtype=np.uint32)]
    T = Tensor((9, 1, 4), np.array([1, 2, 3, 4, 5], dtype=np.float32), idx)
    Y = [rng.standard_normal((d, 3)).astype(np.float32) for d in T.dims]
    gpu_ctx.upload(T)
    gpu_ctx.set_factors(Y, 3)
    try:
        gpu_ctx.set_layout(L.DYN_LAYOUT_LINEAR)
        got = gpu_ctx.mttkrp_all()
    finally:
        gpu_ctx.set_layout(L.DYN_LAYOUT_COPIES)
    for n in range(3):
        assert frob_rel(got[n], oracle.mttkrp_f64(T, Y, n)) <= 1e-6
    # a bad layout code is rejected
    with pytest.raises(L.DynGpuError):
        gpu_ctx.set_layout(7)
