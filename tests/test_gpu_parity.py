"""Device engine vs the oracle and the reference's golden vectors.

Integer/index work (remap permutation, row pointer, copies) must be bit-exact;
FP32 spMTTKRP must agree with the reference --verify outputs within 1e-5
Frobenius-relative (the reference's own check is 1e-4, src/main.cpp:955-968);
the plain CP-ALS fit must track the reference per iteration within 1e-6
(plus half a unit of the printed %.6f).
"""
import sys

import numpy as np
import pytest

from dynasor_plus_amd import GpuContext, Tensor, host
from dynasor_plus_amd import _lib as L
from tests._oracle import tensor_from_npz
from tests.conftest import golden

pytestmark = pytest.mark.gpu

VERIFY = ["verify_3d_7_r8.npz", "verify_3d_7_r16.npz", "verify_3d_7_r32.npz",
          "verify_3d_7_r64.npz", "verify_synth_cp_r5_r8.npz", "verify_g4_r16.npz",
          "verify_g5_r32.npz"]


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def frob_rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def skewed_tensor(rng, dims, nnz, heavy_frac=0.4):
    """Distinct coordinates; up to a `heavy_frac` share of the nonzeros in row 0
    of mode 0 (capped at half of that row's cells), so rows span lane groups,
    workgroups and the fix-up pass."""
    dims = tuple(int(d) for d in dims)
    cells = int(np.prod(dims, dtype=np.float64))
    sub = int(np.prod(dims[1:], dtype=np.float64))
    nh = min(int(nnz * heavy_frac), sub // 2)
    heavy = rng.choice(sub, size=nh, replace=False).astype(np.int64)   # row 0: linear < sub
    chosen = heavy
    while chosen.shape[0] < nnz:
        need = nnz - chosen.shape[0]
        cand = np.unique(rng.integers(0, cells, size=2 * need + 16, dtype=np.int64))
        cand = np.setdiff1d(cand, chosen, assume_unique=False)
        rng.shuffle(cand)
        chosen = np.concatenate([chosen, cand[:need]])
    rng.shuffle(chosen)
    idx = [i.astype(np.uint32) for i in np.unravel_index(chosen, dims)]
    vals = rng.standard_normal(nnz).astype(np.float32)
    return Tensor(dims, vals, idx)


@pytest.mark.parametrize("name", VERIFY)
def test_mttkrp_matches_reference_verify(gpu_ctx, oracle, name):
    say("verify", name)
    z = golden(name)
    T = tensor_from_npz(z)
    R = int(z["R"])
    Y = oracle.init(T.dims, R, int(z["seed"]), kind="uniform")
    gpu_ctx.upload(T)
    gpu_ctx.set_factors(Y, R)
    for n in range(T.num_modes):
        got = gpu_ctx.mttkrp(n)
        err = frob_rel(got, z[f"yhat{n}"])
        assert err <= 1e-5, (name, n, err)


@pytest.mark.parametrize("dims,nnz,R", [
    ((40, 30, 20), 5000, 5), ((64, 48, 32), 20000, 16), ((300, 200, 100), 60000, 32),
    ((50, 40, 30, 20), 30000, 64), ((30, 20, 16, 12, 8), 20000, 32), ((200, 150), 9000, 20),
    ((20, 16, 12, 10, 8, 6), 15000, 8), ((100, 80, 60), 40000, 100),
    ((8, 6, 5, 4, 4, 3, 3), 12000, 16), ((6, 5, 5, 4, 4, 3, 3, 3), 20000, 16)])
def test_mttkrp_vs_oracle_skewed(gpu_ctx, oracle, dims, nnz, R):
    say("mttkrp skewed", dims, nnz, R)
    rng = np.random.default_rng(nnz + R)
    T = skewed_tensor(rng, dims, nnz)
    Y = [rng.standard_normal((d, R)).astype(np.float32) for d in dims]
    gpu_ctx.upload(T)
    gpu_ctx.set_factors(Y, R)
    for n in range(len(dims)):
        want = oracle.mttkrp_f64(T, Y, n)
        got = gpu_ctx.mttkrp(n)
        assert frob_rel(got, want) <= 1e-5, (n, frob_rel(got, want))


def test_mttkrp_hot_split_and_precision_switches(gpu_ctx, oracle):
    """3-mode launches reorder each row hot-first by the next mode's rank
    (rows beyond the LDS capacity stay in memory); FP32 and BF16 launches
    share the split copy, BF16 ones with operand 0's leading rows in the LDS
    half that BF16 rows leave free.  Operand modes larger than the LDS rows
    (1272 at rank 32) give hot and cold heavy pieces and mixed LDS / memory
    gathers of operand 0; Zipf marginals give light rows too.  Alternating
    precisions on one context must stay exact, and every copy must remain a
    row-grouped permutation of the input."""
    say("mttkrp hot split")
    T, _, _ = host.generate((300, 4000, 5000), 400_000, seed=17, coords="zipf", alpha=1.1,
                            values="lognormal")
    R = 32
    rng = np.random.default_rng(17)
    Y = [rng.standard_normal((d, R)).astype(np.float32) for d in T.dims]

    def to_bf16(a):
        u = a.astype(np.float32).view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
        return (u.astype(np.uint32) << 16).view(np.float32)

    want32 = [oracle.mttkrp_f64(T, Y, n) for n in range(3)]
    want16 = [oracle.mttkrp_f64(T, [to_bf16(y) for y in Y], n) for n in range(3)]
    gpu_ctx.upload(T)
    gpu_ctx.set_factors(Y, R)
    for prec in (L.DYN_PREC_FP32, L.DYN_PREC_BF16, L.DYN_PREC_FP32):
        for n in range(3):
            got = gpu_ctx.mttkrp(n, prec)
            want = want32[n] if prec == L.DYN_PREC_FP32 else want16[n]
            assert frob_rel(got, want) <= 1e-5, (prec, n, frob_rel(got, want))
    # after the FP32 launches each copy is split; it still holds the input's
    # (other coordinates, value) multiset
    for n in range(3):
        vals, idx = gpu_ctx.get_copy(n)
        others = [w for w in range(3) if w != n]
        got = np.lexsort([vals.view(np.uint32)] + [idx[w] for w in others][::-1])
        ref = np.lexsort([T.vals.view(np.uint32)] + [T.idx[w] for w in others][::-1])
        for w in others:
            np.testing.assert_array_equal(idx[w][got], T.idx[w][ref])
        np.testing.assert_array_equal(vals[got], T.vals[ref])


@pytest.mark.parametrize("dims", [(120, 90, 70), (60, 50, 40, 30), (30, 24, 20, 16, 12)])
def test_mttkrp_bf16_operands(gpu_ctx, oracle, dims):
    say("mttkrp bf16", dims)
    rng = np.random.default_rng(11)
    T = skewed_tensor(rng, dims, 30000)
    R = 32
    Y = [rng.standard_normal((d, R)).astype(np.float32) for d in T.dims]
    gpu_ctx.upload(T)
    gpu_ctx.set_factors(Y, R)

    def to_bf16(a):   # round-to-nearest-even to bfloat16, back to float
        u = a.astype(np.float32).view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
        return (u.astype(np.uint32) << 16).view(np.float32)

    Yb = [to_bf16(y) for y in Y]
    for n in range(len(dims)):
        got = gpu_ctx.mttkrp(n, L.DYN_PREC_BF16)
        assert frob_rel(got, oracle.mttkrp_f64(T, Yb, n)) <= 1e-5
        assert frob_rel(got, oracle.mttkrp_f64(T, Y, n)) <= 2e-2


@pytest.mark.parametrize("dims,R", [((3000, 2500, 2000, 40), 32),
                                    ((2000, 1500, 1000, 300), 64),
                                    ((4000, 3000, 800, 60, 25), 32)])
def test_mttkrp_lds_whole_and_partial_operands(gpu_ctx, oracle, dims, R):
    """Zipf-skewed coordinates (with repeats, which the MTTKRP sums) on 4- and
    5-mode shapes whose small modes the launcher holds whole in LDS (BF16 rows
    then read by ds_read) while the large ones are split between LDS ranks and
    memory; rank 64 BF16 also reuses the secondary mode's row.  Both operand
    precisions against the FP64 oracle (BF16 on BF16-rounded factors)."""
    say("mttkrp lds whole/partial", dims, R)
    rng = np.random.default_rng(41)
    nnz = 200_000
    idx = [np.minimum(rng.zipf(1.3, nnz) - 1, d - 1).astype(np.uint32) for d in dims]
    for w, d in enumerate(dims):   # scatter the Zipf ranks over the mode
        idx[w] = rng.permutation(d).astype(np.uint32)[idx[w]]
    vals = rng.standard_normal(nnz).astype(np.float32)
    T = Tensor(dims, vals, idx)
    Y = [rng.standard_normal((d, R)).astype(np.float32) for d in dims]
    gpu_ctx.upload(T)
    gpu_ctx.set_factors(Y, R)

    def to_bf16(a):
        u = a.astype(np.float32).view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
        return (u.astype(np.uint32) << 16).view(np.float32)

    Yb = [to_bf16(y) for y in Y]
    for n in range(len(dims)):
        got = gpu_ctx.mttkrp(n, L.DYN_PREC_FP32)
        assert frob_rel(got, oracle.mttkrp_f64(T, Y, n)) <= 1e-5, n
        got = gpu_ctx.mttkrp(n, L.DYN_PREC_BF16)
        assert frob_rel(got, oracle.mttkrp_f64(T, Yb, n)) <= 1e-5, n


def test_mttkrp_all_hot_class_split(gpu_ctx, oracle):
    """BF16 launches on 4+ modes regroup every row of at least 4096 nonzeros so
    that the nonzeros whose gathered rows all sit in LDS come first (their
    pieces then read LDS only).  The copy stays a permutation of the tensor
    (changed order, same entries); FP32 launches on it, BF16 at another rank
    (a fresh split from the plain order) and back all match the oracle."""
    say("mttkrp all-hot class split")
    rng = np.random.default_rng(43)
    dims = (40, 3000, 2500, 60)
    nnz = 300_000
    idx = [np.minimum(rng.zipf(1.3, nnz) - 1, d - 1).astype(np.uint32) for d in dims]
    for w, d in enumerate(dims):
        idx[w] = rng.permutation(d).astype(np.uint32)[idx[w]]
    vals = rng.standard_normal(nnz).astype(np.float32)
    T = Tensor(dims, vals, idx)
    assert np.bincount(idx[0], minlength=dims[0]).max() >= 4 * 4096   # long rows exist
    gpu_ctx.upload(T)

    def to_bf16(a):
        u = a.astype(np.float32).view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
        return (u.astype(np.uint32) << 16).view(np.float32)

    def entries(v, cols):   # the (value, columns) tuples in a canonical order
        m = np.stack([v.view(np.uint32)] + [cols[w] for w in sorted(cols)])
        return m[:, np.lexsort(m[::-1])]

    for step, R in enumerate((32, 64, 32)):
        Y = [rng.standard_normal((d, R)).astype(np.float32) for d in dims]
        Yb = [to_bf16(y) for y in Y]
        gpu_ctx.set_factors(Y, R)
        if step == 0:
            got = gpu_ctx.mttkrp(0, L.DYN_PREC_FP32)          # plain order
            assert frob_rel(got, oracle.mttkrp_f64(T, Y, 0)) <= 1e-5
            v0, c0 = gpu_ctx.get_copy(0)
        for n in range(len(dims)):
            got = gpu_ctx.mttkrp(n, L.DYN_PREC_BF16)
            assert frob_rel(got, oracle.mttkrp_f64(T, Yb, n)) <= 1e-5, (R, n)
            got = gpu_ctx.mttkrp(n, L.DYN_PREC_FP32)
            assert frob_rel(got, oracle.mttkrp_f64(T, Y, n)) <= 1e-5, (R, n)
        if step == 0:
            v1, c1 = gpu_ctx.get_copy(0)
            assert not all(np.array_equal(c0[w], c1[w]) for w in c0)   # regrouped
            np.testing.assert_array_equal(entries(v0, c0), entries(v1, c1))


def test_mttkrp_bitwise_reproducible(gpu_ctx):
    say("mttkrp reproducible")
    rng = np.random.default_rng(12)
    T = skewed_tensor(rng, (80, 70, 60), 50000, heavy_frac=0.6)
    Y = [rng.standard_normal((d, 32)).astype(np.float32) for d in T.dims]
    gpu_ctx.upload(T)
    gpu_ctx.set_factors(Y, 32)
    a = [gpu_ctx.mttkrp(n) for n in range(3)]
    b = [gpu_ctx.mttkrp(n) for n in range(3)]
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_engine_copy_order(gpu_ctx, oracle):
    """dyn_gpu_remap_all's copies order each row by the secondary mode's
    frequency rank (most nonzeros first, ties by index) except for its 1024
    most frequent ranks, which are scattered; dyn_gpu_remap keeps the
    reference's order (secondary input index).  Both hold the same rows."""
    say("engine copy order")
    rng = np.random.default_rng(44)
    dims = (50, 6000, 300)
    nnz = 120_000
    idx = [np.minimum(rng.zipf(1.2, nnz) - 1, d - 1).astype(np.uint32) for d in dims]
    for w, d in enumerate(dims):
        idx[w] = rng.permutation(d).astype(np.uint32)[idx[w]]
    T = Tensor(dims, rng.standard_normal(nnz).astype(np.float32), idx)
    cnt = np.bincount(idx[1], minlength=dims[1])
    order = np.lexsort((np.arange(dims[1]), -cnt))       # rank -> index
    rank = np.empty(dims[1], np.int64)
    rank[order] = np.arange(dims[1])
    gpu_ctx.upload(T)
    gpu_ctx.remap_all()                                  # copy 0: secondary mode 1
    v_e, c_e = gpu_ctx.get_copy(0)
    rowptr, perm = gpu_ctx.remap(0, 1)                   # the reference's order
    operm, orow = oracle.remap(T, 0, 1)
    np.testing.assert_array_equal(perm, operm)
    v_r, c_r = gpu_ctx.get_copy(0)
    assert not np.array_equal(c_e[1], c_r[1])
    for r in range(dims[0]):
        b, e = int(rowptr[r]), int(rowptr[r + 1])
        np.testing.assert_array_equal(np.sort(c_e[1][b:e]), np.sort(c_r[1][b:e]))
        rk = rank[c_e[1][b:e]]
        cold = rk[rk >= 1024]
        assert np.all(np.diff(cold) >= 0), r


@pytest.mark.parametrize("secondary", [-1, 1])
def test_remap_bit_exact(gpu_ctx, oracle, secondary):
    say("remap", secondary)
    rng = np.random.default_rng(13)
    T = skewed_tensor(rng, (300, 257, 70000), 40000)
    gpu_ctx.upload(T)
    for mode in range(3):
        sec = secondary if secondary != mode else (mode + 1) % 3
        rowptr, perm = gpu_ctx.remap(mode, sec)
        operm, orow = oracle.remap(T, mode, sec)
        np.testing.assert_array_equal(perm, operm)
        np.testing.assert_array_equal(rowptr, orow)
        vals, cols = gpu_ctx.get_copy(mode)
        np.testing.assert_array_equal(vals, T.vals[operm])
        for w, c in cols.items():
            np.testing.assert_array_equal(c, T.idx[w][operm])


def test_gram_mfma(gpu_ctx):
    say("gram")
    rng = np.random.default_rng(14)
    for R, dims in ((16, (1000, 37, 5)), (32, (5000, 3, 64)), (64, (3001, 65, 7)),
                    (128, (2000, 130, 9)), (7, (100, 50, 20))):
        T = skewed_tensor(rng, dims, 200, heavy_frac=0.0)
        Y = [rng.standard_normal((d, R)).astype(np.float32) for d in dims]
        gpu_ctx.upload(T)
        gpu_ctx.set_factors(Y, R)
        for n in range(3):
            Yd = Y[n].astype(np.float64)
            want = Yd.T @ Yd
            got = gpu_ctx.gram(n)
            assert frob_rel(got, want) < 1e-6, (R, n, frob_rel(got, want))


def test_cpals_c1_tracks_reference(gpu_ctx):
    say("cpals c1")
    from tests.test_oracle import c1_tensor
    T, z = c1_tensor()
    Y0 = [z[f"init{n}"] for n in range(3)]
    gpu_ctx.upload(T)
    gpu_ctx.set_factors(Y0, 16)
    res = gpu_ctx.run_cpals(int(z["iters"]), float(z["tol"]))
    ref = z["fits"]
    assert res.iters_done == len(ref)
    np.testing.assert_allclose(res.fit_history, ref, rtol=0, atol=1.5e-6)
    Y = gpu_ctx.get_factors()
    for n in range(3):
        assert frob_rel(Y[n], z[f"final{n}"]) < 1e-3


def test_cpals_3d_7_rank2_tracks_reference(gpu_ctx):
    """The reference's own 9-nonzero tensor at rank 2 (padded to 16 rows of
    zeros on the device): fits, final factors and lambda of its 10
    Gauss-Seidel iterations."""
    say("cpals 3d_7 rank 2")
    z = golden("cpals_3d_7_gauss-seidel_r2.npz")
    T = tensor_from_npz(z)
    gpu_ctx.upload(T)
    gpu_ctx.set_factors([z[f"init{n}"] for n in range(3)], 2)
    res = gpu_ctx.run_cpals(int(z["iters"]), float(z["tol"]))
    np.testing.assert_allclose(res.fit_history, z["fits"], rtol=0, atol=1.5e-6)
    Y = gpu_ctx.get_factors()
    for n in range(3):
        assert frob_rel(Y[n], z[f"final{n}"]) < 1e-4
    np.testing.assert_allclose(res.lambda_, z["lam"], rtol=1e-4)


def test_cpals_singular_gram_zero_pivots(gpu_ctx, oracle):
    """Two identical factor columns with small-integer entries and no ridge:
    every Gram, Hadamard product and Cholesky step of the first solve is exact,
    the second pivot is exactly zero, and the reference's rule (zero pivot ->
    zero component, src/dynasor_linalg.cpp:186-189, 235, 243) zeroes that
    column for good.  The device's explicit inverse must agree with the
    oracle's per-row substitutions, and column 1 must stay exactly zero."""
    say("cpals singular gram")
    z = golden("cpals_3d_7_gauss-seidel_r2.npz")
    T = tensor_from_npz(z)
    base = [np.array([3, 1, 0, 2], np.float32), np.array([1, 0, 0, 0], np.float32),
            np.array([2, 0, 0], np.float32)]
    Y0 = [np.stack([b, b], axis=1) for b in base]      # G_1 o G_2 = [[4, 4], [4, 4]]
    fits, Yo, lam = oracle.cpals(T, Y0, 4, 0.0, ridge=0.0, fp64=True)
    gpu_ctx.upload(T)
    gpu_ctx.set_factors(Y0, 2)
    res = gpu_ctx.run_cpals(4, 0.0, ridge=0.0)
    Y = gpu_ctx.get_factors()
    np.testing.assert_allclose(res.fit_history, fits, rtol=0, atol=1e-6)
    np.testing.assert_allclose(res.lambda_, lam, rtol=1e-5, atol=0)
    assert res.lambda_[1] == 0.0 and lam[1] == 0.0
    for n in range(3):
        assert np.all(Y[n][:, 1] == 0.0)
        np.testing.assert_allclose(Y[n], Yo[n], rtol=0, atol=1e-6)


@pytest.mark.parametrize("jacobi", [False, True])
def test_cpals_vs_fp64_oracle(gpu_ctx, oracle, jacobi):
    say("cpals fp64", jacobi)
    rng = np.random.default_rng(15)
    T = skewed_tensor(rng, (60, 50, 40, 30), 30000, heavy_frac=0.2)
    R = 8
    Y0 = [rng.standard_normal((d, R)).astype(np.float32) for d in T.dims]
    fits, Yo, lam = oracle.cpals(T, Y0, 8, 0.0, jacobi=jacobi, fp64=True)
    gpu_ctx.upload(T)
    gpu_ctx.set_factors(Y0, R)
    res = gpu_ctx.run_cpals(8, 0.0, kernel=int(jacobi))
    np.testing.assert_allclose(res.fit_history, fits, rtol=0, atol=1e-5)
    np.testing.assert_allclose(res.lambda_, lam, rtol=1e-3)


def test_forward_and_adjoint(gpu_ctx, oracle):
    say("forward/adjoint")
    rng = np.random.default_rng(16)
    T = skewed_tensor(rng, (70, 60, 50), 20000)
    R = 24
    A = [rng.standard_normal((d, R)).astype(np.float32) for d in T.dims]
    gpu_ctx.upload(T)
    gpu_ctx.set_factors(A, R)
    m = gpu_ctx.forward(use_lambda=False)
    want = oracle.forward(T, A)
    assert frob_rel(m, want) < 1e-5
    g = rng.standard_normal(T.nnz).astype(np.float32)
    Tg = Tensor(T.dims, g, T.idx)
    for n in range(3):
        got = gpu_ctx.adjoint(g, n)
        assert frob_rel(got, oracle.mttkrp_f64(Tg, A, n)) < 1e-5


def test_edge_cases(gpu_ctx, oracle):
    say("edges")
    rng = np.random.default_rng(17)
    # mode of size 1, empty rows, duplicates summed (MTTKRP is linear)
    idx = [np.array([0, 0, 0, 5, 5], dtype=np.uint32), np.array([0, 0, 0, 0, 0], dtype=np.uint32),
           np.array([1, 1, 2, 0, 3], dtype=np.uint32)]
    T = Tensor((9, 1, 4), np.array([1, 2, 3, 4, 5], dtype=np.float32), idx)
    Y = [rng.standard_normal((d, 3)).astype(np.float32) for d in T.dims]
    gpu_ctx.upload(T)
    gpu_ctx.set_factors(Y, 3)
    for n in range(3):
        assert frob_rel(gpu_ctx.mttkrp(n), oracle.mttkrp_f64(T, Y, n)) < 1e-6
    # no nonzeros at all
    T0 = Tensor((3, 4), np.zeros(0, dtype=np.float32),
                [np.zeros(0, dtype=np.uint32), np.zeros(0, dtype=np.uint32)])
    gpu_ctx.upload(T0)
    gpu_ctx.set_factors([np.ones((3, 2), np.float32), np.ones((4, 2), np.float32)], 2)
    assert not gpu_ctx.mttkrp(0).any()


def test_error_paths(gpu_ctx):
    say("errors")
    T = Tensor((3, 3), np.ones(2, np.float32), [np.array([0, 5], np.uint32),
                                                np.array([1, 2], np.uint32)])
    with pytest.raises(L.DynGpuError) as e:
        gpu_ctx.upload(T)
    assert e.value.code == L.DYN_ERR_INVALID
    T = Tensor((3, 3), np.ones(2, np.float32), [np.array([0, 2], np.uint32),
                                                np.array([1, 2], np.uint32)])
    gpu_ctx.upload(T)
    with pytest.raises(L.DynGpuError) as e:
        gpu_ctx.set_factors(None, 200)
    assert e.value.code == L.DYN_ERR_UNSUPPORTED
    with pytest.raises(L.DynGpuError) as e:
        gpu_ctx.mttkrp(0)
    assert e.value.code == L.DYN_ERR_STATE
    # 2^32 - 1 nonzeros or more (32-bit positions): rejected before any buffer
    # is read; below that, null buffers are
    dims = (L._u32p._type_ * 3)(10, 10, 10)
    rc = gpu_ctx.lib.dyn_gpu_upload_coo(gpu_ctx.h, 3, (1 << 32) - 1, dims, None, None)
    assert rc == L.DYN_ERR_INVALID
    rc = gpu_ctx.lib.dyn_gpu_upload_coo(gpu_ctx.h, 3, 1 << 30, dims, None, None)
    assert rc == L.DYN_ERR_INVALID


def test_memory_cap_reports_oom(monkeypatch):
    say("oom cap")
    monkeypatch.setenv("DYN_GPU_MEM_CAP_MB", "1")
    rng = np.random.default_rng(18)
    T = skewed_tensor(rng, (1000, 1000, 1000), 200000, heavy_frac=0.0)
    with GpuContext(0) as ctx:
        with pytest.raises(L.DynGpuError) as e:
            ctx.upload(T)
        assert e.value.code == L.DYN_ERR_OOM
