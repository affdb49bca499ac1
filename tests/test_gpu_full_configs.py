"""The parity-test configurations at their full sizes, through properties that
do not need an oracle run over the whole tensor:

  * configuration 3 (4-mode 1M x 200k x 20k x 500, 150M Zipf(1.2) coordinates,
    rank 64): BF16-gathered CP-ALS reaches the plain FP32 run's fit within
    1e-4 at every iteration from the same initial factors, and the spMTTKRP
    column invariant holds for both precisions;
  * configuration 4 (5-mode 500k x 100k x 50k x 5k x 100, 400M Zipf(1.0)
    coordinates, planted rank 20 + 5 % noise, rank 32): row pointers equal the
    index bincounts in every mode, the column invariant over all five modes,
    one Jacobi and one Gauss-Seidel sweep run to finite fits and lambdas, and
    the configured 30-iteration Jacobi run (BF16 gathers, and FP32) reports
    no clamped fit of 1.0 and a final fit equal to the FP64 true fit of the
    model it returns, or NaN where that fit is not finite (Jacobi diverges on
    this tensor, a property of the reference's update order);
  * configuration 4 on the linearised (Z-order) copy: the single-pass
    all-modes spMTTKRP against the per-mode copies', the column invariant,
    and two Jacobi iterations against the per-mode path;
  * configuration 5 (1M x 100k x 2k, 200M uniform coordinates, Bernoulli
    labels from a planted rank-16 N(0, 0.5) model): 100 Adam steps at rank 32
    on the fused loss/adjoint, the final device loss against the FP64 oracle
    loss, and 10240 device gradient entries (every column of 320 seeded rows)
    against FP64 central differences of the loss restricted to the nonzeros
    that the entry touches (the other terms do not depend on it); the rows'
    nonzeros come from the device remap's row pointers and permutation.
"""
import sys

import numpy as np
import pytest

from dynasor_plus_amd import GpuContext, host, init_factors
from dynasor_plus_amd import _lib as L
from tests._oracle import load as load_oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def column_products(ctx, Y, N):
    """sum_i Yhat_n[i, r] Y_n[i, r] for every mode n (FP64 on the host)."""
    return [np.einsum("ir,ir->r", ctx.mttkrp(n).astype(np.float64), Y[n].astype(np.float64))
            for n in range(N)]


def test_c3_full_mixed_equal_fit():
    say("c3 full: generate")
    T, _, _ = host.generate((1_000_000, 200_000, 20_000, 500), 150_000_000, seed=3, coords="zipf",
                            alpha=1.2, values="lognormal", threads=16)
    R, iters = 64, 3
    Y0 = init_factors(T.dims, R, 42, "normal")
    with GpuContext(0) as ctx:
        say("c3 full: upload + remap")
        ctx.upload(T)
        ctx.remap_all()
        # positive factors: the column sums have no cancellation
        rng = np.random.default_rng(0)
        U = [rng.random((d, R), dtype=np.float32) for d in T.dims]
        ctx.set_factors(U, R)
        cols = column_products(ctx, U, 4)
        for n in range(1, 4):
            np.testing.assert_allclose(cols[n], cols[0], rtol=5e-5)
        ctx.set_factors(Y0, R)
        say("c3 full: cpals plain / bf16")
        plain = ctx.run_cpals(iters, 0.0, precision=L.DYN_PREC_FP32)
        ctx.set_factors(Y0, R)
        mixed = ctx.run_cpals(iters, 0.0, precision=L.DYN_PREC_BF16)
    say("fits", list(plain.fit_history), list(mixed.fit_history))
    assert np.all(np.isfinite(plain.fit_history)) and plain.fit_history[-1] > plain.fit_history[0]
    np.testing.assert_allclose(mixed.fit_history, plain.fit_history, rtol=0, atol=1e-4)


def test_c4_full_remap_invariant_and_jacobi():
    say("c4 full: generate")
    T, _, _ = host.generate((500_000, 100_000, 50_000, 5_000, 100), 400_000_000, seed=4,
                            coords="zipf", alpha=1.0, values="planted", planted_rank=20,
                            noise_frac=0.05, threads=16)
    R = 32
    Y0 = init_factors(T.dims, R, 42, "normal")
    with GpuContext(0) as ctx:
        say("c4 full: upload + remap")
        ctx.upload(T)
        for mode in range(5):
            rowptr, _ = ctx.remap(mode, 1 if mode == 0 else 0)
            counts = np.bincount(T.idx[mode], minlength=T.dims[mode])
            assert rowptr[0] == 0 and rowptr[-1] == T.nnz
            np.testing.assert_array_equal(np.diff(rowptr.astype(np.int64)), counts)
        rng = np.random.default_rng(0)
        U = [rng.random((d, R), dtype=np.float32) for d in T.dims]
        ctx.set_factors(U, R)
        say("c4 full: column invariant")
        cols = column_products(ctx, U, 5)
        for n in range(1, 5):
            np.testing.assert_allclose(cols[n], cols[0], rtol=5e-5)
        say("c4 full: one Jacobi and one Gauss-Seidel sweep")
        ctx.set_factors(Y0, R)
        jac = ctx.run_cpals(1, 0.0, kernel=L.DYN_CP_JACOBI)
        lam_j = ctx.get_lambda()
        ctx.set_factors(Y0, R)
        gs = ctx.run_cpals(1, 0.0, kernel=L.DYN_CP_GAUSS_SEIDEL)
        runs = {}
        for prec in (L.DYN_PREC_BF16, L.DYN_PREC_FP32):
            say("c4 full: 30 Jacobi iterations", "bf16" if prec else "fp32")
            ctx.set_factors(Y0, R)
            res = ctx.run_cpals(30, 0.0, kernel=L.DYN_CP_JACOBI, precision=prec)
            runs[prec] = (res, [y.astype(np.float64) for y in ctx.get_factors()],
                          ctx.get_lambda())
    assert np.isfinite(jac.final_fit) and np.isfinite(gs.final_fit)
    assert np.all(np.isfinite(lam_j)) and np.all(lam_j >= 0)
    # the configured run: every reported fit is the true fit of the model or
    # NaN (flagged divergence), never a clamped perfect fit; the last one is
    # checked against the FP64 true fit of the returned model
    orc = load_oracle()
    v = T.vals.astype(np.float64)
    t2 = float(v @ v)
    for prec, (res, Y, lam) in runs.items():
        say("fits", "bf16" if prec else "fp32", list(res.fit_history))
        assert res.iters_done == 30
        assert not any(f == 1.0 for f in res.fit_history)
        say("c4 full: FP64 true fit of the returned model")
        m = orc.forward(T, Y, lam)
        inner = float(v @ m)
        S = float(np.abs(v) @ np.abs(m))
        H = np.ones((R, R))
        Habs = np.ones((R, R))
        for y in Y:
            H *= y.T @ y
            Habs *= np.abs(y).T @ np.abs(y)
        m2 = float(lam @ H @ lam)
        m2abs = float(np.abs(lam) @ Habs @ np.abs(lam))
        r2 = t2 - 2 * inner + m2
        say("t2", t2, "inner", inner, "sum|v m|", S, "m2", m2, "|m|2", m2abs, "r2", r2)
        if not (np.isfinite(r2) and r2 >= -1e-4 * max(t2, abs(2 * inner), abs(m2))):
            assert np.isnan(res.final_fit), (res.final_fit, t2, inner, m2)
            continue
        fit = 1.0 - np.sqrt(max(r2, 0.0)) / np.sqrt(t2)
        # Forward error of the engine's fit: <T, M> comes from FP32 spMTTKRP
        # sums (error about 1e-5 of sum |v m| at these row lengths) and
        # ||M||^2 from Gram matrices of the FP32 factors with FP32 MFMA
        # partials (error about 1e-6 of |lam|' (*_n |A_n|'|A_n|) |lam|);
        # d fit = d(r2) / (2 |r| |T|).  For a well-conditioned model both
        # terms are far below 1e-5; a diverged one (huge, nearly cancelling
        # components: the fit is a small difference of large terms) is
        # determined only to that.
        den = np.sqrt(max(r2, 1e-300)) * np.sqrt(t2)
        tol = 1e-5 * max(1.0, abs(fit)) + (1e-5 * S + 0.5e-6 * m2abs) / den
        assert abs(res.final_fit - fit) <= tol, (res.final_fit, fit, tol)


def test_c4_full_linear_layout():
    """The linearised copy of the 5-mode 400M tensor (72 index bits, so the key
    drops the largest modes' lowest rank bits): its single-pass all-modes
    spMTTKRP equals the per-mode copies' to FP32 summation-order rounding in
    every mode, the column invariant holds, and Jacobi from the same initial
    factors tracks the per-mode path."""
    say("c4 linear: generate")
    T, _, _ = host.generate((500_000, 100_000, 50_000, 5_000, 100), 400_000_000, seed=4,
                            coords="zipf", alpha=1.0, values="planted", planted_rank=20,
                            noise_frac=0.05, threads=16)
    R = 32
    rng = np.random.default_rng(0)
    U = [rng.random((d, R), dtype=np.float32) for d in T.dims]
    Y0 = init_factors(T.dims, R, 42, "normal")
    with GpuContext(0) as ctx:
        say("c4 linear: upload + Z-order copy")
        ctx.upload(T)
        ctx.set_factors(U, R)
        ctx.set_layout(L.DYN_LAYOUT_LINEAR)
        lin = ctx.mttkrp_all()
        ctx.set_factors(Y0, R)
        say("c4 linear: 2 Jacobi iterations")
        jl = ctx.run_cpals(2, 0.0, kernel=L.DYN_CP_JACOBI)
        say("c4 linear: per-mode copies")
        ctx.set_layout(L.DYN_LAYOUT_COPIES)
        ctx.set_factors(U, R)
        cop = ctx.mttkrp_all()
        ctx.set_factors(Y0, R)
        jc = ctx.run_cpals(2, 0.0, kernel=L.DYN_CP_JACOBI)
    for n in range(5):
        a, b = lin[n].astype(np.float64), cop[n].astype(np.float64)
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-5, n
    cols = [np.einsum("ir,ir->r", lin[n].astype(np.float64), U[n].astype(np.float64))
            for n in range(5)]
    for n in range(1, 5):
        np.testing.assert_allclose(cols[n], cols[0], rtol=5e-5)
    say("fits", list(jl.fit_history), list(jc.fit_history))
    np.testing.assert_allclose(jl.fit_history, jc.fit_history, rtol=0, atol=1e-5)


def test_c5_full_completion_adam_and_gradients():
    torch = pytest.importorskip("torch")
    from dynasor_plus_amd.autograd import completion_loss_and_grad, fit_completion
    say("c5 full: generate")
    T, _, _ = host.generate((1_000_000, 100_000, 2_000), 200_000_000, seed=5, coords="uniform",
                            values="bernoulli", planted_rank=16, planted_factors="normal",
                            planted_var=0.5, threads=16)
    orc = load_oracle()
    steps, R = 100, 32
    rng = np.random.default_rng(7)
    # the seeded gradient subset: 320 (mode, row) pairs, every column of each
    picks = [(int(rng.integers(0, 3)), None) for _ in range(320)]
    picks = [(n, int(rng.integers(0, T.dims[n]))) for n, _ in picks]
    with GpuContext(0) as eng:
        say("c5 full: upload")
        eng.upload(T)
        say("c5 full: adam")
        A, losses = fit_completion(eng, T.dims, rank=R, steps=steps, lr=1e-2, seed=0)
        final, grads = completion_loss_and_grad(eng, A, "bernoulli")
        A64 = [a.cpu().numpy().astype(np.float64) for a in A]
        G = [g.cpu().numpy() for g in grads]
        say("c5 full: row segments of the picked rows")
        segs = {}
        for n in range(3):
            rowptr, perm = eng.remap(n)
            for m, i in picks:
                if m == n:
                    segs[(n, i)] = perm[int(rowptr[i]):int(rowptr[i + 1])].astype(np.int64)
    assert len(losses) == steps and np.all(np.isfinite(losses))
    assert losses[-1] < 0.95 * losses[0]
    say("c5 full: oracle loss")
    L64 = orc.completion_loss(T, A64, 1)
    assert abs(final - L64) / L64 < 1e-5, (final, L64)
    say("c5 full: restricted central differences on 10240 gradient entries")
    h = 1e-4
    y = T.vals.astype(np.float64)
    fd, ad = [], []
    for n, i in picks:
        k = segs[(n, i)]
        assert np.all(T.idx[n][k] == i)
        if k.size == 0:
            continue
        rows = [A64[w][T.idx[w][k]] for w in range(3)]
        m = (rows[0] * rows[1] * rows[2]).sum(axis=1)
        others = [w for w in range(3) if w != n]
        q = rows[others[0]] * rows[others[1]]          # dm / dA_n[i, r] per nonzero
        yk = y[k]

        def loss_at(mm):
            return float(np.sum(np.logaddexp(0.0, mm) - yk * mm))

        for r in range(R):
            fd.append((loss_at(m + h * q[:, r]) - loss_at(m - h * q[:, r])) / (2 * h))
            ad.append(float(G[n][i, r]))
    fd, ad = np.array(fd), np.array(ad)
    say("fd entries", fd.size)
    assert fd.size >= 9000
    assert np.linalg.norm(ad - fd) / np.linalg.norm(fd) < 1e-4
