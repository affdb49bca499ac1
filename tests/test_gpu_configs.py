"""Workload-shaped checks on the device.

  * configuration 2 at full size (12092 x 9184 x 28818, 77M nnz, rank 32):
    size-independent properties (row pointer = bincount, sortedness,
    permutation) and the spMTTKRP column invariant, plus the reference's
    first three Gauss-Seidel fits (golden, printed to 6 digits);
  * configuration-3 shape (4 modes, Zipf(1.2), rank 64) scaled down so the
    FP64 oracle finishes: plain CP-ALS vs the FP64 oracle, BF16-gathered
    CP-ALS vs the FP64 oracle with BF16-rounded gathers;
  * configuration-4 shape (5 modes, Zipf(1.0), planted rank 20 + 5% noise,
    rank 32, Jacobi) scaled down: true fit vs the FP64 Jacobi oracle.
"""
import json
import sys

import numpy as np
import pytest

from dynasor_plus_amd import GpuContext, host, init_factors
from dynasor_plus_amd import _lib as L
from tests._oracle import load as load_oracle
from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def tensor_checksum(T):
    """SHA-256 over dims, values and index columns (the recipe record's
    checksum, tests/golden/make_golden.py)."""
    import hashlib
    h = hashlib.sha256()
    h.update(np.array(T.dims, dtype=np.uint32).tobytes())
    h.update(np.ascontiguousarray(T.vals, dtype=np.float32).tobytes())
    for c in T.idx:
        h.update(np.ascontiguousarray(c, dtype=np.uint32).tobytes())
    return h.hexdigest()


@pytest.fixture(scope="module")
def c2():
    rec = json.loads((GOLDEN / "cpals_c2_gauss-seidel_r32.json").read_text())
    T, _, _ = host.generate(tuple(rec["dims"]), rec["nnz"], seed=rec["seed"], coords="zipf",
                            alpha=rec["alpha"], values=rec["values"])
    return T, rec


def test_c2_remap_properties(c2):
    say("c2 remap")
    T, _ = c2
    with GpuContext(0) as ctx:
        ctx.upload(T)
        for mode in range(3):
            rowptr, perm = ctx.remap(mode, (mode + 1) % 3)
            counts = np.bincount(T.idx[mode], minlength=T.dims[mode])
            np.testing.assert_array_equal(rowptr, np.concatenate([[0], np.cumsum(counts)]))
            assert np.bincount(perm, minlength=T.nnz).max() == 1      # a permutation
            key = T.idx[mode][perm]
            assert (np.diff(key.astype(np.int64)) >= 0).all()


def test_c2_mttkrp_column_invariant(c2):
    """sum_i Yhat_n[i, r] Y_n[i, r] = <T, [[Y]]_r> is the same for every mode."""
    say("c2 invariant")
    T, _ = c2
    rng = np.random.default_rng(0)
    Y = [rng.random((d, 32), dtype=np.float32) for d in T.dims]
    with GpuContext(0) as ctx:
        ctx.upload(T)
        ctx.set_factors(Y, 32)
        cols = [np.einsum("ir,ir->r", ctx.mttkrp(n).astype(np.float64), Y[n].astype(np.float64))
                for n in range(3)]
    for n in (1, 2):
        np.testing.assert_allclose(cols[n], cols[0], rtol=2e-5)


def test_c2_fits_track_reference(c2):
    say("c2 fits")
    T, rec = c2
    # the tensor is regenerated from its recipe: it must be the one the
    # reference ran on
    assert tensor_checksum(T) == rec["checksum"]
    with GpuContext(0) as ctx:
        ctx.upload(T)
        ctx.set_factors(init_factors(T.dims, 32, rec["init_seed"], "normal"), 32)
        res = ctx.run_cpals(3, 0.0)
    # the device algebra is FP64: it follows the FP64 oracle closely ...
    np.testing.assert_allclose(res.fit_history, rec["oracle_fp64_fits"], rtol=0, atol=2e-6)
    # ... while the reference's FP32 Gram / Cholesky / lambda drift from the FP64
    # trajectory by 2.7e-5 at iteration 3 on this tensor (the oracle run with the
    # reference's FP32 algebra reproduces the reference's printed fits)
    np.testing.assert_allclose(rec["oracle_fp32_fits"], rec["fits"], rtol=0, atol=6e-6)
    np.testing.assert_allclose(res.fit_history, rec["fits"], rtol=0, atol=4e-5)


def test_c3_shape_mixed_precision_equal_fit():
    say("c3 mixed")
    orc = load_oracle()
    T, _, _ = host.generate((4000, 1500, 400, 60), 250_000, seed=3, coords="zipf", alpha=1.2,
                            values="lognormal")
    R, iters = 64, 10
    Y0 = init_factors(T.dims, R, 42, "uniform")
    fits64, _, _ = orc.cpals(T, Y0, iters, 0.0, fp64=True)
    fitsbf, _, _ = orc.cpals(T, Y0, iters, 0.0, fp64=True, bf16_gather=True)
    with GpuContext(0) as ctx:
        ctx.upload(T)
        ctx.set_factors(Y0, R)
        plain = ctx.run_cpals(iters, 0.0, precision=L.DYN_PREC_FP32)
        ctx.set_factors(Y0, R)
        mixed = ctx.run_cpals(iters, 0.0, precision=L.DYN_PREC_BF16)
    say("fits", list(plain.fit_history), list(mixed.fit_history), fitsbf.tolist())
    np.testing.assert_allclose(plain.fit_history, fits64, rtol=0, atol=1e-5)
    # The BF16-gathered run follows the oracle run with BF16-rounded gathers.
    # Only the first sweep starts from identical FP32 factors; afterwards FP32
    # accumulation differences (~1e-7) flip BF16 roundings at ties, each a 2^-9
    # relative jump, so the trajectories separate by ~1e-4 over 10 sweeps
    # (the oracle's own BF16 and FP64 runs differ by as much).
    assert abs(mixed.fit_history[0] - fitsbf[0]) < 1e-5, (mixed.fit_history[0], fitsbf[0])
    np.testing.assert_allclose(mixed.fit_history, fitsbf, rtol=0, atol=2e-4)
    # equal fit within the same spread: plain vs mixed on this tensor
    assert abs(mixed.final_fit - plain.final_fit) < 5e-4, (mixed.final_fit, plain.final_fit)


@pytest.mark.parametrize("prec", [L.DYN_PREC_FP32, L.DYN_PREC_BF16])
def test_c4_shape_jacobi_true_fit(prec):
    """Jacobi with the reference's update order (all spMTTKRPs from the
    iteration-start factors, Grams refreshed mode by mode) can diverge on this
    shape, so trajectories are only comparable at the first iteration; the fit
    the device reports must then equal the true FP64 fit of its own model.
    Both operand precisions: FP32 against the FP64 oracle, BF16 against the
    FP64 oracle with BF16-rounded gathers (the configured C4 run)."""
    say("c4 jacobi", "bf16" if prec else "fp32")
    orc = load_oracle()
    T, _, _ = host.generate((3000, 1000, 500, 80, 20), 150_000, seed=4, coords="zipf", alpha=1.0,
                            values="planted", planted_rank=20, noise_frac=0.05)
    R = 32
    # N(0,1) init keeps lambda moderate (U[0,1) init drives it to ~1e10 here,
    # where the fit's three-term difference cancels to noise)
    Y0 = init_factors(T.dims, R, 42, "normal")
    fits64, _, _ = orc.cpals(T, Y0, 1, 0.0, jacobi=True, fp64=True, bf16_gather=bool(prec))
    with GpuContext(0) as ctx:
        ctx.upload(T)
        ctx.set_factors(Y0, R)
        first = ctx.run_cpals(1, 0.0, kernel=L.DYN_CP_JACOBI, precision=prec)
        ctx.set_factors(Y0, R)
        res = ctx.run_cpals(4, 0.0, kernel=L.DYN_CP_JACOBI, precision=prec)
        Y = [y.astype(np.float64) for y in ctx.get_factors()]
        lam = ctx.get_lambda()
    say("fits", list(res.fit_history))
    # a clamped residual would report a perfect fit for a diverged model
    assert not any(f == 1.0 for f in res.fit_history)
    # FP32 spMTTKRP vs the FP64 oracle, amplified by the diverging update
    assert abs(first.fit_history[0] - fits64[0]) <= 1e-4 * max(1.0, abs(fits64[0]))
    v = T.vals.astype(np.float64)
    t2 = float(v @ v)
    inner = float(v @ orc.forward(T, Y, lam))
    H = np.ones((R, R))
    for y in Y:
        H *= y.T @ y
    m2 = float(lam @ H @ lam)
    fit = 1.0 - np.sqrt(max(0.0, t2 - 2 * inner + m2)) / np.sqrt(t2)
    assert abs(res.final_fit - fit) <= 1e-5 * max(1.0, abs(fit)), (res.final_fit, fit)
