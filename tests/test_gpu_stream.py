"""Out-of-HBM tensors streamed from host memory (dyn_gpu_attach_host_coo)
or from a .dnb file (dyn_gpu_attach_dnb), the device counterpart of the
reference's out-of-core Jacobi driver (src/dynasor_ooc.cpp:340-535): the
streamed all-modes spMTTKRP equals the in-core one up to the regrouping of
FP32 sums (chunk partials added in order) and is the same bit for bit from
either source, streamed Jacobi CP-ALS follows the FP64 oracle, a
Gauss-Seidel request runs Jacobi, indices and file headers are checked,
entry points that need the whole tensor refuse, and the CLI's --ooc on path
(from the .dnb beside the tensor, or from host memory with --no-cache)
reproduces the in-core fits."""
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from dynasor_plus_amd import GpuContext, host, init_factors
from dynasor_plus_amd import _lib as L
from dynasor_plus_amd.tensor_io import Tensor, save_dnb, save_tns
from tests._oracle import load as load_oracle

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def frob_rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("dims,nnz,R,chunks", [((900, 700, 500), 120_000, 32, 7),
                                               ((800, 600, 300, 40), 90_000, 16, 5)])
def test_streamed_mttkrp_all_matches_in_core(dims, nnz, R, chunks):
    say("stream mttkrp_all", dims)
    T, _, _ = host.generate(dims, nnz, seed=21, coords="zipf", alpha=1.1, values="lognormal")
    rng = np.random.default_rng(1)
    Y = [rng.random((d, R), dtype=np.float32) for d in dims]
    with GpuContext(0) as a:
        a.upload(T)
        a.set_factors(Y, R)
        want = {p: a.mttkrp_all(p) for p in (L.DYN_PREC_FP32, L.DYN_PREC_BF16)}
    with GpuContext(0) as b:
        b.attach_host(T, chunk_nnz=nnz // chunks + 1)
        b.set_factors(Y, R)
        for p, outs in want.items():
            got = b.mttkrp_all(p)
            for n in range(len(dims)):
                assert frob_rel(got[n], outs[n]) < 1e-6, (p, n)
            np.testing.assert_allclose(b.mttkrp(1, p), got[1], rtol=0, atol=0)


@pytest.mark.parametrize("dims,nnz,R,chunk", [((900, 700, 500), 100_000, 32, 30_001),
                                              ((500, 400, 300, 60, 7), 80_000, 16, 80_000)])
def test_dnb_stream_equals_host_stream(tmp_path, dims, nnz, R, chunk):
    say("stream dnb", dims, chunk)
    T, _, _ = host.generate(dims, nnz, seed=33, coords="zipf", alpha=1.2, values="lognormal")
    path = tmp_path / "x.dnb"
    save_dnb(path, T)
    rng = np.random.default_rng(4)
    Y = [rng.standard_normal((d, R)).astype(np.float32) for d in dims]
    with GpuContext(0) as a:
        a.upload(T)
        a.set_factors(Y, R)
        core = a.mttkrp_all(L.DYN_PREC_FP32)
    with GpuContext(0) as h, GpuContext(0) as f:
        h.attach_host(T, chunk_nnz=chunk)
        f.attach_dnb(path, chunk_nnz=chunk)
        assert f.nnz == nnz and f.dims == tuple(dims)
        for ctx in (h, f):
            ctx.set_factors(Y, R)
        for p in (L.DYN_PREC_FP32, L.DYN_PREC_BF16):
            got_h, got_f = h.mttkrp_all(p), f.mttkrp_all(p)
            for n in range(len(dims)):
                np.testing.assert_array_equal(got_f[n], got_h[n])
                if p == L.DYN_PREC_FP32:
                    assert frob_rel(got_f[n], core[n]) < 1e-6, n
        # the same Jacobi run from either source
        rh = h.run_cpals(3, 0.0, kernel=L.DYN_CP_JACOBI)
        rf = f.run_cpals(3, 0.0, kernel=L.DYN_CP_JACOBI)
        np.testing.assert_array_equal(rf.fit_history, rh.fit_history)


def test_dnb_stream_rejects_bad_files(tmp_path):
    say("stream dnb refusals")
    T, _, _ = host.generate((50, 40, 30), 2_000, seed=2, coords="uniform", values="lognormal")
    good = tmp_path / "g.dnb"
    save_dnb(good, T)
    raw = good.read_bytes()
    cases = {"missing": None, "magic": b"XNB1" + raw[4:], "truncated": raw[:-4],
             "short_header": raw[:40]}
    with GpuContext(0) as ctx:
        for name, data in cases.items():
            p = tmp_path / f"{name}.dnb"
            if data is not None:
                p.write_bytes(data)
            with pytest.raises(L.DynGpuError) as e:
                ctx.attach_dnb(p)
            assert e.value.code == L.DYN_ERR_INVALID, name
        bad = [c.copy() for c in T.idx]
        bad[0][5] = 50
        bp = tmp_path / "oob.dnb"
        save_dnb(bp, Tensor(T.dims, T.vals, bad))
        with pytest.raises(L.DynGpuError) as e:
            ctx.attach_dnb(bp, chunk_nnz=700)
        assert e.value.code == L.DYN_ERR_INVALID
        ctx.attach_dnb(good, chunk_nnz=700)   # the context is usable afterwards


def test_streamed_jacobi_tracks_fp64_oracle():
    say("stream jacobi")
    orc = load_oracle()
    from tests.test_gpu_parity import skewed_tensor
    rng = np.random.default_rng(15)
    # the in-core Jacobi parity tensor (Jacobi converges on it; on tensors
    # where the reference's Jacobi order diverges, FP32 trajectories separate)
    T = skewed_tensor(rng, (60, 50, 40, 30), 30000, heavy_frac=0.2)
    R = 8
    Y0 = [rng.standard_normal((d, R)).astype(np.float32) for d in T.dims]
    fits, _, lam = orc.cpals(T, Y0, 8, 0.0, jacobi=True, fp64=True)
    with GpuContext(0) as ctx:
        ctx.attach_host(T, chunk_nnz=7_001)
        ctx.set_factors(Y0, R)
        res = ctx.run_cpals(8, 0.0, kernel=L.DYN_CP_JACOBI)
        ctx.set_factors(Y0, R)
        gs = ctx.run_cpals(8, 0.0, kernel=L.DYN_CP_GAUSS_SEIDEL)   # runs Jacobi
    np.testing.assert_allclose(res.fit_history, fits, rtol=0, atol=1e-5)
    np.testing.assert_allclose(res.lambda_, lam, rtol=1e-3)
    np.testing.assert_allclose(gs.fit_history, res.fit_history, rtol=0, atol=0)


def test_streamed_checks_and_refusals():
    say("stream checks")
    T, _, _ = host.generate((300, 200, 100), 20_000, seed=3, coords="uniform", values="lognormal")
    bad = [c.copy() for c in T.idx]
    bad[2][-1] = 100                                   # one index past mode 2's size
    with GpuContext(0) as ctx:
        with pytest.raises(L.DynGpuError) as e:
            ctx.attach_host(Tensor(T.dims, T.vals, bad), chunk_nnz=3_000)
        assert e.value.code == L.DYN_ERR_INVALID
        ctx.attach_host(T, chunk_nnz=3_000)
        ctx.set_factors(init_factors(T.dims, 8, 1, "normal"), 8)
        for call in (lambda: ctx.remap(0), lambda: ctx.forward(),
                     lambda: ctx.get_copy(0)):
            with pytest.raises(L.DynGpuError) as e:
                call()
            assert e.value.code == L.DYN_ERR_STATE
        # an in-core upload afterwards works normally
        ctx.upload(T)
        ctx.set_factors(init_factors(T.dims, 8, 1, "normal"), 8)
        ctx.remap(0)


def test_cli_ooc_matches_in_core_jacobi(tmp_path):
    say("stream cli")
    exe = ROOT / "dynasor_plus_amd" / "lib" / "dynasor_plus_gpu"
    T, _, _ = host.generate((400, 300, 200, 50), 60_000, seed=6, coords="zipf", alpha=1.1,
                            values="lognormal")
    t = tmp_path / "t.tns"
    save_dnb(str(t) + ".dnb", T)
    common = [str(exe), str(t), "--decompose", "cpals", "--rank", "8", "--cp-iters", "5",
              "--cp-tol", "0", "--cp-kernel", "jacobi"]
    a = subprocess.run(common, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    b = subprocess.run(common + ["--ooc", "on", "--ooc-chunk", "9000"], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    # the same tensor as text without a cache: streamed from host memory
    u = tmp_path / "u.tns"
    save_tns(u, T)
    c = subprocess.run([str(exe), str(u)] + common[2:] + ["--no-cache", "--ooc", "on",
                                                         "--ooc-chunk", "9000"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert a.returncode == 0 and b.returncode == 0 and c.returncode == 0, \
        (a.stdout[-800:], b.stdout[-800:], c.stdout[-800:])
    assert "OUT-OF-CORE: the tensor stays in" in b.stdout
    assert f"[ooc] streaming the tensor from {t}.dnb" in b.stdout
    assert "[ooc] streaming the tensor from host memory" in c.stdout
    fits = lambda out: [float(l.split("fit = ")[1].split()[0]) for l in out.splitlines()
                        if l.startswith("[cpals] iter")]
    assert len(fits(a.stdout)) == 5
    np.testing.assert_allclose(fits(b.stdout), fits(a.stdout), rtol=0, atol=2e-6)
    np.testing.assert_allclose(fits(c.stdout), fits(b.stdout), rtol=0, atol=0)
