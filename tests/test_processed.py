"""Processed-state files ("DGP1": the device counterpart of the reference's
.dnp, include/dynasor_dnp.h): the header check on the CPU, and on the GPU a
save / reload round trip whose copies and spMTTKRP results are bit-exact with
the saved context's, rejection of damaged permutations, and the CLI's
--save-processed / --runtime pair giving the same CP-ALS fits."""
import struct
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from dynasor_plus_amd import host
from dynasor_plus_amd import _lib as L
from dynasor_plus_amd.engine import processed_info

ROOT = Path(__file__).resolve().parents[1]
HEADER = 4 + 5 * 4 + 8 + 8 * 4 + 8 * 4 + 8 * 4


def write_dgp(path, dims, vals, idx, perms, secondary=None, magic=b"DGP1", version=1,
              truncate=0):
    """A processed-state file written from its published layout (header
    fields in order, then vals, idx[N], one permutation per stored copy)."""
    N, nnz = len(dims), len(vals)
    mask = sum(1 << n for n in perms)
    sec = [-1] * 8
    for n, s in (secondary or {}).items():
        sec[n] = s
    h = magic + struct.pack("<5I", version, 4, 4, N, mask) + struct.pack("<Q", nnz)
    h += struct.pack("<8I", *(list(dims) + [0] * (8 - N))) + struct.pack("<8i", *sec)
    h += struct.pack("<8I", *([0] * 8))
    assert len(h) == HEADER
    body = np.asarray(vals, np.float32).tobytes()
    body += b"".join(np.asarray(c, np.uint32).tobytes() for c in idx)
    body += b"".join(np.asarray(perms[n], np.uint32).tobytes() for n in sorted(perms))
    data = h + body
    Path(path).write_bytes(data[: len(data) - truncate])


def tiny():
    dims = (4, 4, 3)
    idx = [np.array([0, 0, 1, 3, 2]), np.array([1, 3, 0, 2, 2]), np.array([0, 2, 1, 1, 0])]
    vals = np.array([1.0, 2.0, 3.0, 4.0, 5.0])
    perm0 = np.argsort(idx[0], kind="stable")
    return dims, vals, idx, perm0


def test_processed_info_reads_the_header(tmp_path):
    dims, vals, idx, perm0 = tiny()
    p = tmp_path / "t.dgp"
    write_dgp(p, dims, vals, idx, {0: perm0}, secondary={0: 1})
    assert processed_info(p) == (3, 5, dims, 1)


@pytest.mark.parametrize("damage", ["magic", "version", "truncated", "mask", "secondary"])
def test_processed_info_rejects_damaged_files(tmp_path, damage):
    dims, vals, idx, perm0 = tiny()
    p = tmp_path / "t.dgp"
    kw = {}
    perms = {0: perm0}
    if damage == "magic":
        kw["magic"] = b"DNP1"
    elif damage == "version":
        kw["version"] = 2
    elif damage == "truncated":
        kw["truncate"] = 4
    elif damage == "secondary":
        kw["secondary"] = {0: 0}
    write_dgp(p, dims, vals, idx, perms, **kw)
    if damage == "mask":
        b = bytearray(p.read_bytes())
        b[4 + 4 * 4:4 + 5 * 4] = struct.pack("<I", 1 << 5)
        p.write_bytes(bytes(b))
    with pytest.raises(ValueError):
        processed_info(p)


@pytest.mark.gpu
def test_save_reload_is_bit_exact(tmp_path):
    from dynasor_plus_amd import GpuContext
    print("[case] processed round trip", flush=True)
    rng = np.random.default_rng(2)
    for dims, nnz, R in (((3000, 2000, 4000), 300_000, 32), ((700, 500, 300, 40), 50_000, 16)):
        T, _, _ = host.generate(dims, nnz, seed=9, coords="zipf", alpha=1.1, values="lognormal")
        Y = [rng.random((d, R), dtype=np.float32) for d in dims]
        p = tmp_path / f"t{len(dims)}.dgp"
        with GpuContext(0) as a:
            a.upload(T)
            a.remap_all()
            a.set_factors(Y, R)
            want = {prec: [a.mttkrp(n, prec) for n in range(len(dims))]
                    for prec in (L.DYN_PREC_FP32, L.DYN_PREC_BF16)}
            copies = [a.get_copy(n) for n in range(len(dims))]
            a.save_processed(p)
        assert processed_info(p) == (len(dims), nnz, dims, (1 << len(dims)) - 1)
        with GpuContext(0) as b:
            b.load_processed(p)
            for n in range(len(dims)):
                v, ix = b.get_copy(n)
                np.testing.assert_array_equal(v, copies[n][0])
                for w in ix:
                    np.testing.assert_array_equal(ix[w], copies[n][1][w])
            b.set_factors(Y, R)
            for prec, outs in want.items():
                for n in range(len(dims)):
                    np.testing.assert_array_equal(b.mttkrp(n, prec), outs[n])


@pytest.mark.gpu
def test_damaged_permutations_are_rejected(tmp_path):
    from dynasor_plus_amd import GpuContext
    print("[case] processed damaged", flush=True)
    T, _, _ = host.generate((300, 200, 100), 20_000, seed=4, coords="zipf", alpha=1.1,
                            values="lognormal")
    perm = np.argsort(T.idx[1], kind="stable").astype(np.uint32)
    dup = perm.copy()
    dup[1] = dup[0]                                   # not a permutation
    unsorted = perm[::-1].copy()                      # a permutation that does not sort
    big = perm.copy()
    big[7] = T.nnz + 3                                # out of range
    with GpuContext(0) as ctx:
        good = tmp_path / "good.dgp"
        write_dgp(good, T.dims, T.vals, T.idx, {1: perm}, secondary={1: 0})
        ctx.load_processed(good)
        v, _ = ctx.get_copy(1)
        np.testing.assert_array_equal(v, T.vals[perm])
        for name, bad in (("dup", dup), ("unsorted", unsorted), ("big", big)):
            p = tmp_path / f"{name}.dgp"
            write_dgp(p, T.dims, T.vals, T.idx, {1: bad}, secondary={1: 0})
            with pytest.raises(L.DynGpuError) as e:
                ctx.load_processed(p)
            assert e.value.code == L.DYN_ERR_INVALID, name


@pytest.mark.gpu
def test_cli_save_processed_then_runtime(tmp_path):
    print("[case] processed cli", flush=True)
    exe = ROOT / "dynasor_plus_amd" / "lib" / "dynasor_plus_gpu"
    T, _, _ = host.generate((400, 300, 200), 40_000, seed=6, coords="zipf", alpha=1.1,
                            values="lognormal")
    from dynasor_plus_amd.tensor_io import save_dnb
    t = tmp_path / "t.tns"
    save_dnb(str(t) + ".dnb", T)
    proc = tmp_path / "t.dgp"
    common = ["--decompose", "cpals", "--rank", "8", "--cp-iters", "6", "--cp-tol", "0"]
    a = subprocess.run([str(exe), str(t), *common, "--save-processed", str(proc)],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert a.returncode == 0, a.stdout
    b = subprocess.run([str(exe), str(proc), "--runtime", *common], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    assert b.returncode == 0, b.stdout
    fits = lambda out: [l.split("fit = ")[1].split()[0] for l in out.splitlines()
                        if l.startswith("[cpals] iter")]
    assert fits(a.stdout) == fits(b.stdout) and len(fits(a.stdout)) == 6
    assert "[runtime] loading processed state" in b.stdout
