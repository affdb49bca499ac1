"""Tensor file formats and the host-side loaders (Python and the C++ host
library), against the reference's own fixtures and writer.  CPU only."""
import ctypes as C
import os
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from dynasor_plus_amd import tensor_io as tio
from tests._oracle import tensor_from_npz
from tests.conftest import GOLDEN, golden

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "dynasor_plus_amd" / "lib"
CLI = LIB / "dynasor_plus_gpu"
REF = ROOT / "oracle" / "_ref" / "dynasor_plus"

D37 = """1 1 1 1.000000
1 2 1 2.000000
1 4 1 10.000000
2 1 2 7.000000
2 2 1 8.000000
2 3 1 6.000000
3 1 3 3.000000
3 3 1 4.000000
4 1 2 5.000000
"""


@pytest.fixture()
def d37(tmp_path):
    p = tmp_path / "3d_7.tns"
    p.write_text(D37)
    return p


def host_lib():
    C.CDLL(str(LIB / "libdynasor_gpu.so"), mode=C.RTLD_GLOBAL)
    lib = C.CDLL(str(LIB / "libdynasor_host.so"))
    lib.dyn_host_tensor_load.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_void_p)]
    lib.dyn_host_tensor_info.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_uint64),
                                         C.POINTER(C.c_uint32)]
    lib.dyn_host_tensor_copy.argtypes = [C.c_void_p, C.POINTER(C.c_float),
                                         C.POINTER(C.POINTER(C.c_uint32))]
    lib.dyn_host_tensor_save_dnb.argtypes = [C.c_void_p, C.c_char_p]
    lib.dyn_host_tensor_free.argtypes = [C.c_void_p]
    return lib


def host_load(path, cache=0):
    lib = host_lib()
    h = C.c_void_p()
    rc = lib.dyn_host_tensor_load(str(path).encode(), cache, C.byref(h))
    if rc != 0:
        return rc, None
    nm, nnz = C.c_int(), C.c_uint64()
    dims = (C.c_uint32 * 8)()
    lib.dyn_host_tensor_info(h, C.byref(nm), C.byref(nnz), dims)
    vals = np.zeros(nnz.value, np.float32)
    cols = [np.zeros(nnz.value, np.uint32) for _ in range(nm.value)]
    arr = (C.POINTER(C.c_uint32) * nm.value)(*[c.ctypes.data_as(C.POINTER(C.c_uint32)) for c in cols])
    lib.dyn_host_tensor_copy(h, vals.ctypes.data_as(C.POINTER(C.c_float)), arr)
    lib.dyn_host_tensor_free(h)
    return 0, tio.Tensor(tuple(dims[: nm.value]), vals, cols)


def same(a, b):
    assert a.dims == b.dims
    np.testing.assert_array_equal(a.vals, b.vals)
    for x, y in zip(a.idx, b.idx):
        np.testing.assert_array_equal(x, y)


def test_tns_matches_fixture(d37):
    T = tio.load_tns(d37)
    ref = tensor_from_npz(golden("verify_3d_7_r16.npz"))
    same(T, ref)
    rc, H = host_load(d37)
    assert rc == 0
    same(H, ref)


def test_synth_cp_r5_both_loaders(tmp_path):
    T = tensor_from_npz(golden("verify_synth_cp_r5_r8.npz"))
    p = tmp_path / "s.tns"
    tio.save_tns(p, T, fmt="%.9g")
    same(tio.load_tns(p), T)
    rc, H = host_load(p)
    assert rc == 0
    same(H, T)


def test_splatt_header_comments_and_blank_lines(tmp_path):
    p = tmp_path / "h.tns"
    p.write_text("# comment\n\n3\n10 20 30\n1 2 3 0.5\n# mid\n2 3 4 1.5\n")
    T = tio.load_tns(p)
    assert T.dims == (10, 20, 30) and T.nnz == 2
    rc, H = host_load(p)
    assert rc == 0
    same(H, T)


@pytest.mark.parametrize("text", ["1 1 1 1.0\n0 2 1 2.0\n", "1 1 1 1.0\n1 2 2.0\n",
                                  "1 1 1 1 1 1 1 1 1 1.0\n", "1 x 1 1.0\n"])
def test_bad_text_is_rejected(tmp_path, text):
    p = tmp_path / "bad.tns"
    p.write_text(text)
    with pytest.raises(tio.TensorFormatError):
        tio.load_tns(p)
    rc, _ = host_load(p)
    assert rc == 2


def test_dnb_roundtrip_and_byte_compat(tmp_path, d37):
    T = tio.load_tns(d37)
    a = tmp_path / "a.dnb"
    tio.save_dnb(a, T)
    same(tio.load_dnb(a), T)
    # the C++ writer produces the same bytes
    lib = host_lib()
    h = C.c_void_p()
    assert lib.dyn_host_tensor_load(str(d37).encode(), 0, C.byref(h)) == 0
    b = tmp_path / "b.dnb"
    assert lib.dyn_host_tensor_save_dnb(h, str(b).encode()) == 0
    lib.dyn_host_tensor_free(h)
    assert a.read_bytes() == b.read_bytes()
    assert len(a.read_bytes()) == 88 + 9 * 4 * 4


@pytest.mark.skipif(not REF.exists(), reason="reference build (oracle/_ref) not present")
def test_dnb_matches_reference_writer(tmp_path, d37):
    """The reference binary writes <path>.dnb on first load; ours must be identical."""
    subprocess.run([str(REF), str(d37), "--rank", "4", "--threads", "1"], cwd=tmp_path,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, check=True)
    ref_bytes = (tmp_path / "3d_7.tns.dnb").read_bytes()
    ours = tmp_path / "ours.dnb"
    tio.save_dnb(ours, tio.load_tns(d37))
    assert ours.read_bytes() == ref_bytes


@pytest.mark.parametrize("mutate,what", [
    (lambda b: b"XNB1" + b[4:], "magic"),
    (lambda b: b[:8] + (8).to_bytes(4, "little") + b[12:], "value_bytes"),
    (lambda b: b[:16] + (9).to_bytes(4, "little") + b[20:], "num_modes"),
    (lambda b: b[:-5], "truncated")])
def test_bad_dnb_is_rejected(tmp_path, d37, mutate, what):
    T = tio.load_tns(d37)
    p = tmp_path / "x.tns.dnb"
    tio.save_dnb(p, T)
    p.write_bytes(mutate(p.read_bytes()))
    with pytest.raises(tio.TensorFormatError):
        tio.load_dnb(p)
    # load_coo_tensor falls back to the text when the cache is invalid (C++
    # loader first: the Python loader rewrites a valid cache after its fallback)
    shutil.copy(d37, tmp_path / "x.tns")
    os.utime(p, None)
    bad = p.read_bytes()
    rc, H = host_load(tmp_path / "x.tns", cache=1)
    assert rc == 0
    same(H, T)
    p.write_bytes(bad)
    os.utime(p, None)
    same(tio.load_coo_tensor(tmp_path / "x.tns"), T)


def test_dnb_cache_is_used_when_fresh(tmp_path, d37):
    T = tio.load_coo_tensor(d37)            # writes 3d_7.tns.dnb
    dnb = Path(str(d37) + ".dnb")
    assert dnb.exists()
    d37.unlink()                             # no source: the cache is trusted
    same(tio.load_coo_tensor(d37), T)
    rc, H = host_load(d37, cache=1)
    assert rc == 0
    same(H, T)


def test_dyfm_roundtrip(tmp_path):
    Y = np.arange(15, dtype=np.float32).reshape(5, 3)
    p = tmp_path / "f.bin"
    tio.save_dyfm(p, Y)
    raw = p.read_bytes()
    assert raw[:4] == b"DYFM" and len(raw) == 24 + 15 * 4
    np.testing.assert_array_equal(tio.load_dyfm(p), Y)


def test_cli_usage_and_errors(tmp_path, d37):
    assert subprocess.run([str(CLI)], capture_output=True).returncode == 1
    assert subprocess.run([str(CLI), str(d37), "--bogus"], capture_output=True).returncode == 1
    # --runtime needs --decompose cpals (exit 7, as the reference's conflicts), and
    # a processed-state file (a .tns is not one: exit 2)
    assert subprocess.run([str(CLI), str(d37), "--runtime"], capture_output=True).returncode == 7
    assert subprocess.run([str(CLI), str(d37), "--runtime", "--decompose", "cpals"],
                          capture_output=True).returncode == 2
    assert subprocess.run([str(CLI), str(tmp_path / "missing.tns")],
                          capture_output=True).returncode == 2
    r = subprocess.run([str(CLI), str(d37), "--decompose", "tucker", "--no-cache"],
                       capture_output=True, text=True)
    assert r.returncode == 5
