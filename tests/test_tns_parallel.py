"""The parallel FROSTT parser (the reference's load_tns_mmap,
src/dynasor_io.cpp:960-1176) against the serial one, CPU only.

Both parsers read every value with the same double parse and cast it to
float, so their tensors must be identical bit for bit, including comments
and blank lines between data lines, CRLF endings, a last line without a
newline and a SPLATT header.  Files the parallel parser declines (an index
written as 3.0) go to the serial parser, and malformed files fail either way.
"""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from tests.test_formats import host_load, same

ROOT = Path(__file__).resolve().parents[1]
CLI = ROOT / "dynasor_plus_amd" / "lib" / "dynasor_plus_gpu"


def parse_with(path, parser):
    old = os.environ.get("DYN_TNS_PARSER")
    os.environ["DYN_TNS_PARSER"] = parser
    try:
        return host_load(path, cache=0)
    finally:
        if old is None:
            os.environ.pop("DYN_TNS_PARSER", None)
        else:
            os.environ["DYN_TNS_PARSER"] = old


def loaded_line(path):
    """The CLI's "[io] loaded" line (it stops later for want of a device)."""
    env = dict(os.environ)
    env.pop("DYN_TNS_PARSER", None)
    env["OMP_NUM_THREADS"] = "8"
    p = subprocess.run([str(CLI), str(path), "--no-cache", "--rank", "4"], capture_output=True,
                       text=True, timeout=120, env=env)
    return next((l for l in p.stdout.splitlines() if l.startswith("[io] loaded")), "")


def write_tensor(path, rng, dims, nnz, header=False, crlf=False, noise=False, last_newline=True):
    idx = [rng.integers(1, d + 1, size=nnz) for d in dims]
    vals = rng.standard_normal(nnz) * 10.0 ** rng.integers(-8, 8, size=nnz)
    fmts = ["%.9g", "%.17g", "%.6e", "%+.4f"]
    nl = "\r\n" if crlf else "\n"
    lines = []
    if header:
        lines.append(str(len(dims)))
        lines.append(" ".join(str(d + 3) for d in dims))
    for k in range(nnz):
        if noise and k % 9973 == 0:
            lines.append("# a comment between data lines")
        if noise and k % 7919 == 0:
            lines.append("   ")
        f = fmts[k % len(fmts)]
        sep = "\t" if k % 5 == 0 else " "
        lines.append(sep.join(str(int(i[k])) for i in idx) + sep + (f % vals[k]))
    text = nl.join(lines) + (nl if last_newline else "")
    path.write_bytes(text.encode())


@pytest.mark.parametrize("case", [
    dict(dims=(900, 700, 500), nnz=400_000, noise=True),
    dict(dims=(90, 70, 50, 30), nnz=300_000, crlf=True, last_newline=False),
    dict(dims=(50_000, 20_000), nnz=350_000, header=True, noise=True),
    dict(dims=(12, 10, 8, 6, 5), nnz=2000),                      # below one slice
])
def test_parallel_parser_equals_serial(tmp_path, case):
    rng = np.random.default_rng(len(case["dims"]) * 7 + case["nnz"])
    p = tmp_path / "t.tns"
    write_tensor(p, rng, **case)
    rc1, a = parse_with(p, "parallel")
    rc2, b = parse_with(p, "serial")
    assert rc1 == 0 and rc2 == 0
    same(a, b)
    assert a.nnz == case["nnz"]
    if case.get("header"):
        assert a.dims == tuple(d + 3 for d in case["dims"])


def test_parallel_parser_is_used_and_uses_threads(tmp_path):
    rng = np.random.default_rng(3)
    p = tmp_path / "t.tns"
    write_tensor(p, rng, (900, 700, 500), 400_000)
    line = loaded_line(p)
    assert "parallel parser on" in line, line
    assert int(line.split("parallel parser on")[1].split()[0]) > 1, line


def test_declined_files_go_to_the_serial_parser(tmp_path):
    p = tmp_path / "f.tns"
    p.write_text("1 2 3 1.5\n2.0 1 1 2.5\n4 4 3 -1e-3\n")
    line = loaded_line(p)
    assert line.endswith("(text)"), line
    rc, T = parse_with(p, "parallel")
    assert rc == 0 and T.nnz == 3 and T.dims == (4, 4, 3)
    np.testing.assert_array_equal(T.idx[0], [0, 1, 3])


@pytest.mark.parametrize("text", ["1 2 3 1.5\n1 2 2.5\n", "1 2 3 1.5\n0 1 1 2.0\n",
                                  "1 2 3 abc\n", "# only a comment\n"])
def test_malformed_files_fail_either_way(tmp_path, text):
    p = tmp_path / "bad.tns"
    p.write_text(text)
    for parser in ("parallel", "serial"):
        rc, _ = parse_with(p, parser)
        assert rc != 0, parser
