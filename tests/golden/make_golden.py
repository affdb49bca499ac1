"""Regenerate the golden fixtures from the reference binary.

Runs oracle/_ref/dynasor_plus (compiled from the reference sources by
`make -C oracle ref`) on small inputs and stores inputs + outputs as .npz
files next to this script.  Only needed when the fixtures change; the tests
read the .npz files and never run the reference.

    python tests/golden/make_golden.py

Fixtures
  verify_<name>_r<R>.npz   --verify runs: tensor, R, seed, Yhat per mode
                           (reference --dump-bin output, DYFM)
  cpals_<name>_<kernel>_r<R>.npz
                           --decompose cpals runs (--threads 1): tensor (or a
                           generator recipe + checksum), init factors
                           (--cp-dump-init), final factors + lambda (--cp-dump),
                           printed fit per iteration (%.6f)
"""
from __future__ import annotations

import hashlib
import re
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

from dynasor_plus_amd import tensor_io as tio  # noqa: E402
from dynasor_plus_amd import host  # noqa: E402

REF_BIN = ROOT / "oracle" / "_ref" / "dynasor_plus"
REF_DATA = Path("/root/reference/data")


def run(args, cwd):
    p = subprocess.run([str(REF_BIN), *map(str, args)], cwd=cwd, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"reference run failed ({p.returncode}): {args}\n{p.stdout[-2000:]}")
    return p.stdout


def tensor_arrays(T):
    return {"dims": np.array(T.dims, dtype=np.uint32), "vals": T.vals,
            **{f"idx{n}": T.idx[n] for n in range(T.num_modes)}}


def checksum(T) -> str:
    h = hashlib.sha256()
    h.update(np.array(T.dims, dtype=np.uint32).tobytes())
    h.update(T.vals.tobytes())
    for c in T.idx:
        h.update(c.tobytes())
    return h.hexdigest()


def verify_fixture(name, tns_path, R, seed, tmp, store_tensor=True):
    out = run([tns_path, "--rank", R, "--threads", 1, "--seed", seed, "--verify", "--no-cache",
               "--dump", "o", "--dump-bin"], tmp)
    if "PASS" not in out:
        raise RuntimeError(out)
    T = tio.load_tns(tns_path)
    yh = {f"yhat{n}": tio.load_dyfm(Path(tmp) / f"o.fm{n}_out.bin") for n in range(T.num_modes)}
    np.savez_compressed(HERE / f"verify_{name}_r{R}.npz", R=R, seed=seed,
                        **(tensor_arrays(T) if store_tensor else {}), **yh)
    print(f"verify_{name}_r{R}: {T.nnz} nnz, modes={T.dims}")


FIT_RE = re.compile(r"^\[cpals\] iter\s+(\d+)\s+fit = ([-0-9.]+)")


def cpals_fixture(name, tns_path, R, kernel, iters, tol, seed, tmp, extra=None, tensor=None):
    out = run([tns_path, "--decompose", "cpals", "--cp-kernel", kernel, "--rank", R, "--threads", 1,
               "--cp-iters", iters, "--cp-tol", tol, "--seed", seed, "--no-cache",
               "--cp-dump", "p", "--cp-dump-init", "p"], tmp)
    fits = [float(m.group(2)) for m in (FIT_RE.match(l) for l in out.splitlines()) if m]
    T = tensor if tensor is not None else tio.load_tns(tns_path)
    init = {f"init{n}": tio.load_dyfm(Path(tmp) / f"p.cp_init_fm{n}.bin") for n in range(T.num_modes)}
    fin = {f"final{n}": tio.load_dyfm(Path(tmp) / f"p.cp_fm{n}.bin") for n in range(T.num_modes)}
    lam = np.fromfile(Path(tmp) / "p.cp_lambda.bin", dtype="<f4")
    arrays = dict(R=R, seed=seed, iters=iters, tol=tol, kernel=kernel, fits=np.array(fits),
                  lam=lam, checksum=checksum(T), **init, **fin)
    if extra:
        arrays.update(extra)
    else:
        arrays.update(tensor_arrays(T))
    np.savez_compressed(HERE / f"cpals_{name}_{kernel}_r{R}.npz", **arrays)
    print(f"cpals_{name}_{kernel}_r{R}: {len(fits)} iterations, last fit {fits[-1] if fits else None}")


def main():
    if not REF_BIN.exists():
        sys.exit("build the reference first: make -C oracle ref")
    with tempfile.TemporaryDirectory() as tmp:
        d37 = REF_DATA / "3d_7.tns"
        scp = REF_DATA / "synth_cp_r5.tns"
        for R in (8, 16, 32, 64):
            verify_fixture("3d_7", d37, R, 42, tmp)
        verify_fixture("synth_cp_r5", scp, 8, 42, tmp)
        # small 4- and 5-mode tensors from the engine's generator (uniform values)
        for name, dims, nnz, R in (("g4", (30, 20, 10, 8), 3000, 16),
                                   ("g5", (12, 10, 9, 7, 5), 2500, 32)):
            T, _, _ = host.generate(dims, nnz, seed=7, coords="uniform", values="uniform")
            p = Path(tmp) / f"{name}.tns"
            tio.save_tns(p, T, fmt="%.9g")
            verify_fixture(name, p, R, 42, tmp)
        # CP-ALS on the exact rank-5 fixture, both kernels
        cpals_fixture("synth_cp_r5", scp, 5, "gauss-seidel", 30, 1e-4, 42, tmp)
        cpals_fixture("synth_cp_r5", scp, 5, "jacobi", 6, 0.0, 42, tmp)
        cpals_fixture("3d_7", d37, 2, "gauss-seidel", 10, 0.0, 42, tmp)
        # configuration 1: 500x400x300, 200k distinct uniform coords, planted rank 10 + 1% noise
        T, v64, _ = host.generate((500, 400, 300), 200_000, seed=1, coords="uniform",
                                  values="planted", planted_rank=10, noise_frac=0.01, want_f64=True)
        p = Path(tmp) / "c1.tns"
        tio.save_tns(p, T, values=v64, fmt="%.17g")
        recipe = dict(gen_dims=np.array(T.dims), gen_nnz=T.nnz, gen_seed=1, gen_noise=0.01,
                      gen_rank=10)
        cpals_fixture("c1", p, 16, "gauss-seidel", 50, 1e-4, 42, tmp, extra=recipe, tensor=T)


if __name__ == "__main__":
    main()
