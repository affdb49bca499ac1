#!/usr/bin/env python3
"""CP-ALS s/iter and spMTTKRP HBM bandwidth on one MI355X.

Workload (default): 3-mode tensor 12092 x 9184 x 28818 with 77M distinct
coordinates drawn Zipf(1.1) per mode, lognormal(0, 1) FP32 values; CP-ALS
Gauss-Seidel at rank 32 with FP64 Cholesky / normalisation / fit.  A step is
one CP-ALS iteration (three per-mode spMTTKRPs + the R x R algebra + the fit).

Prints one JSON line.  `value` = seconds per CP-ALS iteration (plain FP32
gathers), measured over --steps iterations after --warmup iterations, with the
tensor and its per-mode copies already resident in HBM.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import statistics
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    "c2": dict(dims=(12092, 9184, 28818), nnz=77_000_000, coords="zipf", alpha=1.1,
               values="lognormal", rank=32, seed=2,
               desc="3-mode 12092x9184x28818, 77M distinct Zipf(1.1) coords, lognormal(0,1) "
                    "values, CP-ALS Gauss-Seidel rank 32"),
    "small": dict(dims=(1200, 900, 2800), nnz=2_000_000, coords="zipf", alpha=1.1,
                  values="lognormal", rank=32, seed=2,
                  desc="3-mode 1200x900x2800, 2M distinct Zipf(1.1) coords, rank 32"),
    # parity-test configurations, runnable by name (not the default line)
    "c3": dict(dims=(1_000_000, 200_000, 20_000, 500), nnz=150_000_000, coords="zipf", alpha=1.2,
               values="lognormal", rank=64, seed=3,
               desc="4-mode 1Mx200kx20kx500, 150M distinct Zipf(1.2) coords, lognormal(0,1) "
                    "values, CP-ALS Gauss-Seidel rank 64"),
    "c4": dict(dims=(500_000, 100_000, 50_000, 5_000, 100), nnz=400_000_000, coords="zipf",
               alpha=1.0, values="planted", planted_rank=20, noise_frac=0.05, rank=32, seed=4,
               kernel="jacobi",
               desc="5-mode 500kx100kx50kx5kx100, 400M distinct Zipf(1.0) coords, planted rank 20 "
                    "+ 5% noise, CP-ALS Jacobi rank 32"),
}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def algorithmic_bytes(dims, nnz, R, mode, bf=4):
    """Compulsory HBM bytes of one mode-n spMTTKRP (the reference's CSR-compact
    accounting, src/dynasor_kernel.cpp:585-595):
    B_n = nnz (4 + 4 (N-1))            values + the other modes' index columns
        + 8 (I_n + 1)                  row pointer / segment table
        + sum_{w != n} I_w R b_f       every gathered factor row read once
        + 4 I_n R                      output rows written once."""
    N = len(dims)
    gathered = sum(dims[w] * R * bf for w in range(N) if w != mode)
    return nnz * (4 + 4 * (N - 1)) + 8 * (dims[mode] + 1) + gathered + 4 * dims[mode] * R


TRAFFIC_FILE = ROOT / "profiles" / "spmttkrp_c2_hbm_traffic.md"


def measured_traffic(workload):
    """HBM bytes per spMTTKRP launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (calibrated against a plain copy), or None."""
    if workload != "c2" or not TRAFFIC_FILE.exists():
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", TRAFFIC_FILE.read_text(), re.M)
    return float(m.group(1)) if m else None


def copy_calibration(mib=1024, reps=5):
    """A plain device-to-device copy of known size, for calibrating the HBM
    byte counters (FETCH_SIZE / WRITE_SIZE) under rocprofv3."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"copy calibration: {reps} copies of {mib} MiB (read {mib} MiB + write {mib} MiB each)")


def cpu_baseline(T, wl, threads, iters=3):
    """The reference binary (oracle/_ref, compiled from the reference sources)
    on the same tensor: Gauss-Seidel CP-ALS, same rank; s/iter = median of the
    iterations after the first, preprocessing reported separately."""
    ref = ROOT / "oracle" / "_ref" / "dynasor_plus"
    if not ref.exists():
        return None
    from dynasor_plus_amd.tensor_io import save_dnb
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        base = Path(tmp) / "bench.tns"
        save_dnb(str(base) + ".dnb", T)
        t0 = time.time()
        env = dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores")
        p = subprocess.run([str(ref), str(base), "--decompose", "cpals", "--cp-kernel",
                            "gauss-seidel", "--rank", str(wl["rank"]), "--cp-iters", str(iters),
                            "--cp-tol", "0", "--threads", str(threads), "--seed", "42"],
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env,
                           timeout=600)
        wall = time.time() - t0
    for line in p.stdout.splitlines():
        if line.startswith(("[io]", "[layout] NCopy", "[flycoo] preprocess", "[cpals] iter",
                            "[summary]")):
            log("ref>", line)
    if p.returncode != 0:
        log("reference CPU run failed:", p.stdout[-500:])
        return None
    its = [float(m.group(1)) for m in re.finditer(r"\[cpals\] iter\s+\d+ .*iter = ([0-9.]+) s",
                                                   p.stdout)]
    if not its:
        return None
    steady = its[1:] if len(its) > 1 else its
    v = statistics.median(steady)
    return {"value": v, "unit": "s/iter", "cores": threads, "kind": "reference",
            "sample": f"full workload, {iters} Gauss-Seidel iterations at rank {wl['rank']} on "
                      f"{threads} host threads; median of iterations 2..{iters}; "
                      f"whole run incl. load+preprocess {wall:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the BF16-gather comparison")
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
              "(the engine is single-device; multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return

    from dynasor_plus_amd import GpuContext, host, init_factors
    from dynasor_plus_amd import _lib as L

    wl = WORKLOADS[args.workload]
    threads = args.threads or min(16, os.cpu_count() or 1)
    t0 = time.time()
    gen_extra = {k: wl[k] for k in ("planted_rank", "noise_frac") if k in wl}
    T, _, st = host.generate(wl["dims"], wl["nnz"], seed=wl["seed"], coords=wl["coords"],
                             alpha=wl["alpha"], values=wl["values"], threads=threads, **gen_extra)
    log(f"generated {T.nnz} nnz in {time.time() - t0:.1f}s (rounds={st.rounds}, "
        f"max row={st.max_row_nnz_mode0})")
    R = wl["rank"]
    dims = T.dims
    ctx = GpuContext(0)
    ctx.upload(T)
    t0 = time.time()
    ctx.remap_all()
    remap_s = time.time() - t0
    log(f"remap ({len(dims)} per-mode copies) {remap_s:.3f}s")
    Y0 = init_factors(dims, R, 42, "normal")
    kern = L.DYN_CP_JACOBI if wl.get("kernel") == "jacobi" else L.DYN_CP_GAUSS_SEIDEL

    # ---- plain FP32 gathers ----
    ctx.set_factors(Y0, R)
    ctx.run_cpals(args.warmup, 0.0, kernel=kern)
    ctx.enable_timing(True)
    ctx.kernel_stats()                      # reset
    ctx.synchronize()
    t0 = time.perf_counter()
    res = ctx.run_cpals(args.steps, 0.0, kernel=kern)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    launches, kms = ctx.kernel_stats()
    ctx.enable_timing(False)
    s_per_iter = elapsed / args.steps
    N = len(dims)
    bytes_sweep = sum(algorithmic_bytes(dims, T.nnz, R, n) for n in range(N))
    sweeps = launches / N
    achieved = bytes_sweep * sweeps / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    log(f"plain: {s_per_iter * 1e3:.3f} ms/iter, spMTTKRP {launches} launches {kms:.3f} ms "
        f"-> {achieved:.1f} GB/s, fit {res.final_fit:.6f}")

    # ---- mixed: BF16 gathered operand, same init, same iteration count ----
    mixed = None
    if not args.no_mixed:
        iters = args.warmup + args.steps
        ctx.set_factors(Y0, R)
        rp = ctx.run_cpals(iters, 0.0, kernel=kern, precision=L.DYN_PREC_FP32)
        ctx.set_factors(Y0, R)
        ctx.enable_timing(True)
        ctx.kernel_stats()
        rm = ctx.run_cpals(iters, 0.0, kernel=kern, precision=L.DYN_PREC_BF16)
        ml, mms = ctx.kernel_stats()
        ctx.enable_timing(False)
        med = lambda h: statistics.median(h[1:]) if len(h) > 1 else h[0]
        bytes_bf = sum(algorithmic_bytes(dims, T.nnz, R, n, bf=2) for n in range(N))
        mixed = {"s_per_iter": med(rm.iter_time_s), "plain_s_per_iter": med(rp.iter_time_s),
                 "speedup": med(rp.iter_time_s) / med(rm.iter_time_s),
                 "final_fit": rm.final_fit, "plain_final_fit": rp.final_fit,
                 "fit_abs_diff": abs(rm.final_fit - rp.final_fit), "iters": iters,
                 "spmttkrp_GBps": bytes_bf * (ml / N) / (mms * 1e-3) / 1e9 if mms > 0 else 0.0}
        log(f"mixed: {mixed}")
    ctx.close()

    cpu = None
    if not args.no_cpu and args.workload in ("c2", "small"):
        t0 = time.time()
        cpu = cpu_baseline(T, wl, threads)
        log(f"cpu baseline {cpu} ({time.time() - t0:.1f}s)")

    out = {
        "metric": "CP-ALS s/iter + spMTTKRP HBM GB/s at equal fit, mixed vs plain GPU, 1 GPU",
        "value": s_per_iter,
        "unit": "s/iter",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": s_per_iter * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (gathers/accumulate), f64 (Cholesky, lambda, fit)",
        "data": "synthetic (engine generator, distinct coordinates)",
        "config": {"workload": wl["desc"], "dims": list(dims), "nnz": T.nnz, "rank": R,
                   "kernel": wl.get("kernel", "gauss-seidel"), "parallelism": "single GPU"},
        "spmttkrp_GBps": achieved,
        "spmttkrp_ms_per_mode": kms / launches if launches else None,
        "remap_s": remap_s,
        "final_fit": res.final_fit,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(args.workload)},
        "mixed": mixed,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
