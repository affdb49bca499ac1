#!/usr/bin/env python3
"""CP-ALS s/iter and spMTTKRP HBM bandwidth on one MI355X, mixed vs plain.

Headline workload (default, the largest single-GPU configuration and the
mixed-precision all-modes one): a 5-mode 500k x 100k x 50k x 5k x 100 tensor
with 400M distinct Zipf(1.0) coordinates and values from a planted rank-20
model plus 5 % Gaussian noise; CP-ALS with the all-modes (Jacobi) update at
rank 32, gathered factor operands in BF16, FP32 accumulation, FP64 Cholesky /
lambda / fit.  A step is one CP-ALS iteration (every mode's spMTTKRP from one
factor snapshot, the five R x R solves and normalisations, and the fit).

Prints one JSON line.  `value` = seconds per iteration of the mixed (BF16
gathers) run over --steps iterations after --warmup iterations, with the
tensor and its per-mode copies resident in HBM; `mixed` holds the plain FP32
run on the same initial factors and whether the two final fits agree within
1e-4.  `extra` holds the same comparison for the 3-mode (rank 32) and
4-mode (rank 64) Gauss-Seidel configurations.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import re
import shutil
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
FP32_MFMA_PEAK_TFLOPS = 157.3
EQUAL_FIT_TOL = 1e-4
WORKLOADS = {
    "c4": dict(dims=(500_000, 100_000, 50_000, 5_000, 100), nnz=400_000_000, coords="zipf",
               alpha=1.0, values="planted", planted_rank=20, noise_frac=0.05, rank=32, seed=4,
               kernel="jacobi", steps=30, cpu_sample=8,
               desc="5-mode 500kx100kx50kx5kx100, 400M distinct Zipf(1.0) coords, planted rank 20 "
                    "+ 5% noise, CP-ALS all-modes (Jacobi) rank 32, BF16 gathered operands"),
    "c2": dict(dims=(12092, 9184, 28818), nnz=77_000_000, coords="zipf", alpha=1.1,
               values="lognormal", rank=32, seed=2, kernel="gauss-seidel", steps=20, cpu_sample=1,
               desc="3-mode 12092x9184x28818, 77M distinct Zipf(1.1) coords, lognormal(0,1) "
                    "values, CP-ALS Gauss-Seidel rank 32"),
    "c3": dict(dims=(1_000_000, 200_000, 20_000, 500), nnz=150_000_000, coords="zipf", alpha=1.2,
               values="lognormal", rank=64, seed=3, kernel="gauss-seidel", steps=20, cpu_sample=2,
               desc="4-mode 1Mx200kx20kx500, 150M distinct Zipf(1.2) coords, lognormal(0,1) "
                    "values, CP-ALS Gauss-Seidel rank 64"),
    "c5": dict(dims=(1_000_000, 100_000, 2_000), nnz=200_000_000, coords="uniform",
               values="bernoulli", planted_rank=16, planted_factors="normal", planted_var=0.5,
               rank=32, seed=5, kernel="completion", steps=20, cpu_sample=0,
               desc="3-mode 1Mx100kx2k, 200M uniform coords, Bernoulli labels of a planted "
                    "rank-16 N(0,0.5) model; Adam on the fused Bernoulli-logit loss/adjoint, "
                    "rank 32"),
    "small": dict(dims=(1200, 900, 2800), nnz=2_000_000, coords="zipf", alpha=1.1,
                  values="lognormal", rank=32, seed=2, kernel="gauss-seidel", steps=20,
                  cpu_sample=1, desc="3-mode 1200x900x2800, 2M distinct Zipf(1.1) coords, rank 32"),
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


def jacobi_pass_bytes(dims, nnz, R, bf=4):
    """Compulsory bytes of one all-modes pass over a single copy:
    B_J = nnz (4 + 4 N) + sum_n I_n R (b_f + 4)."""
    return nnz * (4 + 4 * len(dims)) + sum(d * R * (bf + 4) for d in dims)


def profile_figure(name, key):
    """A figure from a committed rocprofv3 summary under profiles/, or None."""
    f = ROOT / "profiles" / name
    if not f.exists():
        return None
    m = re.search(rf"^{key}:\s*([0-9.eE+-]+)", f.read_text(), re.M)
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


def host_threads():
    """Host threads for the CPU baseline: every CPU this process may run on,
    capped by OMP_NUM_THREADS when set (the GPU box allots a share of its
    cores per GPU and exports that share there)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def reference_binary():
    """oracle/_ref build of the reference sources: the AVX-512 one where the
    host supports x86-64-v4, else the AVX2 one."""
    ref = ROOT / "oracle" / "_ref"
    flags = set()
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("flags"):
                flags = set(line.split(":", 1)[1].split())
                break
    except OSError:
        pass
    v4 = {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"} <= flags
    for name, ok in (("dynasor_plus_avx512", v4), ("dynasor_plus", True)):
        p = ref / name
        if ok and p.exists():
            return p, ("avx512 (x86-64-v4)" if name.endswith("avx512") else "avx2+fma")
    return None, None


def cpu_baseline(T, wl, threads, iters=3):
    """The reference binary (compiled from the reference sources) on the same
    workload: its CP-ALS with the same kernel and rank.  For large tensors it
    runs on every k-th nonzero (same mode sizes), and the per-iteration time is
    scaled by k (work is linear in nonzeros); preprocessing is not counted."""
    ref, isa = reference_binary()
    k = int(wl.get("cpu_sample", 1))
    if ref is None or k <= 0:
        return None
    from dynasor_plus_amd.tensor_io import Tensor, save_dnb
    S = T if k == 1 else Tensor(T.dims, T.vals[::k].copy(), [c[::k].copy() for c in T.idx])
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        base = Path(tmp) / "bench.tns"
        save_dnb(str(base) + ".dnb", S)
        t0 = time.time()
        env = dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores",
                   OMP_NUM_THREADS=str(threads))
        p = subprocess.run([str(ref), str(base), "--decompose", "cpals", "--cp-kernel",
                            wl["kernel"], "--rank", str(wl["rank"]), "--cp-iters", str(iters),
                            "--cp-tol", "0", "--threads", str(threads), "--seed", "42"],
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env,
                           timeout=900)
        wall = time.time() - t0
    for line in p.stdout.splitlines():
        if line.startswith(("[io]", "[flycoo] preprocess", "[cpals] iter", "[cpals] kernel",
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
    v = statistics.median(steady) * k
    sample = (f"{'full workload' if k == 1 else f'every {k}th nonzero ({S.nnz} of {T.nnz}), same mode sizes'}"
              f"; {iters} {wl['kernel']} iterations at rank {wl['rank']}; median of iterations "
              f"2..{iters}{'' if k == 1 else f', times {k}'}; reference build {isa}; "
              f"whole run incl. load+preprocess {wall:.1f} s")
    return {"value": v, "unit": "s/iter", "cores": threads, "kind": "reference", "sample": sample}


def timed_run(ctx, Y0, R, kern, prec, warmup, steps):
    """Warmup iterations, then exactly `steps` timed iterations from Y0, with
    the spMTTKRP launches timed by HIP events on their own stream."""
    ctx.set_factors(Y0, R)
    if warmup:
        ctx.run_cpals(warmup, 0.0, kernel=kern, precision=prec)
        ctx.set_factors(Y0, R)
    ctx.enable_timing(True)
    ctx.kernel_stats()
    ctx.synchronize()
    t0 = time.perf_counter()
    res = ctx.run_cpals(steps, 0.0, kernel=kern, precision=prec)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    launches, kms = ctx.kernel_stats()
    ctx.enable_timing(False)
    return res, elapsed, launches, kms


def kernel_figures(dims, nnz, R, jacobi, steps, launches, kms, bf, chunk_nnz=0):
    """Algorithmic bytes and flops of the timed spMTTKRP launches.  A
    Gauss-Seidel sweep launches every mode once; a Jacobi run of K iterations
    launches every mode K times plus the smallest mode once more for the last
    iteration's fit.  With BF16 gathers the mode whose Yhat carries the fit
    (Gauss-Seidel: the last, Jacobi: the smallest) gathers FP32 rows.  A
    streamed tensor (chunk_nnz > 0) launches every mode once per chunk; each
    chunk's launch is credited with its own stream bytes plus the row
    pointer, the gathered rows and the output once."""
    N = len(dims)
    fit_mode = int(np.argmin(dims)) if jacobi else N - 1
    chunks = [nnz] if not chunk_nnz else \
        [min(chunk_nnz, nnz - b) for b in range(0, nnz, chunk_nnz)]
    per_mode = [sum(algorithmic_bytes(dims, c, R, n, 4 if n == fit_mode else bf) for c in chunks)
                for n in range(N)]
    total = steps * sum(per_mode)
    n_expected = steps * N * len(chunks)
    if jacobi:
        total += per_mode[fit_mode]
        n_expected += len(chunks)
    if launches != n_expected or kms <= 0:
        log(f"unexpected launch count {launches} (expected {n_expected})")
    flops = N * R * nnz * (n_expected // len(chunks))
    gbs = total / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    return {"GBps": gbs, "GFLOPs": flops / (kms * 1e-3) / 1e9 if kms > 0 else 0.0,
            "ms_per_launch": kms / launches if launches else None, "launches": launches,
            "bytes_per_launch": total / launches if launches else None}


def linear_figures(dims, nnz, R, steps, launches, kms):
    """Algorithmic bytes and flops of the timed all-modes passes over the
    linearised copy: each pass streams the copy once and reads and writes
    every computed mode's rows once (B_J, FP32 rows); a Jacobi run of K
    iterations makes K full passes plus one pass restricted to the smallest
    mode for the last fit (its stream and every gathered mode once, one
    output)."""
    N = len(dims)
    s = int(np.argmin(dims))
    full = jacobi_pass_bytes(dims, nnz, R, 4)
    last = nnz * (4 + 4 * N) + sum(d * R * 4 for d in dims) + dims[s] * R * 4
    total = steps * full + last
    if launches != steps + 1 or kms <= 0:
        log(f"unexpected linear launch count {launches} (expected {steps + 1})")
    flops = N * N * R * nnz * steps + N * R * nnz
    return {"GBps": total / (kms * 1e-3) / 1e9 if kms > 0 else 0.0,
            "GFLOPs": flops / (kms * 1e-3) / 1e9 if kms > 0 else 0.0,
            "ms_per_launch": kms / launches if launches else None, "launches": launches,
            "bytes_per_launch": total / launches if launches else None}


def cpals_leg(name, wl, warmup, steps, threads, want_cpu, mixed=True, gs_too=False,
              ooc_chunk=None, ooc_source="host", layout="copies", linear_too=False):
    """Plain FP32 and mixed BF16-gather CP-ALS on one workload: s/iter, fits,
    equal-fit check and the spMTTKRP roofline figures of each.  `gs_too`
    adds the same comparison under Gauss-Seidel on the same copies (for a
    Jacobi workload); `ooc_chunk` streams the tensor from host memory in
    chunks of that many nonzeros (0: the engine default) instead of keeping
    it resident; layout "linear" runs Jacobi on the single Z-order copy (FP32
    gathers only); `linear_too` adds that run after the others, on the same
    tensor and initial factors."""
    from dynasor_plus_amd import GpuContext, host, init_factors
    from dynasor_plus_amd import _lib as L
    t0 = time.time()
    gen_extra = {k: wl[k] for k in ("planted_rank", "noise_frac") if k in wl}
    T, _, st = host.generate(wl["dims"], wl["nnz"], seed=wl["seed"], coords=wl["coords"],
                             alpha=wl.get("alpha", 1.0), values=wl["values"], threads=threads,
                             **gen_extra)
    log(f"{name}: generated {T.nnz} nnz in {time.time() - t0:.1f}s (max row={st.max_row_nnz_mode0})")
    R, dims, N = wl["rank"], T.dims, T.num_modes
    jacobi = wl["kernel"] == "jacobi"
    kern = L.DYN_CP_JACOBI if jacobi else L.DYN_CP_GAUSS_SEIDEL
    Y0 = init_factors(dims, R, 42, "normal")
    out = {"workload": wl["desc"], "dims": list(dims), "nnz": T.nnz, "rank": R,
           "kernel": wl["kernel"], "steps": steps, "warmup": warmup}
    legs = (("mixed", L.DYN_PREC_BF16, 2), ("plain", L.DYN_PREC_FP32, 4))

    def run_pair(ctx, kern_id, is_jacobi, tag, mixed=mixed):
        runs = {}
        for label, prec, bf in (legs if mixed else legs[1:]):
            res, elapsed, launches, kms = timed_run(ctx, Y0, R, kern_id, prec, warmup, steps)
            if layout == "linear":
                k = linear_figures(dims, T.nnz, R, steps, launches, kms)
            else:
                k = kernel_figures(dims, T.nnz, R, is_jacobi, steps, launches, kms, bf,
                                   chunk_nnz=0 if ooc_chunk is None else
                                   (ooc_chunk or min(T.nnz, 1 << 27)))
            k["s_per_iter"] = elapsed / steps
            k["final_fit"] = res.final_fit
            k["fits"] = [round(f, 7) if math.isfinite(f) else None for f in res.fit_history]
            if is_jacobi and ooc_chunk is None and layout != "linear":
                # all-modes pass over one copy as the yardstick
                bj = jacobi_pass_bytes(dims, T.nnz, R, bf)
                k["jacobi_pass_frac"] = bj / (kms * 1e-3 / (launches / N)) / 1e9 / HBM_PEAK_GBS
            if ooc_chunk is not None:
                # host -> device bytes per sweep over the whole iteration time
                k["stream_GBps"] = T.nnz * (4 + 4 * N) / k["s_per_iter"] / 1e9
            runs[label] = k
            log(f"{name}{tag} {label}: {k['s_per_iter'] * 1e3:.3f} ms/iter, spMTTKRP "
                f"{k['ms_per_launch']:.3f} ms/launch -> {k['GBps']:.0f} GB/s, "
                f"{k['GFLOPs']:.0f} GFLOP/s, final fit {res.final_fit}")
        if not mixed:
            runs["mixed"] = runs["plain"]
        fm, fp = runs["mixed"]["final_fit"], runs["plain"]["final_fit"]
        equal = math.isfinite(fm) and math.isfinite(fp) and abs(fm - fp) <= EQUAL_FIT_TOL
        runs["equal_fit"] = equal
        runs["fit_abs_diff"] = abs(fm - fp) if math.isfinite(fm) and math.isfinite(fp) else None
        runs["speedup"] = runs["plain"]["s_per_iter"] / runs["mixed"]["s_per_iter"]
        if not equal:
            log(f"{name}{tag}: final fits differ (mixed {fm}, plain {fp}): the speedup is not "
                "at equal fit")
        return runs

    with GpuContext(0) as ctx:
        t0 = time.time()
        if layout == "linear":
            ctx.upload(T)
            ctx.set_layout(L.DYN_LAYOUT_LINEAR)
            out["linear_build_s"] = time.time() - t0
            out["layout"] = "linear (one Z-order copy, single-pass all-modes)"
        elif ooc_chunk is None:
            ctx.upload(T)
            ctx.remap_all()
            out["remap_s"] = time.time() - t0
        elif ooc_source == "dnb":
            from dynasor_plus_amd.tensor_io import save_dnb
            tmpdir = tempfile.mkdtemp(prefix="dynooc_")
            path = os.path.join(tmpdir, "t.dnb")
            save_dnb(path, T)
            log(f"{name}: wrote {os.path.getsize(path) / 1e9:.2f} GB to {path} in "
                f"{time.time() - t0:.1f}s")
            t0 = time.time()
            ctx.attach_dnb(path, ooc_chunk)
            out["attach_s"] = time.time() - t0
            out["source"] = "dnb"
            out["chunk_nnz"] = ooc_chunk or min(T.nnz, 1 << 27)
        else:
            ctx.attach_host(T, ooc_chunk)
            out["attach_s"] = time.time() - t0
            out["source"] = "host"
            out["chunk_nnz"] = ooc_chunk or min(T.nnz, 1 << 27)
        try:
            out.update(run_pair(ctx, kern, jacobi, "", mixed=mixed and layout != "linear"))
            if gs_too and jacobi:
                out["gauss_seidel"] = run_pair(ctx, L.DYN_CP_GAUSS_SEIDEL, False, " gauss-seidel")
            if linear_too and ooc_chunk is None and layout != "linear":
                t0 = time.time()
                ctx.set_layout(L.DYN_LAYOUT_LINEAR)
                build_s = time.time() - t0
                layout = "linear"
                lin = run_pair(ctx, L.DYN_CP_JACOBI, True, " linear", mixed=False)
                out["linear"] = {"build_s": build_s, **lin["plain"]}
        finally:
            if ooc_source == "dnb" and ooc_chunk is not None:
                # the context keeps the file open; the unlinked file stays readable
                shutil.rmtree(tmpdir, ignore_errors=True)
    if want_cpu:
        t0 = time.time()
        out["cpu_baseline"] = cpu_baseline(T, wl, threads)
        log(f"{name}: cpu baseline {out['cpu_baseline']} ({time.time() - t0:.1f}s)")
    return out


def completion_leg(wl, warmup, steps, threads):
    """Sparse CP completion: Adam steps on the fused Bernoulli-logit
    loss + adjoint kernel (every mode's gradient per step)."""
    import torch
    from dynasor_plus_amd import GpuContext, host
    from dynasor_plus_amd.autograd import completion_loss_and_grad
    T, _, _ = host.generate(wl["dims"], wl["nnz"], seed=wl["seed"], coords="uniform",
                            values="bernoulli", planted_rank=wl["planted_rank"],
                            planted_factors="normal", planted_var=wl["planted_var"],
                            threads=threads)
    g = torch.Generator(device="cpu").manual_seed(0)
    A = [(0.1 * torch.randn(int(d), wl["rank"], generator=g)).cuda().requires_grad_(True)
         for d in T.dims]
    opt = torch.optim.Adam(A, lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    losses = []

    def step():
        val, grads = completion_loss_and_grad(eng, A, "bernoulli")
        for a, gr in zip(A, grads):
            a.grad = gr.contiguous()
        opt.step()
        losses.append(val)

    with GpuContext(0) as eng:
        eng.upload(T)
        eng.remap_all()
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
    return {"workload": wl["desc"], "s_per_step": el / steps, "loss_first": float(losses[0]),
            "loss_last": float(losses[-1]), "steps": steps, "warmup": warmup}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed iterations (default per workload)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c4", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the 3- and 4-mode Gauss-Seidel comparison legs")
    ap.add_argument("--no-mixed", action="store_true",
                    help="plain FP32 gathers only (the line then reports that run)")
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--ooc", type=int, default=None, metavar="CHUNK_NNZ",
                    help="stream the tensor from host memory in chunks of CHUNK_NNZ nonzeros "
                         "(0: the engine default) instead of keeping it in device memory")
    ap.add_argument("--ooc-source", default="host", choices=("host", "dnb"),
                    help="with --ooc: stream from the host arrays, or from a .dnb file written "
                         "to a temporary directory first")
    ap.add_argument("--layout", default="copies", choices=("copies", "linear"),
                    help="copies: one sorted copy per mode (default); linear: one Z-order "
                         "copy and the single-pass all-modes kernel (Jacobi, FP32 gathers)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
              "(the engine is single-device; sharding over GPUs is design-only)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    wl = WORKLOADS[args.workload]
    threads = args.threads or host_threads()
    steps = args.steps or wl["steps"]
    if wl["kernel"] == "completion":
        print(json.dumps({"workload": args.workload, **completion_leg(wl, args.warmup, steps,
                                                                      threads)}), flush=True)
        return

    linear = args.layout == "linear"
    if args.ooc is not None or linear:
        # a streamed or linearised tensor runs the all-modes update
        wl = dict(wl, kernel="jacobi")
    plain_head = args.workload == "c4" and not args.no_extra and args.ooc is None and not linear
    head = cpals_leg(args.workload, wl, args.warmup, steps, threads,
                     not args.no_cpu and args.ooc is None, mixed=not args.no_mixed and not linear,
                     gs_too=plain_head, ooc_chunk=args.ooc, ooc_source=args.ooc_source,
                     layout=args.layout, linear_too=plain_head)
    keys = ("s_per_iter", "final_fit", "GBps", "GFLOPs", "ms_per_launch")

    def summary(leg):
        d = {k: leg[k] for k in ("equal_fit", "fit_abs_diff", "speedup")}
        for label in ("mixed", "plain"):
            d[label] = {k: leg[label][k] for k in keys + ("stream_GBps",) if k in leg[label]}
        return d

    extra = {}
    if args.workload == "c4" and not args.no_extra and args.ooc is None and not linear:
        for name in ("c2", "c3"):
            w = WORKLOADS[name]
            leg = cpals_leg(name, w, args.warmup, w["steps"], threads, not args.no_cpu)
            extra[name] = {"workload": leg["workload"], **summary(leg),
                           "cpu_baseline": leg.get("cpu_baseline")}
        # the 3-mode tensor streamed from host memory in 5 chunks (Jacobi)
        w = dict(WORKLOADS["c2"], kernel="jacobi",
                 desc=WORKLOADS["c2"]["desc"].replace("Gauss-Seidel", "all-modes (Jacobi)") +
                 ", streamed from host memory in 16M-nonzero chunks")
        leg = cpals_leg("c2-ooc", w, 1, 5, threads, False, ooc_chunk=1 << 24)
        extra["c2_streamed"] = {"workload": leg["workload"], "chunk_nnz": leg["chunk_nnz"],
                                "attach_s": leg["attach_s"], "steps": 5, **summary(leg)}

    m = head["mixed"]
    tr_file = f"spmttkrp_{args.workload}_bf16_hbm_traffic.md"
    gram_util = profile_figure("gram_mfma_counters.md", "mfma_busy_frac")
    out = {
        "metric": "CP-ALS s/iter + spMTTKRP HBM GB/s at equal fit, mixed vs plain GPU, 1 GPU",
        "value": m["s_per_iter"],
        "unit": "s/iter",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": m["s_per_iter"] * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f32 gathered operands" if args.no_mixed or linear else "bf16 gathered operands") +
                 ", f32 accumulate, f64 Cholesky/lambda/fit",
        "data": "synthetic (engine generator, distinct coordinates); N(0,1) initial factors, seed 42",
        "config": {"workload": head["workload"], "dims": head["dims"], "nnz": head["nnz"],
                   "rank": head["rank"], "kernel": head["kernel"], "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": m["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": m["GBps"] / HBM_PEAK_GBS,
                     "traffic": None if args.ooc is not None or linear else
                     profile_figure(tr_file, "traffic_bytes_per_launch")},
        "spmttkrp_GFLOPs": m["GFLOPs"],
        "spmttkrp_ms_per_launch": m["ms_per_launch"],
        "gram_mfma_busy_frac": gram_util,
        "remap_s": head.get("remap_s"),
        "final_fit": m["final_fit"],
        "mixed": {"s_per_iter": m["s_per_iter"], "plain_s_per_iter": head["plain"]["s_per_iter"],
                  "speedup": head["speedup"], "equal_fit": head["equal_fit"],
                  "final_fit": m["final_fit"], "plain_final_fit": head["plain"]["final_fit"],
                  "fit_abs_diff": head["fit_abs_diff"], "fits": m["fits"],
                  "plain_fits": head["plain"]["fits"],
                  "spmttkrp_GBps": m["GBps"], "plain_spmttkrp_GBps": head["plain"]["GBps"],
                  "jacobi_pass_frac": m.get("jacobi_pass_frac"),
                  "plain_jacobi_pass_frac": head["plain"].get("jacobi_pass_frac")},
        "gauss_seidel": summary(head["gauss_seidel"]) if "gauss_seidel" in head else None,
        "linear": ({k: head["linear"][k] for k in ("build_s",) + keys + ("fits",)}
                   if "linear" in head else None),
        "layout": head.get("layout", "copies (one sorted copy per mode)"),
        "streamed": ({"chunk_nnz": head["chunk_nnz"], "attach_s": head["attach_s"],
                      "source": head["source"], "stream_GBps": m.get("stream_GBps")}
                     if args.ooc is not None else None),
        "extra": extra or None,
        "cpu_baseline": head.get("cpu_baseline"),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
