"""Build the engine's native libraries in-tree.

    python -m dynasor_plus_amd.build            # HIP library + host library + CLI
    python -m dynasor_plus_amd.build --oracle   # also the CPU oracle (test infrastructure)

Outputs (all git-ignored, shipped with the tree):
    dynasor_plus_amd/lib/libdynasor_gpu.so   HIP kernels + C-ABI (include/dynasor_gpu.h)
    dynasor_plus_amd/lib/libdynasor_host.so  C++ drop-in host API (cpals(), loaders, plan)
    dynasor_plus_amd/lib/dynasor_plus_gpu    CLI mirroring `dynasor_plus --decompose cpals`
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
HOST = CSRC / "host"
LIB = PKG / "lib"
BUILD = PKG / "_build"

GPU_ARCH = "gfx950"
HIP_SOURCES = ["dyn_mttkrp.hip", "dyn_mttkrp_big.hip", "dyn_ctx.hip", "dyn_remap.hip", "dyn_algebra.hip",
               "dyn_linear.hip", "dyn_persist.hip", "dyn_stream.hip"]
HOST_SOURCES = ["dyn_io.cpp", "dyn_plan.cpp", "dyn_cpals_host.cpp", "dyn_gen.cpp", "dyn_dnp.cpp"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build the engine)")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}")


def hip_flags() -> list[str]:
    return [
        "-O3", "-std=c++17", "-fPIC", f"--offload-arch={GPU_ARCH}",
        # packed FP32 (v_pk_*_f32) stays off: scalar FP32 code paths only
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument", "-Wno-pass-failed",
        f"-I{ROOT / 'include'}",
    ]


def build_gpu(jobs: int = 8) -> Path:
    hipcc = _hipcc()
    BUILD.mkdir(exist_ok=True)
    LIB.mkdir(exist_ok=True)
    objs = []

    def one(src: str) -> Path:
        obj = BUILD / (src.replace(".hip", ".o"))
        _run([hipcc, *hip_flags(), "-x", "hip", "-c", str(CSRC / src), "-o", str(obj)])
        return obj

    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(HIP_SOURCES)))) as ex:
        objs = list(ex.map(one, HIP_SOURCES))
    out = LIB / "libdynasor_gpu.so"
    tmp = out.with_suffix(".so.tmp")
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={GPU_ARCH}", *map(str, objs), "-o", str(tmp)])
    tmp.replace(out)
    return out


def build_host() -> list[Path]:
    srcs = [HOST / s for s in HOST_SOURCES if (HOST / s).exists()]
    if not srcs:
        return []
    LIB.mkdir(exist_ok=True)
    cxx = os.environ.get("CXX", "g++")
    # -O3 with FMA matches the reference build, whose std::normal_distribution<float>
    # init stream depends on the contraction of its polar-method arithmetic
    flags = ["-O3", "-mavx2", "-mfma", "-std=c++17", "-fPIC", "-fopenmp", "-Wall",
             "-Wno-unused-function",
             f"-I{ROOT / 'include'}", f"-I{HOST}"]
    out = LIB / "libdynasor_host.so"
    tmp = out.with_suffix(".so.tmp")
    rpath = "-Wl,-rpath,$ORIGIN"
    _run([cxx, *flags, "-shared", *map(str, srcs), "-o", str(tmp),
          f"-L{LIB}", "-ldynasor_gpu", rpath])
    tmp.replace(out)
    built = [out]
    main = HOST / "dyn_main.cpp"
    if main.exists():
        exe = LIB / "dynasor_plus_gpu"
        _run([cxx, *flags, str(main), "-o", str(exe), f"-L{LIB}", "-ldynasor_host",
              "-ldynasor_gpu", rpath])
        built.append(exe)
    return built


def build_oracle(with_ref: bool) -> None:
    oracle = ROOT / "oracle"
    if not (oracle / "Makefile").exists():
        return
    _run(["make", "-s", "-C", str(oracle), "-j8"])
    if with_ref and Path("/root/reference/src").is_dir():
        _run(["make", "-s", "-C", str(oracle), "ref", "-j8"])


def build_all(oracle: bool = True) -> None:
    build_gpu()
    build_host()
    if oracle:
        build_oracle(with_ref=True)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--oracle", action="store_true")
    ap.add_argument("--gpu-only", action="store_true")
    args = ap.parse_args()
    build_gpu()
    if not args.gpu_only:
        build_host()
    if args.oracle:
        build_oracle(with_ref=True)
    print("built:", ", ".join(p.name for p in sorted(LIB.iterdir())))


if __name__ == "__main__":
    main()
