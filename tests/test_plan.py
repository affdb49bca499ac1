"""The drop-in C++ boundary on the CPU:

  * a caller written like the reference's `--decompose cpals` block
    (src/main.cpp:770-879) compiles and links against the engine's headers
    and libraries, and plans a tensor without a device;
  * the planner's kernel choice follows the reference's decision tree
    (src/dynasor_preprocess.cpp:1253-1308): a CLI or DYN_CP_KERNEL override,
    else Gauss-Seidel below 1M nonzeros, for the all-modes fallback shapes
    ((N == 3 and Rp >= 64) or Rp >= 128) and when the all-modes accumulators
    exceed min(free / 4, 8 GiB), else Jacobi; `fits` compares the device
    footprint with the free device memory.
"""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HOST = ROOT / "dynasor_plus_amd" / "csrc" / "host"
LIB = ROOT / "dynasor_plus_amd" / "lib"
CPP = ROOT / "tests" / "cpp"
GiB = 1 << 30


def build(name, out_dir):
    if not (LIB / "libdynasor_host.so").exists():
        pytest.skip("host library not built")
    exe = Path(out_dir) / name
    cmd = [shutil.which("g++") or "g++", "-std=c++17", "-O1", f"-I{ROOT / 'include'}", f"-I{HOST}",
           str(CPP / f"{name}.cpp"), "-o", str(exe), f"-L{LIB}", "-ldynasor_host",
           "-ldynasor_gpu", f"-Wl,-rpath,{LIB}"]
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert p.returncode == 0, p.stdout
    return exe


@pytest.fixture(scope="module")
def plan_exe(tmp_path_factory):
    return build("plan_cases", tmp_path_factory.mktemp("plan"))


def plan(exe, nnz, rank, kernel, dims, have=1, free=64 * GiB, total=288 * GiB, env=None):
    e = dict(os.environ)
    e.pop("DYN_CP_KERNEL", None)
    e.pop("DYN_OOC", None)
    e.update(env or {})
    p = subprocess.run([str(exe), str(nnz), str(rank), kernel, str(have), str(free), str(total),
                        *map(str, dims)], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, env=e, timeout=60)
    assert p.returncode == 0, p.stderr
    out = dict(kv.split("=", 1) for kv in p.stdout.strip().split("\t") if "=" in kv)
    return out, p.stderr


def test_drop_in_caller_compiles_links_and_plans(tmp_path):
    exe = build("drop_in_caller", tmp_path)
    t = tmp_path / "t.tns"
    t.write_text("1 1 1 1.0\n2 3 1 2.0\n4 2 3 0.5\n")
    p = subprocess.run([str(exe), str(t), "8", "auto", "--plan-only"], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    assert p.returncode == 0, p.stdout
    assert "[plan] kernel" in p.stdout and "[cpals] kernel -> gauss-seidel" in p.stdout
    assert "choice=5" in p.stdout          # Layout::DeviceCopies


@pytest.mark.parametrize("nnz,rank,dims,kernel,reason", [
    (999_999, 32, (1000, 1000, 1000, 10), 0, "nnz<1M (linalg-dominated)"),
    (2_000_000, 64, (1000, 1000, 1000), 0, "all-modes-fallback-shape"),
    (2_000_000, 100, (1000, 1000, 1000, 10), 0, "all-modes-fallback-shape"),
    (2_000_000, 32, (1000, 1000, 1000), 1, "all-modes fits"),
    (400_000_000, 32, (500_000, 100_000, 50_000, 5_000, 100), 1, "all-modes fits"),
])
def test_auto_kernel_follows_reference_tree(plan_exe, nnz, rank, dims, kernel, reason):
    out, _ = plan(plan_exe, nnz, rank, "auto", dims)
    assert int(out["kernel"]) == kernel and out["reason"] == reason, out
    assert out["source"] == "auto"


def test_auto_kernel_accumulator_budget(plan_exe):
    dims = (1_000_000, 100_000, 10_000, 500)
    acc = sum(dims) * 32 * 4                       # rank-ordered FP32 accumulators
    out, _ = plan(plan_exe, 2_000_000, 32, "auto", dims, free=4 * acc + 4096)
    assert int(out["kernel"]) == 1 and int(out["ofibs"]) == acc
    # (out-of-core off: at this little free memory the copies would not fit either)
    out, _ = plan(plan_exe, 2_000_000, 32, "auto", dims, free=4 * acc - 4096,
                  env={"DYN_OOC": "off"})
    assert int(out["kernel"]) == 0 and out["reason"] == "ofibs>budget"
    # no visible device: nothing to size the pass against
    out, _ = plan(plan_exe, 2_000_000, 32, "auto", dims, have=0)
    assert int(out["kernel"]) == 0 and int(out["fits"]) == 1


@pytest.mark.parametrize("cli,env,kernel,reason,source", [
    ("jacobi", None, 1, "user-forced-jacobi", "cli-kernel"),
    ("all-modes", None, 1, "user-forced-jacobi", "cli-kernel"),
    ("gs", None, 0, "user-forced-gs", "cli-kernel"),
    ("per-mode", None, 0, "user-forced-gs", "cli-kernel"),
    ("gs", "jacobi", 1, "user-forced-jacobi", "env-kernel"),
    ("auto", "gauss-seidel", 0, "user-forced-gs", "env-kernel"),
])
def test_kernel_overrides(plan_exe, cli, env, kernel, reason, source):
    out, _ = plan(plan_exe, 2_000_000, 32, cli, (1000, 1000, 1000),
                  env={"DYN_CP_KERNEL": env} if env else None)
    assert (int(out["kernel"]), out["reason"], out["source"]) == (kernel, reason, source), out


def test_unknown_kernel_warns_and_falls_back_to_auto(plan_exe):
    out, err = plan(plan_exe, 500_000, 32, "simplex", (100, 100, 100))
    assert "unknown kernel 'simplex'" in err
    assert int(out["kernel"]) == 0 and out["source"] == "cli-kernel"


def test_fits_gate(plan_exe):
    dims = (12092, 9184, 28818)
    out, _ = plan(plan_exe, 77_000_000, 32, "auto", dims)
    need = int(out["need"])
    assert int(out["fits"]) == 1 and need > 77_000_000 * 16 * 3
    out, _ = plan(plan_exe, 77_000_000, 32, "auto", dims, free=need - 1)
    assert int(out["fits"]) == 0


@pytest.mark.parametrize("env,cli,kernel,reason,choice", [
    ({"DYN_OOC": "on"}, "auto", 1, "forced-jacobi-by-layout", 3),
    ({"DYN_OOC": "on"}, "gs", 1, "user-wanted-gs-but-layout-forces-jacobi", 3),
    ({"DYN_OOC": "off"}, "auto", 1, "all-modes fits", 5),
])
def test_out_of_core_choice(plan_exe, env, cli, kernel, reason, choice):
    """An out-of-core request streams the tensor from host memory and forces
    Jacobi, as the reference's OutOfCore layout does."""
    out, _ = plan(plan_exe, 2_000_000, 32, cli, (1000, 1000, 1000), env=env)
    assert (int(out["kernel"]), out["reason"], int(out["choice"])) == (kernel, reason, choice), out


def test_memory_tiers_copies_linear_stream(plan_exe):
    """The per-mode copies when they fit; else one linearised copy when that
    fits (Layout::Morton, Jacobi); else streaming from host memory unless
    out-of-core is off, in which case the plan does not fit."""
    dims = (12092, 9184, 28818)
    out, _ = plan(plan_exe, 77_000_000, 32, "auto", dims)
    need, lin = int(out["need"]), int(out["linear"])
    assert int(out["choice"]) == 5 and out["layout"] == "device-remap"
    assert lin < need
    out, _ = plan(plan_exe, 77_000_000, 32, "auto", dims, free=need - 1)
    assert (int(out["choice"]), int(out["kernel"]), out["layout"]) == (4, 1, "device-zorder"), out
    assert out["reason"] == "forced-jacobi-by-layout"
    out, _ = plan(plan_exe, 77_000_000, 32, "auto", dims, free=need - 1, env={"DYN_OOC": "off"})
    assert int(out["choice"]) == 4
    out, _ = plan(plan_exe, 77_000_000, 32, "auto", dims, free=lin - 1)
    assert (int(out["choice"]), int(out["kernel"]), out["layout"]) == (3, 1, "host-stream"), out
    out, _ = plan(plan_exe, 77_000_000, 32, "auto", dims, free=lin - 1, env={"DYN_OOC": "off"})
    assert int(out["choice"]) == 5 and int(out["fits"]) == 0


@pytest.mark.parametrize("layout", ["morton", "linear", "MORTON"])
@pytest.mark.parametrize("cli,reason", [("auto", "forced-jacobi-by-layout"),
                                        ("gs", "user-wanted-gs-but-layout-forces-jacobi"),
                                        ("jacobi", "user-forced-jacobi")])
def test_linear_layout_request(plan_exe, layout, cli, reason):
    """--layout morton selects the linearised copy, which runs Jacobi only, as
    the reference's Morton layout does (src/dynasor_preprocess.cpp:1254-1300)."""
    out, _ = plan(plan_exe, 2_000_000, 32, cli, (1000, 1000, 1000), env={"PLAN_LAYOUT": layout})
    assert (int(out["choice"]), int(out["kernel"]), out["reason"]) == (4, 1, reason), out


def test_cpu_layouts_are_ignored(plan_exe):
    out, _ = plan(plan_exe, 2_000_000, 32, "auto", (1000, 1000, 1000), env={"PLAN_LAYOUT": "ncopy"})
    assert int(out["choice"]) == 5 and int(out["kernel"]) == 1
