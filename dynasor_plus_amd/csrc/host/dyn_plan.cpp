// ComputePlan for the device engine: the reference's planner entry point
// (dyn_build_compute_plan, src/dynasor_preprocess.cpp:1179-1363) with the
// same kernel-selection surface (CLI string, DYN_CP_KERNEL environment
// override, "auto"), sized against device memory instead of host RAM.
#include "dynasor_common.h"
#include "dynasor_gpu.h"
#include "dynasor_jit.h"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <string>

namespace dynasor {

namespace {

std::string lower(std::string s) {
    for (auto& ch : s) ch = (char)std::tolower((unsigned char)ch);
    return s;
}

int device_rank(int R) {
    int p = 16;
    while (p < R) p <<= 1;
    return p;
}

}  // namespace

ComputePlan dyn_build_compute_plan(const Tensor& T, int rank_padded, int num_threads,
                                   const std::string& layout_cli, const std::string& ooc_cli,
                                   const std::string& kernel_cli) {
    uint64_t fr = 0, tot = 0;
    const bool have_dev = dyn_gpu_device_memory(0, &fr, &tot) == DYN_OK;
    // a context cap (DYN_GPU_MEM_CAP_MB) bounds what a run may allocate
    if (const char* cap = std::getenv("DYN_GPU_MEM_CAP_MB")) {
        const long long mb = std::atoll(cap);
        if (mb > 0) fr = std::min<uint64_t>(fr, (uint64_t)mb << 20);
    }
    return dyn_build_compute_plan_with_memory(T, rank_padded, num_threads, layout_cli, ooc_cli,
                                              kernel_cli, have_dev, fr, tot);
}

ComputePlan dyn_build_compute_plan_with_memory(const Tensor& T, int rank_padded, int num_threads,
                                               const std::string& layout_cli,
                                               const std::string& ooc_cli,
                                               const std::string& kernel_cli, bool have_dev,
                                               uint64_t fr, uint64_t tot) {
    ComputePlan p;
    const int N = T.num_modes;
    p.num_modes = N;
    p.rank_padded = rank_padded;
    p.num_threads = num_threads;
    const int Rd = device_rank(rank_padded);
    p.device_rank_padded = Rd;

    // ---- storage: device memory in place of host RAM --------------------
    size_t fb = 0, dfb = 0, acc = 0;
    for (int n = 0; n < N; ++n) {
        fb += 2 * (size_t)T.mode_size[n] * rank_padded * sizeof(value_t);
        dfb += (size_t)T.mode_size[n] * Rd * (4 + 4 + 2 + 4);   // Y, rank-order Y, BF16, Yhat
        acc += (size_t)T.mode_size[n] * Rd * sizeof(float);
    }
    p.factor_bytes = fb;
    p.device_factor_bytes = dfb;
    // COO upload + per mode: vals, N-1 index columns, permutation, row pointer
    size_t tb = (size_t)T.nnz * (4 + 4 * (size_t)N);
    for (int n = 0; n < N; ++n)
        tb += (size_t)T.nnz * (4 + 4 * (size_t)(N - 1) + 4) + ((size_t)T.mode_size[n] + 1) * 4;
    p.device_tensor_bytes = tb;
    const size_t sort_scratch = (size_t)T.nnz * 12;
    p.device_total_bytes = tb + dfb + sort_scratch;
    // linearised layout: the COO upload, one Z-order copy (values, N rank
    // columns, permutation), the rank-order FP64 accumulators and its key
    // sort's scratch (four u32 columns)
    p.device_linear_bytes = (size_t)T.nnz * (4 + 4 * (size_t)N) +
                            (size_t)T.nnz * (4 + 4 * (size_t)N + 4) + 2 * acc + dfb +
                            (size_t)T.nnz * 16;
    if (!have_dev) fr = tot = 0;
    if (have_dev) {
        p.device_free_bytes = fr;
        p.fits = p.device_total_bytes < fr;
    }
    StoragePlan& sp = p.storage;
    sp.choice = Layout::DeviceCopies;
    sp.avail_bytes = fr;
    sp.total_bytes = tot;
    sp.budget_bytes = fr;
    sp.ncopy_bytes = tb;
    sp.overhead_bytes = dfb + sort_scratch;
    sp.ncopy_need = p.device_total_bytes;
    sp.reason = !have_dev ? "device memory unknown"
                : p.fits  ? "per-mode sorted copies fit in device memory"
                          : "per-mode sorted copies exceed free device memory";
    sp.override_source = "auto";
    // out-of-core (src/dynasor_preprocess.cpp:951-1104): the tensor streams
    // from host memory when asked (--ooc on / --layout ooc / DYN_OOC=on), or
    // when its copies do not fit and out-of-core is not switched off.  The CPU
    // layouts (PingPong / NCopy / InPlace / Morton) have no device counterpart.
    const std::string lc = lower(layout_cli);
    std::string oc = lower(ooc_cli);
    const char* osrc = "cli-ooc";
    if (const char* env = std::getenv("DYN_OOC"); env && env[0] && (oc.empty() || oc == "auto")) {
        oc = lower(env);
        osrc = "env-ooc";
    }
    const bool ooc_forced = oc == "on" || lc == "ooc" || lc == "outofcore" || lc == "out-of-core";
    const bool linear_req = lc == "morton" || lc == "linear" || lc == "zorder" || lc == "alto";
    const bool linear_fits = !have_dev || p.device_linear_bytes < fr;
    sp.inplace_bytes = p.device_linear_bytes;
    sp.inplace_need = p.device_linear_bytes;
    if (ooc_forced) {
        sp.choice = Layout::OutOfCore;
        sp.reason = "streamed from host memory (requested)";
        sp.override_source = oc == "on" ? osrc : "cli-layout";
    } else if (linear_req) {
        sp.choice = Layout::Morton;
        sp.reason = "one linearised (Z-order) copy (requested)";
        sp.override_source = "cli-layout";
    } else if (have_dev && !p.fits && linear_fits) {
        // src/dynasor_preprocess.cpp:951-1104 picks the layout the memory
        // allows; on the device one linearised copy sits between the per-mode
        // copies and streaming from host memory
        sp.choice = Layout::Morton;
        sp.reason = "per-mode sorted copies exceed free device memory; one linearised copy fits";
    } else if (have_dev && !p.fits && oc != "off") {
        sp.choice = Layout::OutOfCore;
        sp.reason = "per-mode sorted copies exceed free device memory: streamed from host memory";
    } else if (!lc.empty() && lc != "auto") {
        sp.override_source = "cli-layout (ignored: device copies)";
    }
    // two chunk buffers, one sorted copy and the sort scratch of a chunk
    sp.ooc_need = (size_t)std::min<uint64_t>(T.nnz, 1ull << 27) * (4 + 4 * (size_t)N) * 4 + dfb;
    // out-of-core and the linearised copy run the all-modes update only
    const bool ooc = sp.choice == Layout::OutOfCore || sp.choice == Layout::Morton;
    if (sp.choice == Layout::Morton) p.layout = "device-zorder";
    else if (sp.choice == Layout::OutOfCore) p.layout = "host-stream";

    // ---- kernel: the reference's decision tree ---------------------------
    // (src/dynasor_preprocess.cpp:1253-1308) with the all-modes pass's
    // rank-ordered accumulators in place of the per-thread private outputs
    p.ofibs_bytes = acc;
    const uint64_t cap8 = 8ull << 30;
    const uint64_t budget = std::min<uint64_t>(fr / 4, cap8);
    std::string k = kernel_cli;
    const char* src = "auto";
    if (const char* env = std::getenv("DYN_CP_KERNEL"); env && env[0]) {
        k = env;
        src = "env-kernel";
    } else if (!kernel_cli.empty() && lower(kernel_cli) != "auto") {
        src = "cli-kernel";
    }
    k = lower(k);
    auto pick_auto = [&]() -> int {
        if (ooc) {
            p.kernel_reason = "forced-jacobi-by-layout";
            return 1;
        }
        if (T.nnz < 1000000ull) {
            p.kernel_reason = "nnz<1M (linalg-dominated)";
            return 0;
        }
        if ((N == 3 && Rd >= 64) || Rd >= 128) {
            p.kernel_reason = "all-modes-fallback-shape";
            return 0;
        }
        if (!have_dev || (uint64_t)acc >= budget) {
            p.kernel_reason = "ofibs>budget";
            return 0;
        }
        p.kernel_reason = "all-modes fits";
        return 1;
    };
    if (k == "jacobi" || k == "all-modes" || k == "all_modes" || k == "js") {
        p.kernel_enum = 1;
        p.kernel_reason = "user-forced-jacobi";
    } else if (k == "gauss-seidel" || k == "gs" || k == "classical" || k == "per-mode" ||
               k == "per_mode") {
        p.kernel_enum = 0;
        p.kernel_reason = "user-forced-gs";
        if (ooc) {
            p.kernel_enum = 1;
            p.kernel_reason = "user-wanted-gs-but-layout-forces-jacobi";
        }
    } else {
        if (!k.empty() && k != "auto")
            std::fprintf(stderr,
                         "warning: dyn_build_compute_plan: unknown kernel '%s'; falling back to "
                         "auto.\n",
                         kernel_cli.c_str());
        p.kernel_enum = pick_auto();
    }
    p.kernel_source = src;
    for (int n = 0; n < N && n < DYN_MAX_MODES; ++n) p.per_mode[n] = KernelImpl::FiberCSR;
    p.pf_far = false;
    p.use_jit = false;

    if (const char* env = std::getenv("DYN_CP_PRECISION"); env && env[0]) {
        const std::string e = lower(env);
        p.precision = (e == "bf16" || e == "mixed") ? 1 : 0;
        p.precision_source = "env-precision";
    }
    return p;
}

namespace {
bool g_jit_forced = false;
bool g_jit_on = false;
}  // namespace

bool dyn_jit_enabled() {
    if (g_jit_forced) return g_jit_on;
    const char* e = std::getenv("DYN_JIT");
    return e && e[0] && e[0] != '0';
}

void dyn_jit_force_enable(bool on) {
    g_jit_forced = true;
    g_jit_on = on;
}

void dyn_print_compute_plan(const ComputePlan& p, const char* prefix) {
    const double GiB = 1024.0 * 1024.0 * 1024.0;
    const char* what = p.storage.choice == Layout::Morton      ? "one Z-order copy on the device"
                       : p.storage.choice == Layout::OutOfCore ? "streamed from host memory"
                                                               : "one sorted copy per mode on the device";
    std::printf("%s layout : %s (%s; reason=%s; src=%s)\n", prefix, p.layout, what,
                p.storage.reason, p.storage.override_source);
    std::printf("%s linear : %.2f GiB for one linearised copy\n", prefix,
                p.device_linear_bytes / GiB);
    std::printf("%s device : tensor=%.2f GiB  factors=%.2f GiB  total=%.2f GiB  free=%.2f GiB%s\n",
                prefix, p.device_tensor_bytes / GiB, p.device_factor_bytes / GiB,
                p.device_total_bytes / GiB, p.device_free_bytes / GiB,
                p.fits ? "" : "  (DOES NOT FIT)");
    std::printf("%s kernel : %s  (reason=%s; src=%s)\n", prefix,
                p.kernel_enum ? "Jacobi" : "GaussSeidel", p.kernel_reason, p.kernel_source);
    std::printf("%s gather : %s  (src=%s)  Rp(host)=%d  Rp(device)=%d\n", prefix,
                p.precision ? "bf16" : "fp32", p.precision_source, p.rank_padded,
                p.device_rank_padded);
    std::printf("%s per-mode : FiberCSR (segmented device spMTTKRP) x %d  pf_far=%d  jit=%d  "
                "all-modes accumulators=%.1f MiB\n",
                prefix, p.num_modes, (int)p.pf_far, (int)p.use_jit,
                p.ofibs_bytes / (1024.0 * 1024.0));
}

}  // namespace dynasor
