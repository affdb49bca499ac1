// ComputePlan for the device engine: the reference's planner entry point
// (dyn_build_compute_plan, src/dynasor_preprocess.cpp:1179-1363) with the
// same kernel-selection surface (CLI string, DYN_CP_KERNEL environment
// override, "auto"), sized against device memory instead of host RAM.
#include "dynasor_common.h"
#include "dynasor_gpu.h"

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
    (void)layout_cli;   // CPU layouts (PingPong / NCopy / InPlace / Morton) do not apply:
    (void)ooc_cli;      // the device always builds one sorted copy per mode
    ComputePlan p;
    const int N = T.num_modes;
    p.num_modes = N;
    p.rank_padded = rank_padded;
    p.num_threads = num_threads;
    const int Rd = device_rank(rank_padded);
    p.device_rank_padded = Rd;

    size_t fb = 0, dfb = 0;
    for (int n = 0; n < N; ++n) {
        fb += 2 * (size_t)T.mode_size[n] * rank_padded * sizeof(value_t);
        dfb += (size_t)T.mode_size[n] * Rd * (4 + 2 + 4);    // Y, BF16 shadow, Yhat
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
    uint64_t fr = 0, tot = 0;
    if (dyn_gpu_device_memory(0, &fr, &tot) == DYN_OK) {
        p.device_free_bytes = fr;
        p.fits = p.device_total_bytes < fr;
    }

    std::string k = kernel_cli;
    const char* src = "auto";
    if (const char* env = std::getenv("DYN_CP_KERNEL"); env && env[0]) {
        k = env;
        src = "env-kernel";
    } else if (!kernel_cli.empty() && lower(kernel_cli) != "auto") {
        src = "cli-kernel";
    }
    k = lower(k);
    if (k == "jacobi" || k == "all-modes" || k == "all_modes" || k == "js") {
        p.kernel_enum = 1;
        p.kernel_reason = "user-forced-jacobi";
    } else if (k == "gauss-seidel" || k == "gs" || k == "classical" || k == "per-mode" ||
               k == "per_mode") {
        p.kernel_enum = 0;
        p.kernel_reason = "user-forced-gs";
    } else {
        if (!k.empty() && k != "auto")
            std::fprintf(stderr,
                         "warning: dyn_build_compute_plan: unknown kernel '%s'; falling back to "
                         "auto.\n",
                         kernel_cli.c_str());
        // every per-mode copy is resident in HBM, so a Gauss-Seidel sweep costs
        // the same tensor traffic as a Jacobi pass and converges faster
        p.kernel_enum = 0;
        p.kernel_reason = "per-mode copies resident in HBM";
    }
    p.kernel_source = src;

    if (const char* env = std::getenv("DYN_CP_PRECISION"); env && env[0]) {
        const std::string e = lower(env);
        p.precision = (e == "bf16" || e == "mixed") ? 1 : 0;
        p.precision_source = "env-precision";
    }
    return p;
}

void dyn_print_compute_plan(const ComputePlan& p, const char* prefix) {
    const double GiB = 1024.0 * 1024.0 * 1024.0;
    std::printf("%s layout : %s (one sorted copy per mode on the device)\n", prefix, p.layout);
    std::printf("%s device : tensor=%.2f GiB  factors=%.2f GiB  total=%.2f GiB  free=%.2f GiB%s\n",
                prefix, p.device_tensor_bytes / GiB, p.device_factor_bytes / GiB,
                p.device_total_bytes / GiB, p.device_free_bytes / GiB,
                p.fits ? "" : "  (DOES NOT FIT)");
    std::printf("%s kernel : %s  (reason=%s; src=%s)\n", prefix,
                p.kernel_enum ? "Jacobi" : "GaussSeidel", p.kernel_reason, p.kernel_source);
    std::printf("%s gather : %s  (src=%s)  Rp(host)=%d  Rp(device)=%d\n", prefix,
                p.precision ? "bf16" : "fp32", p.precision_source, p.rank_padded,
                p.device_rank_padded);
}

}  // namespace dynasor
