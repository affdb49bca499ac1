// A caller of the drop-in C++ API written the way the reference's CLI drives
// CP-ALS (src/main.cpp:770-879): it touches every Tensor / FactorMatrices /
// CpOptions / ComputePlan / CpResult member and free function that block
// uses, so compiling and linking it shows that such a caller builds against
// the engine's headers and libraries unchanged.
//
//   drop_in_caller <tensor.tns> <rank> <kernel> [--plan-only]
#include "dynasor_common.h"
#include "dynasor_cpals.h"
#include "dynasor_jit.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

using namespace dynasor;

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s <tensor.tns> <rank> <kernel> [--plan-only]\n", argv[0]);
        return 1;
    }
    const std::string path = argv[1];
    const int rank = std::atoi(argv[2]);
    const std::string kernel = argv[3];
    const bool plan_only = argc > 4 && std::strcmp(argv[4], "--plan-only") == 0;

    Tensor T;
    if (!load_coo_tensor(path, T)) return 2;

    FactorMatrices F;
    F.allocate(T.num_modes, T.mode_size, rank);
    CpOptions opt;
    opt.max_iters = 5;
    opt.tol = 1e-4;
    opt.seed = 42;
    opt.num_threads = 0;
    opt.ridge = (value_t)1e-10;
    opt.verbose = true;
    opt.init_dump_prefix = "";

    ComputePlan cp = dyn_build_compute_plan(T, F.rank_padded, 1, "auto", "auto", kernel);
    const bool runtime_skip_preprocess = false;
    if (runtime_skip_preprocess) {
        cp.storage.choice = T.layout;
        cp.storage.reason = "loaded from a processed-state file";
        cp.storage.override_source = "dnp-runtime";
    }
    if (cp.pf_far) setenv("DYN_PF_FAR", "1", 1);
    if (cp.use_jit) dyn_jit_force_enable(true);
    cp.use_jit = dyn_jit_enabled();
    dyn_print_compute_plan(cp, "[plan]");
    opt.kernel = (cp.kernel_enum == 1) ? CpKernel::Jacobi : CpKernel::GaussSeidel;
    std::printf("[cpals] kernel -> %s  (reason=%s; src=%s)\n",
                opt.kernel == CpKernel::Jacobi ? "jacobi" : "gauss-seidel", cp.kernel_reason,
                cp.kernel_source);
    std::printf("[caller] storage choice=%d reason=%s src=%s per_mode0=%d ofibs=%zu\n",
                (int)cp.storage.choice, cp.storage.reason, cp.storage.override_source,
                (int)cp.per_mode[0], cp.ofibs_bytes);
    if (plan_only) return 0;

    CpResult res = cpals(T, F, opt);
    for (int n = 0; n < T.num_modes; ++n) {
        const std::string p = std::string("/tmp/drop_in_caller.cp_fm") + std::to_string(n) + ".bin";
        if (!dump_factor_matrix_bin(p, F.Y[n], T.mode_size[n], F.rank, F.rank_padded))
            std::fprintf(stderr, "[io] WARN: failed to write %s\n", p.c_str());
    }
    if (std::FILE* fp = std::fopen("/tmp/drop_in_caller.cp_lambda.bin", "wb")) {
        std::fwrite(res.lambda.data(), sizeof(value_t), (size_t)rank, fp);
        std::fclose(fp);
    }
    std::printf("[summary] cpals iters=%d  final_fit=%.6f  residual=%.3e  total=%.3f s  "
                "spMTTKRP=%.3f s  linalg=%.3f s\n",
                res.iters_done, res.final_fit, res.final_residual, res.total_time_s,
                res.mttkrp_time_s, res.linalg_time_s);
    F.free_all();
    return res.iters_done > 0 ? 0 : 4;
}
