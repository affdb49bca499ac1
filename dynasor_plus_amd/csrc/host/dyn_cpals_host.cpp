// cpals(): the reference's CP-ALS entry point (include/dynasor_cpals.h:121,
// src/dynasor_cpals.cpp:120-396) on top of the device engine's C-ABI.
#include "dynasor_cpals.h"
#include "dynasor_ooc.h"
#include "dynasor_gpu.h"
#include "dynasor_host.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

namespace dynasor {

namespace {

struct Ctx {
    dyn_gpu_ctx* h = nullptr;
    ~Ctx() {
        if (h) dyn_gpu_destroy(h);
    }
};

}  // namespace

CpResult cpals(Tensor& T, FactorMatrices& F, const CpOptions& opt) {
    CpResult result;
    const int N = T.num_modes;
    const int R = F.rank;
    const int Rp = F.rank_padded;
    if (N < 2) {
        std::fprintf(stderr, "[cpals] need num_modes >= 2, got %d\n", N);
        return result;
    }
    if (R <= 0) {
        std::fprintf(stderr, "[cpals] need rank >= 1, got %d\n", R);
        return result;
    }
    if ((int)F.Y.size() != N) {
        std::fprintf(stderr, "[cpals] FactorMatrices not allocated for %d modes\n", N);
        return result;
    }

    // initial factors: the reference's seeded N(0,1) stream, U[0,1), or DYFM files
    if (!opt.init_load_prefix.empty()) {
        for (int n = 0; n < N; ++n) {
            const std::string p = opt.init_load_prefix + ".cp_init_fm" + std::to_string(n) + ".bin";
            if (!load_factor_matrix_bin(p, F.Y[n], T.mode_size[n], R, Rp)) {
                std::fprintf(stderr, "[cpals] cannot read initial factor %s\n", p.c_str());
                return result;
            }
        }
    } else if (opt.uniform_init) {
        dyn_host_init_uniform(N, T.mode_size, R, Rp, opt.seed, F.Y.data());
    } else {
        dyn_host_init_normal(N, T.mode_size, R, Rp, opt.seed, F.Y.data());
    }
    if (!opt.init_dump_prefix.empty()) {
        for (int n = 0; n < N; ++n) {
            const std::string p = opt.init_dump_prefix + ".cp_init_fm" + std::to_string(n) + ".bin";
            if (!dump_factor_matrix_bin(p, F.Y[n], T.mode_size[n], R, Rp))
                std::fprintf(stderr, "[cpals] WARN: failed to dump init factor %s\n", p.c_str());
        }
    }

    Ctx ctx;
    auto fail = [&](const char* what, int rc) {
        std::fprintf(stderr, "[cpals] %s failed (status %d): %s\n", what, rc,
                     ctx.h ? dyn_gpu_last_error(ctx.h) : "no device context");
        CpResult empty;
        return empty;
    };
    int rc = dyn_gpu_create(opt.device, &ctx.h);
    if (rc != DYN_OK) return fail("dyn_gpu_create", rc);
    if (!opt.processed_path.empty()) {
        // runtime-only: the device state comes from the processed file, which
        // must hold the tensor T describes
        int fN = 0;
        uint64_t fnnz = 0;
        uint32_t fdims[DYN_MAX_MODES] = {0}, mask = 0;
        rc = dyn_gpu_processed_info(opt.processed_path.c_str(), &fN, &fnnz, fdims, &mask);
        bool same = rc == DYN_OK && fN == N && fnnz == T.nnz;
        for (int n = 0; same && n < N; ++n) same = fdims[n] == T.mode_size[n];
        if (!same) {
            std::fprintf(stderr, "[cpals] '%s' does not hold this tensor\n",
                         opt.processed_path.c_str());
            return CpResult();
        }
        rc = dyn_gpu_load_processed(ctx.h, opt.processed_path.c_str());
        if (rc != DYN_OK) return fail("dyn_gpu_load_processed", rc);
    } else if (!opt.stream_path.empty()) {
        // the file must hold the tensor T describes
        Tensor hdr;
        bool same = ooc_load_header(opt.stream_path, hdr) && hdr.num_modes == N &&
                    hdr.nnz == T.nnz;
        for (int n = 0; same && n < N; ++n) same = hdr.mode_size[n] == T.mode_size[n];
        if (!same) {
            std::fprintf(stderr, "[cpals] '%s' does not hold this tensor\n",
                         opt.stream_path.c_str());
            return CpResult();
        }
        rc = dyn_gpu_attach_dnb(ctx.h, opt.stream_path.c_str(), opt.stream_chunk_nnz);
        if (rc != DYN_OK) return fail("dyn_gpu_attach_dnb", rc);
    } else if (opt.stream_from_host) {
        rc = dyn_gpu_attach_host_coo(ctx.h, N, T.nnz, T.mode_size, T.vals, T.idx_buf,
                                     opt.stream_chunk_nnz);
        if (rc != DYN_OK) return fail("dyn_gpu_attach_host_coo", rc);
    } else {
        rc = dyn_gpu_upload_coo(ctx.h, N, T.nnz, T.mode_size, T.vals, T.idx_buf);
        if (rc != DYN_OK) return fail("dyn_gpu_upload_coo", rc);
    }
    if (!opt.save_processed_path.empty()) {
        rc = dyn_gpu_remap_all(ctx.h);
        if (rc == DYN_OK) rc = dyn_gpu_save_processed(ctx.h, opt.save_processed_path.c_str());
        if (rc != DYN_OK) return fail("dyn_gpu_save_processed", rc);
        if (opt.verbose)
            std::printf("[io] wrote processed state to %s\n", opt.save_processed_path.c_str());
    }
    if (T.layout == Layout::Morton && opt.stream_path.empty() && !opt.stream_from_host) {
        // the linearised copy replaces the per-mode copies (all-modes only,
        // src/dynasor_cpals.cpp:232-240)
        rc = dyn_gpu_set_layout(ctx.h, DYN_LAYOUT_LINEAR);
        if (rc != DYN_OK) return fail("dyn_gpu_set_layout", rc);
    }
    rc = dyn_gpu_set_factors(ctx.h, R, F.Y.data(), Rp);
    if (rc != DYN_OK) return fail("dyn_gpu_set_factors", rc);

    dyn_cp_opts o{};
    o.max_iters = opt.max_iters;
    o.tol = opt.tol;
    o.ridge = (double)opt.ridge;
    o.kernel = opt.kernel == CpKernel::Jacobi ? DYN_CP_JACOBI : DYN_CP_GAUSS_SEIDEL;
    o.precision = opt.precision == CpPrecision::BF16 ? DYN_PREC_BF16 : DYN_PREC_FP32;
    o.verbose = opt.verbose ? 1 : 0;
    std::vector<double> hist(std::max(opt.max_iters, 1)), itim(std::max(opt.max_iters, 1));
    dyn_cp_result r{};
    r.fit_history = hist.data();
    r.iter_time_s = itim.data();
    rc = dyn_gpu_cpals(ctx.h, &o, &r);
    if (rc != DYN_OK) return fail("dyn_gpu_cpals", rc);

    for (int n = 0; n < N; ++n) {
        rc = dyn_gpu_get_factors(ctx.h, n, F.Y[n], Rp);
        if (rc != DYN_OK) return fail("dyn_gpu_get_factors", rc);
    }
    std::vector<double> lam(R);
    rc = dyn_gpu_get_lambda(ctx.h, lam.data());
    if (rc != DYN_OK) return fail("dyn_gpu_get_lambda", rc);

    result.iters_done = r.iters_done;
    result.final_fit = r.final_fit;
    result.final_residual = r.final_residual;
    result.tensor_norm = r.tensor_norm;
    result.total_time_s = r.total_time_s;
    result.mttkrp_time_s = r.mttkrp_time_s;
    result.linalg_time_s = r.linalg_time_s;
    result.remap_time_s = r.remap_time_s;
    result.fit_history.assign(hist.begin(), hist.begin() + r.iters_done);
    result.iter_time_s.assign(itim.begin(), itim.begin() + r.iters_done);
    result.lambda.resize(R);
    for (int i = 0; i < R; ++i) result.lambda[i] = (value_t)lam[i];
    if (opt.verbose) {
        std::printf("[cpals] total = %.3f s   spMTTKRP = %.3f s (%.1f%%)   linalg  = %.3f s (%.1f%%)"
                    "   remap = %.3f s (one-time)\n",
                    result.total_time_s, result.mttkrp_time_s,
                    100.0 * result.mttkrp_time_s / std::max(result.total_time_s, 1e-30),
                    result.linalg_time_s,
                    100.0 * result.linalg_time_s / std::max(result.total_time_s, 1e-30),
                    result.remap_time_s);
    }
    return result;
}

}  // namespace dynasor
