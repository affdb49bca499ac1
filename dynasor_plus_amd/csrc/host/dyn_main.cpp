// dynasor_plus_gpu -- command line of the MI355X engine.  Mirrors the
// reference CLI (src/main.cpp:39-323, 770-976) for the paths it implements:
//   --decompose cpals ...      CP-ALS on the device (same flags, same
//                              [cpals] / [summary] lines and DYFM dumps)
//   (no --decompose)           one spMTTKRP sweep over every mode, with
//                              --verify / --dump / --dump-bin / --all-modes, so
//                              the reference's scripts/run_tests.sh can drive it
// --layout morton selects the linearised Z-order copy and --ooc / --layout ooc
// the host-streamed tensor; the CPU-only layout flags of the reference
// (--layout pingpong|ncopy|inplace, --ncopy-*, --shard-size, --jit,
// --threads) are accepted and ignored; --save-processed / --runtime write and
// read the device's processed state (dynasor_dnp.h).
#include "dynasor_common.h"
#include "dynasor_cpals.h"
#include "dynasor_dnp.h"
#include "dynasor_gpu.h"
#include "dynasor_host.h"
#include "dynasor_ooc.h"

#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

using namespace dynasor;

namespace {

struct Args {
    std::string tensor_path;
    int rank = 16;
    int threads = 0;
    bool verify = false;
    std::string dump_prefix;
    bool dump_binary = false;
    bool no_cache = false;
    uint64_t seed = 42;
    std::string decompose;
    int cp_iters = 50;
    double cp_tol = 1e-4;
    std::string cp_kernel = "auto";
    double cp_ridge = 1e-10;
    std::string cp_dump, cp_dump_init, cp_init_load;
    std::string precision;
    bool all_modes = false;
    bool uniform_init = false;
    int device = 0;
    std::string save_processed;
    bool runtime_only = false;
    std::string ooc = "auto";
    std::string layout = "auto";
    uint64_t ooc_chunk = 0;
};

void usage(const char* a0) {
    std::printf(
        "Usage: %s <tensor.tns> [options]\n"
        "  --rank R               factor rank (default 16)\n"
        "  --seed S               RNG seed (default 42)\n"
        "  --no-cache             do not read or write the .dnb binary cache\n"
        "  --verify               spMTTKRP sweep: also run a CPU check, PASS if the\n"
        "                         worst per-mode Frobenius-relative error <= 1e-4\n"
        "  --dump PREFIX          write Yhat_n to <PREFIX>.fm<n>_out.txt (.bin with --dump-bin)\n"
        "  --all-modes            spMTTKRP sweep from one factor snapshot (Jacobi pass)\n"
        "  --decompose cpals      run CP-ALS\n"
        "  --cp-iters N           max iterations (default 50)\n"
        "  --cp-tol TOL           |dfit| threshold (default 1e-4)\n"
        "  --cp-kernel K          auto | gauss-seidel | jacobi (env DYN_CP_KERNEL)\n"
        "  --cp-ridge R           relative Cholesky ridge (default 1e-10)\n"
        "  --cp-dump PREFIX       write <PREFIX>.cp_fm<n>.bin and .cp_lambda.bin\n"
        "  --cp-dump-init PREFIX  write the initial factors <PREFIX>.cp_init_fm<n>.bin\n"
        "  --cp-init-load PREFIX  start from <PREFIX>.cp_init_fm<n>.bin\n"
        "  --cp-init-uniform      U[0,1) initial factors instead of N(0,1)\n"
        "  --precision P          fp32 | bf16 gathered factors (env DYN_CP_PRECISION)\n"
        "  --device D             GPU ordinal (default 0)\n"
        "  --ooc on|off|auto      stream the tensor from host memory once per sweep\n"
        "                         (Jacobi); auto: when the device copies do not fit\n"
        "  --ooc-chunk NNZ        streaming chunk in nonzeros (default 2^27)\n"
        "  --layout L             auto | morton (one linearised Z-order copy, all-modes\n"
        "                         pass; auto picks it when the per-mode copies do not fit\n"
        "                         but one copy does) | ooc; CPU layouts are accepted and\n"
        "                         ignored\n"
        "  --save-processed P     write the processed device state (tensor + per-mode\n"
        "                         permutations) to P after the copies are built\n"
        "  --runtime              <tensor> is a processed-state file: load it and skip\n"
        "                         the device sort (needs --decompose cpals)\n",
        a0);
}

bool parse(int argc, char** argv, Args& a) {
    if (argc < 2) return false;
    a.tensor_path = argv[1];
    if (a.tensor_path == "-h" || a.tensor_path == "--help") return false;
    for (int i = 2; i < argc; ++i) {
        const std::string k = argv[i];
        auto need = [&](const char* n) -> const char* {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "missing value for %s\n", n);
                std::exit(1);
            }
            return argv[++i];
        };
        if (k == "--rank") a.rank = std::atoi(need("--rank"));
        else if (k == "--threads") a.threads = std::atoi(need("--threads"));
        else if (k == "--seed") a.seed = (uint64_t)std::atoll(need("--seed"));
        else if (k == "--verify") a.verify = true;
        else if (k == "--dump") a.dump_prefix = need("--dump");
        else if (k == "--dump-bin") a.dump_binary = true;
        else if (k == "--no-cache") a.no_cache = true;
        else if (k == "--decompose") a.decompose = need("--decompose");
        else if (k == "--cp-iters") a.cp_iters = std::atoi(need("--cp-iters"));
        else if (k == "--cp-tol") a.cp_tol = std::atof(need("--cp-tol"));
        else if (k == "--cp-kernel") a.cp_kernel = need("--cp-kernel");
        else if (k == "--cp-ridge") a.cp_ridge = std::atof(need("--cp-ridge"));
        else if (k == "--cp-dump") a.cp_dump = need("--cp-dump");
        else if (k == "--cp-dump-init") a.cp_dump_init = need("--cp-dump-init");
        else if (k == "--cp-init-load") a.cp_init_load = need("--cp-init-load");
        else if (k == "--cp-init-uniform") a.uniform_init = true;
        else if (k == "--precision") a.precision = need("--precision");
        else if (k == "--device") a.device = std::atoi(need("--device"));
        else if (k == "--all-modes") a.all_modes = true;
        else if (k == "--ooc") a.ooc = need("--ooc");
        else if (k == "--layout") a.layout = need("--layout");
        else if (k == "--ooc-chunk") a.ooc_chunk = (uint64_t)std::atoll(need("--ooc-chunk"));
        else if (k == "--shard-size" || k == "--ncopy-fiber-sort" || k == "--ncopy-csr" ||
                 k == "--ncopy-csr-compact") {
            (void)need(k.c_str());   // CPU layout knobs: no device counterpart
        } else if (k == "--jit") {
        } else if (k == "--save-processed") {
            a.save_processed = need("--save-processed");
        } else if (k == "--runtime") {
            a.runtime_only = true;
        } else if (k == "-h" || k == "--help") {
            return false;
        } else {
            std::fprintf(stderr, "unknown option: %s\n", k.c_str());
            return false;
        }
    }
    return true;
}

// --layout values that select the linearised copy (the reference's "morton")
bool is_linear_layout(const std::string& layout) {
    std::string l = layout;
    for (char& ch : l) ch = (char)std::tolower((unsigned char)ch);
    return l == "morton" || l == "linear" || l == "zorder" || l == "alto";
}

// spMTTKRP by definition (src/dynasor_reference.cpp:18-50), the CLI's own
// --verify check of the device result
void cpu_check_mttkrp(const Tensor& T, const FactorMatrices& F, int n, std::vector<float>& out) {
    const int R = F.rank, Rp = F.rank_padded;
    out.assign((size_t)T.mode_size[n] * Rp, 0.0f);
    std::vector<float> tmp(R);
    for (uint64_t k = 0; k < T.nnz; ++k) {
        if (T.vals[k] == 0.0f) continue;
        for (int r = 0; r < R; ++r) tmp[r] = T.vals[k];
        for (int w = 0; w < T.num_modes; ++w) {
            if (w == n) continue;
            const float* row = F.Y[w] + (size_t)T.idx_buf[w][k] * Rp;
            for (int r = 0; r < R; ++r) tmp[r] *= row[r];
        }
        float* o = out.data() + (size_t)T.idx_buf[n][k] * Rp;
        for (int r = 0; r < R; ++r) o[r] += tmp[r];
    }
}

int mttkrp_mode_main(const Args& args, Tensor& T) {
    FactorMatrices F;
    F.allocate(T.num_modes, T.mode_size, args.rank);
    {
        // U[0,1) inputs from one std::mt19937_64(seed), mode by mode (src/main.cpp:900-907)
        std::mt19937_64 rng(args.seed);
        std::uniform_real_distribution<float> U(0.0f, 1.0f);
        for (int n = 0; n < T.num_modes; ++n)
            for (idx_t i = 0; i < T.mode_size[n]; ++i) {
                float* row = F.Y[n] + (size_t)i * F.rank_padded;
                for (int r = 0; r < F.rank; ++r) row[r] = U(rng);
            }
    }
    dyn_gpu_ctx* h = nullptr;
    int rc = dyn_gpu_create(args.device, &h);
    if (rc == DYN_OK) rc = dyn_gpu_upload_coo(h, T.num_modes, T.nnz, T.mode_size, T.vals, T.idx_buf);
    if (rc == DYN_OK) rc = dyn_gpu_set_factors(h, F.rank, F.Y.data(), F.rank_padded);
    const bool linear = is_linear_layout(args.layout);
    if (rc == DYN_OK) rc = linear ? dyn_gpu_set_layout(h, DYN_LAYOUT_LINEAR) : dyn_gpu_remap_all(h);
    if (rc != DYN_OK) {
        std::fprintf(stderr, "[gpu] setup failed (status %d): %s\n", rc,
                     h ? dyn_gpu_last_error(h) : "no device");
        if (h) dyn_gpu_destroy(h);
        return 2;
    }
    const int prec = (args.precision == "bf16") ? DYN_PREC_BF16 : DYN_PREC_FP32;
    dyn_gpu_synchronize(h);
    const auto t0 = std::chrono::steady_clock::now();
    if (args.all_modes) {
        rc = dyn_gpu_mttkrp_all(h, prec, F.Yhat.data(), F.rank_padded);
    } else {
        for (int n = 0; n < T.num_modes && rc == DYN_OK; ++n)
            rc = dyn_gpu_mttkrp(h, n, prec, F.Yhat[n], F.rank_padded);
    }
    dyn_gpu_synchronize(h);
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc != DYN_OK) {
        std::fprintf(stderr, "[gpu] spMTTKRP failed (status %d): %s\n", rc, dyn_gpu_last_error(h));
        dyn_gpu_destroy(h);
        return 2;
    }
    dyn_gpu_destroy(h);
    std::printf("[dynasor] device spMTTKRP (%s, %s%s) over %d modes = %.6f s\n",
                args.all_modes ? "all-modes" : "per-mode", (prec && !linear) ? "bf16" : "fp32",
                linear ? ", linearised Z-order copy" : "", T.num_modes, secs);
    if (!args.dump_prefix.empty()) {
        const char* ext = args.dump_binary ? ".bin" : ".txt";
        for (int n = 0; n < T.num_modes; ++n) {
            const std::string p = args.dump_prefix + ".fm" + std::to_string(n) + "_out" + ext;
            const bool ok = args.dump_binary
                                ? dump_factor_matrix_bin(p, F.Yhat[n], T.mode_size[n], F.rank,
                                                         F.rank_padded)
                                : dump_factor_matrix(p, F.Yhat[n], T.mode_size[n], F.rank,
                                                     F.rank_padded);
            if (!ok) std::fprintf(stderr, "[io] WARN: failed to write %s\n", p.c_str());
        }
        std::printf("[io] wrote Yhat[*] to %s.fm<n>_out%s\n", args.dump_prefix.c_str(), ext);
    }
    int ret = 0;
    if (args.verify) {
        double worst = 0.0;
        std::vector<float> ref;
        for (int n = 0; n < T.num_modes; ++n) {
            cpu_check_mttkrp(T, F, n, ref);
            double sd = 0.0, sr = 0.0;
            for (idx_t i = 0; i < T.mode_size[n]; ++i)
                for (int r = 0; r < F.rank; ++r) {
                    const size_t o = (size_t)i * F.rank_padded + r;
                    const double d = (double)F.Yhat[n][o] - (double)ref[o];
                    sd += d * d;
                    sr += (double)ref[o] * ref[o];
                }
            const double fro = sr > 0.0 ? std::sqrt(sd / sr) : std::sqrt(sd);
            worst = std::max(worst, fro);
            std::printf("[compare] mode %d : frob_rel = %.3e\n", n, fro);
        }
        std::printf("[compare] worst Frobenius rel = %.6g\n", worst);
        if (worst > 1e-4) {
            std::printf("FAIL: Frobenius relative error %.3e exceeds %.1e\n", worst, 1e-4);
            ret = 3;
        } else {
            std::printf("PASS: Frobenius relative error %.3e within %.1e\n", worst, 1e-4);
        }
    }
    std::printf("[summary] dynasor = %.6f s (%.3f Mnnz/s)\n", secs,
                (double)T.nnz / 1e6 / std::max(secs, 1e-12));
    return ret;
}

}  // namespace

int main(int argc, char** argv) {
    std::setvbuf(stdout, nullptr, _IONBF, 0);
    Args args;
    if (!parse(argc, argv, args)) {
        usage(argv[0]);
        return 1;
    }
    std::printf("=== Dynasor+ MI355X engine (backend: HIP gfx950) ===\n");
    std::unique_ptr<Tensor> Tp(new Tensor());
    Tensor& T = *Tp;
    std::string stream_dnb;   // non-empty: stream the tensor from this .dnb file
    if (args.runtime_only) {
        // src/main.cpp:362-427: a processed file replaces load + preprocessing
        if (args.verify || args.decompose != "cpals") {
            std::fprintf(stderr, "[main] --runtime needs --decompose cpals and excludes --verify\n");
            return 7;
        }
        if (!load_dnp_header(args.tensor_path, T)) return 2;
        std::printf("[runtime] loading processed state: %s\n          modes=%d  nnz=%llu  "
                    "layout=device-copies\n",
                    args.tensor_path.c_str(), T.num_modes, (unsigned long long)T.nnz);
    } else {
        // src/main.cpp:444-513: with a .dnb cache beside the tensor, plan from
        // its header alone; a plan that streams leaves the tensor on disk
        if (args.decompose == "cpals" && !args.verify && !args.no_cache) {
            Tensor hdr;
            if (ooc_load_header(args.tensor_path + ".dnb", hdr)) {
                const int Rp = std::max((args.rank + 7) / 8 * 8, (args.rank + 15) / 16 * 16);
                const ComputePlan pre = dyn_build_compute_plan(hdr, Rp, args.threads, args.layout,
                                                               args.ooc, args.cp_kernel);
                if (pre.storage.choice == Layout::OutOfCore) {
                    stream_dnb = args.tensor_path + ".dnb";
                    T.num_modes = hdr.num_modes;
                    T.nnz = hdr.nnz;
                    for (int n = 0; n < hdr.num_modes; ++n) T.mode_size[n] = hdr.mode_size[n];
                    std::printf("[main] OUT-OF-CORE: the tensor stays in %s (nnz %llu, %d modes)\n",
                                stream_dnb.c_str(), (unsigned long long)T.nnz, T.num_modes);
                }
            }
        }
        if (stream_dnb.empty() && !load_coo_tensor(args.tensor_path, T, !args.no_cache)) {
            std::fprintf(stderr, "failed to load tensor '%s'\n", args.tensor_path.c_str());
            return 2;
        }
    }
    if (args.decompose == "cpals") {
        FactorMatrices F;
        F.allocate(T.num_modes, T.mode_size, args.rank);
        CpOptions opt;
        opt.max_iters = args.cp_iters;
        opt.tol = args.cp_tol;
        opt.seed = args.seed;
        opt.num_threads = args.threads;
        opt.ridge = (value_t)args.cp_ridge;
        opt.verbose = true;
        opt.init_dump_prefix = args.cp_dump_init;
        opt.init_load_prefix = args.cp_init_load;
        opt.uniform_init = args.uniform_init;
        opt.device = args.device;
        opt.save_processed_path = args.save_processed;
        if (args.runtime_only) opt.processed_path = args.tensor_path;
        ComputePlan cp = dyn_build_compute_plan(T, F.rank_padded, args.threads, args.layout,
                                                args.ooc, args.cp_kernel);
        if (!args.precision.empty()) {
            cp.precision = (args.precision == "bf16" || args.precision == "mixed") ? 1 : 0;
            cp.precision_source = "cli-precision";
        }
        dyn_print_compute_plan(cp, "[plan]");
        if (cp.storage.choice == Layout::OutOfCore || !stream_dnb.empty()) {
            if (args.runtime_only) {
                std::fprintf(stderr, "[main] --runtime and out-of-core streaming are mutually "
                                     "exclusive\n");
                return 7;
            }
            opt.stream_chunk_nnz = args.ooc_chunk ? args.ooc_chunk : ooc_default_chunk_nnz(T);
            if (!stream_dnb.empty())
                opt.stream_path = stream_dnb;
            else
                opt.stream_from_host = true;
            std::printf("[ooc] streaming the tensor from %s in chunks of %llu nonzeros\n",
                        stream_dnb.empty() ? "host memory" : stream_dnb.c_str(),
                        (unsigned long long)opt.stream_chunk_nnz);
        } else if (cp.storage.choice == Layout::Morton) {
            // src/main.cpp:712-733: the tensor is converted to the linearised
            // layout, which cpals() runs with the all-modes update
            T.layout = Layout::Morton;
            std::printf("[layout] linearised Z-order copy (%.2f GiB on the device)\n",
                        cp.device_linear_bytes / 1073741824.0);
        } else if (!cp.fits) {
            // --ooc off: the device engine needs the tensor and its copies resident
            std::fprintf(stderr,
                         "[plan] refusing: the device tensor and its copies need %.2f GiB, "
                         "%.2f GiB of device memory are free (out-of-core is off)\n",
                         cp.device_total_bytes / 1073741824.0,
                         cp.device_free_bytes / 1073741824.0);
            return 6;
        }
        opt.kernel = cp.kernel_enum == 1 ? CpKernel::Jacobi : CpKernel::GaussSeidel;
        opt.precision = cp.precision == 1 ? CpPrecision::BF16 : CpPrecision::FP32;
        std::printf("[cpals] kernel -> %s  (reason=%s; src=%s)\n",
                    opt.kernel == CpKernel::Jacobi ? "jacobi (all-modes snapshot)"
                                                   : "gauss-seidel (classical per-mode)",
                    cp.kernel_reason, cp.kernel_source);
        CpResult res = cpals(T, F, opt);
        if (!args.cp_dump_init.empty() && res.iters_done > 0)
            std::printf("[io] wrote initial factors to %s.cp_init_fm<n>.bin\n",
                        args.cp_dump_init.c_str());
        if (!args.cp_dump.empty() && res.iters_done > 0) {
            for (int n = 0; n < T.num_modes; ++n) {
                const std::string p = args.cp_dump + ".cp_fm" + std::to_string(n) + ".bin";
                if (!dump_factor_matrix_bin(p, F.Y[n], T.mode_size[n], F.rank, F.rank_padded))
                    std::fprintf(stderr, "[io] WARN: failed to write %s\n", p.c_str());
            }
            const std::string lp = args.cp_dump + ".cp_lambda.bin";
            if (std::FILE* fp = std::fopen(lp.c_str(), "wb")) {
                std::fwrite(res.lambda.data(), sizeof(value_t), (size_t)args.rank, fp);
                std::fclose(fp);
            }
            std::printf("[io] wrote CP factors to %s.cp_fm<n>.bin and lambda to %s\n",
                        args.cp_dump.c_str(), lp.c_str());
        }
        std::printf("[summary] cpals iters=%d  final_fit=%.6f  residual=%.3e  total=%.3f s  "
                    "spMTTKRP=%.3f s  linalg=%.3f s\n",
                    res.iters_done, res.final_fit, res.final_residual, res.total_time_s,
                    res.mttkrp_time_s, res.linalg_time_s);
        return res.iters_done > 0 ? 0 : 4;
    }
    if (!args.decompose.empty()) {
        std::fprintf(stderr, "[main] unknown --decompose algorithm: %s (supported: cpals)\n",
                     args.decompose.c_str());
        return 5;
    }
    if (!args.save_processed.empty()) {
        if (!save_dnp_tensor(args.save_processed, T, args.threads, args.device)) return 2;
        std::printf("[io] wrote processed state to %s\n", args.save_processed.c_str());
    }
    return mttkrp_mode_main(args, T);
}
