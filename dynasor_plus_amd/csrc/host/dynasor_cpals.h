// ============================================================================
//  dynasor_cpals.h -- CP-ALS entry point of the MI355X engine.
//
//  Drop-in for the reference's include/dynasor_cpals.h:77-121: the same
//  CpKernel / CpOptions / CpResult / cpals() names and meanings.  The
//  decomposition runs on the GPU (libdynasor_gpu, include/dynasor_gpu.h):
//  spMTTKRP in FP32 (or BF16-gathered operands), the R x R Hadamard /
//  Cholesky / solve, lambda normalisation and the fit in FP64.
//
//  Differences a caller can observe:
//    * num_threads is accepted and ignored (no OpenMP in the solve),
//    * the Jacobi fit is the true fit <T, M> from a forward pass, not the
//      Yhat[N-1] identity (which is stale under Jacobi),
//    * a GPU failure returns iters_done = 0 and prints the device error.
// ============================================================================
#ifndef DYNASOR_GPU_CPALS_H
#define DYNASOR_GPU_CPALS_H

#include "dynasor_common.h"

#include <cstdint>
#include <string>
#include <vector>

namespace dynasor {

enum class CpKernel {
    GaussSeidel,   // one spMTTKRP per mode, freshest factors
    Jacobi,        // every spMTTKRP from the iteration-start factors
};

enum class CpPrecision {
    FP32,          // FP32 gathered factors, FP32 accumulation
    BF16,          // BF16 gathered factors, FP32 accumulation
};

struct CpOptions {
    int         max_iters = 50;
    double      tol = 1e-4;            // stop when |fit_i - fit_{i-1}| < tol (i > 0)
    uint64_t    seed = 42;
    int         num_threads = 0;
    value_t     ridge = 1e-10f;        // relative ridge for the Cholesky solve
    bool        verbose = true;
    CpKernel    kernel = CpKernel::GaussSeidel;
    std::string init_dump_prefix;      // writes <prefix>.cp_init_fm<n>.bin
    // engine extensions
    CpPrecision precision = CpPrecision::FP32;
    int         device = 0;
    bool        uniform_init = false;  // U[0,1) (std::mt19937_64(seed)) instead of N(0,1)
    std::string init_load_prefix;      // read <prefix>.cp_init_fm<n>.bin instead of seeding
    // processed state (dynasor_dnp.h): load the device tensor and its copies
    // from this file instead of uploading T and sorting (T needs only its
    // header), and/or write the processed state after the copies are built
    std::string processed_path;
    std::string save_processed_path;
    // out-of-HBM: stream T through device chunks of this many nonzeros (0:
    // 2^27) once per sweep, from T's host arrays (stream_from_host) or from
    // the .dnb file stream_path (T then needs only num_modes / nnz /
    // mode_size, as ooc_load_header fills them); CP-ALS then runs Jacobi
    bool        stream_from_host = false;
    std::string stream_path;
    uint64_t    stream_chunk_nnz = 0;
};

struct CpResult {
    int                  iters_done = 0;
    double               final_fit = 0.0;
    double               final_residual = 0.0;
    double               tensor_norm = 0.0;
    double               total_time_s = 0.0;
    double               mttkrp_time_s = 0.0;
    double               linalg_time_s = 0.0;
    std::vector<double>  fit_history;
    std::vector<double>  iter_time_s;
    std::vector<value_t> lambda;
    double               remap_time_s = 0.0;   // one-time per-mode device copies
};

CpResult cpals(Tensor& T, FactorMatrices& F, const CpOptions& opt);

}  // namespace dynasor

#endif  // DYNASOR_GPU_CPALS_H
