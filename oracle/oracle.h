/*
 * oracle.h -- CPU oracle for the sparse CP-ALS path.  TEST INFRASTRUCTURE:
 * only tests/, the smoke check and the bench's CPU-baseline leg load it, and
 * only as the checker.  Nothing in the engine links or calls it.
 *
 * Each function restates one piece of the reference (paths relative to the
 * reference tree) in plain single-threaded C++, in the same operation order,
 * so its results can be compared with the reference binary's own outputs:
 *
 *   orc_mttkrp          src/dynasor_reference.cpp:18-50
 *   orc_init_normal     src/dynasor_cpals.cpp:42-69   (CP-ALS factor init)
 *   orc_init_uniform    src/main.cpp:326-337, 900-907 (--verify inputs)
 *   orc_remap           src/dynasor_preprocess.cpp:471-521, 1915-1951
 *                       (stable counting sort by row + CSR row pointer)
 *   orc_cpals           src/dynasor_cpals.cpp:120-396 with
 *                       src/dynasor_linalg.cpp:47-402 (FP32 algebra exactly
 *                       as the reference; fp64 = 1 runs everything in FP64)
 *   orc_forward /
 *   orc_completion_loss no reference counterpart: the CP model at the stored
 *                       coordinates and the sparse-completion losses in FP64.
 *
 * How each is pinned (tests/test_oracle.py): against the reference binary's
 * dumps (tests/golden/), against closed forms (rank-1 MTTKRP, dense
 * reconstruction of the fit, numpy Cholesky), and by invariants (row
 * multisets of the remap, <Yhat_n, Y_n> equal for every n).
 */
#ifndef DYN_ORACLE_H
#define DYN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_cp_opts {
    int    max_iters;
    double tol;
    double ridge;
    int    jacobi;             /* 0: Gauss-Seidel, 1: Jacobi (all MTTKRPs first) */
    int    fp64;               /* 0: FP32 factors and algebra as the reference; 1: FP64 */
    int    stale_jacobi_fit;   /* 1: Jacobi fit from the Yhat identity as the reference
                                  prints it (not the true fit); 0: true fit */
    int    bf16_gather;        /* 1: the spMTTKRP reads factor rows rounded to bfloat16
                                  (round to nearest even of the FP32 value), except
                                  for the mode whose Yhat carries the fit (Gauss-
                                  Seidel: the last; Jacobi: the smallest), which
                                  reads FP32 rows: the engine's mixed mode */
} orc_cp_opts;

int orc_mttkrp(int N, uint64_t nnz, const uint32_t* dims, const float* vals,
               const uint32_t* const* idx, int R, int ld, const float* const* Y,
               int mode, float* out);
int orc_mttkrp_f64(int N, uint64_t nnz, const uint32_t* dims, const float* vals,
                   const uint32_t* const* idx, int R, int ld, const double* const* Y,
                   int mode, double* out);

void orc_init_normal(int N, const uint32_t* dims, int R, int ld, uint64_t seed,
                     float* const* Y);
void orc_init_uniform(int N, const uint32_t* dims, int R, int ld, uint64_t seed,
                      float* const* Y);

int orc_remap(int N, uint64_t nnz, const uint32_t* dims, const uint32_t* const* idx,
              int mode, int secondary, uint32_t* perm, uint64_t* rowptr);

/* Y: in = initial factors (row-major I_n x R), out = final normalised factors.
 * fit_hist: >= max_iters doubles; lambda_out: R doubles. Returns iterations. */
int orc_cpals(int N, uint64_t nnz, const uint32_t* dims, const float* vals,
              const uint32_t* const* idx, int R, float* const* Y,
              const orc_cp_opts* opt, double* fit_hist, double* lambda_out,
              double* final_fit);

void orc_forward(int N, uint64_t nnz, const uint32_t* const* idx, int R, int ld,
                 const double* const* A, const double* lambda, double* m_out);
double orc_completion_loss(int N, uint64_t nnz, const uint32_t* const* idx,
                           const float* y, int R, int ld, const double* const* A,
                           int loss);

#ifdef __cplusplus
}
#endif

#endif /* DYN_ORACLE_H */
