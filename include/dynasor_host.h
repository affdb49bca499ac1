/*
 * dynasor_host.h -- C-ABI of the host-side helpers of the engine
 * (libdynasor_host.so): factor initialisation streams and the synthetic
 * tensor generator.  The C++ drop-in API (cpals(), Tensor, loaders, plan)
 * lives in dynasor_plus_amd/csrc/host/dynasor_cpals.h.
 */
#ifndef DYNASOR_HOST_H
#define DYNASOR_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* CP-ALS initial factors as the reference draws them
 * (src/dynasor_cpals.cpp:42-69): per mode n, std::mt19937_64 seeded with a
 * splitmix64 step of (seed + 0x9E3779B97F4A7C15 * (n + 1)), then
 * std::normal_distribution<float>(0, 1) row-major over I_n x R; columns
 * [R, ld) are zeroed. */
void dyn_host_init_normal(int N, const uint32_t* dims, int R, int ld, uint64_t seed,
                          float* const* Y);
/* U[0,1) factors as the reference's MTTKRP benchmark draws them
 * (src/main.cpp:326-337, 900-907): one std::mt19937_64(seed) and
 * std::uniform_real_distribution<float>(0, 1) over the modes in order. */
void dyn_host_init_uniform(int N, const uint32_t* dims, int R, int ld, uint64_t seed,
                           float* const* Y);

/* Synthetic sparse tensors with DISTINCT coordinates (no reference
 * counterpart: scripts/gen_tensor.cpp allows duplicates).  Coordinates are
 * drawn one nonzero at a time from a counter-based stream keyed by
 * (seed, mode, draw index); the first `nnz` distinct coordinates in draw
 * order are kept and emitted in row-major (lexicographic) order.  Zipf(alpha)
 * on {0..I-1} has P(k) ~ (k+1)^-alpha, mapped through a seeded permutation
 * of [0, I) so heavy rows are scattered. */
#define DYN_GEN_COORD_UNIFORM   0
#define DYN_GEN_COORD_ZIPF      1
#define DYN_GEN_VAL_LOGNORMAL   0   /* exp(mu + sigma z) */
#define DYN_GEN_VAL_PLANTED     1   /* sum_r prod_n A_n[i_n, r] + N(0, (noise_frac*RMS)^2) */
#define DYN_GEN_VAL_BERNOULLI   2   /* y ~ Bernoulli(sigmoid(planted model)) as 0/1 */
#define DYN_GEN_VAL_UNIFORM     3   /* U[0,1) */
#define DYN_GEN_FACTOR_UNIFORM  0   /* planted factor entries U[0,1) */
#define DYN_GEN_FACTOR_NORMAL   1   /* planted factor entries N(0, planted_var) */

typedef struct dyn_gen_spec {
    int      num_modes;
    uint32_t dims[8];
    uint64_t nnz;
    uint64_t seed;
    int      coord_dist;
    double   alpha;
    int      value_kind;
    double   mu, sigma;
    int      planted_rank;
    int      planted_factor_dist;
    double   planted_var;
    double   noise_frac;
} dyn_gen_spec;

/* idx_out[n]: nnz uint32 each; vals_out: nnz floats; vals64_out (nullable):
 * the same values in double (before the float cast); stats_out (nullable,
 * 4 doubles): rounds, total draws, max row nnz of mode 0, rms of the planted
 * model.  threads <= 0: OpenMP default.  Returns 0 on success. */
int dyn_gen_tensor(const dyn_gen_spec* spec, uint32_t* const* idx_out, float* vals_out,
                   double* vals64_out, double* stats_out, int threads);

/* Tensor loaders behind a C-ABI (replace the reference's load_coo_tensor /
 * save_dnb_tensor, src/dynasor_io.cpp:1184-1261, 854-893, for callers that
 * cannot use the C++ API).  dyn_host_tensor_load: <path>.dnb when fresh,
 * else the FROSTT text (then <path>.dnb is written when cache_on_success).
 * Returns 0 on success, 2 on a parse / format error. */
typedef struct dyn_host_tensor dyn_host_tensor;
int  dyn_host_tensor_load(const char* path, int cache_on_success, dyn_host_tensor** out);
int  dyn_host_tensor_info(const dyn_host_tensor* t, int* num_modes, uint64_t* nnz,
                          uint32_t* dims);
int  dyn_host_tensor_copy(const dyn_host_tensor* t, float* vals, uint32_t* const* idx);
int  dyn_host_tensor_save_dnb(const dyn_host_tensor* t, const char* path);
void dyn_host_tensor_free(dyn_host_tensor* t);

#ifdef __cplusplus
}
#endif

#endif /* DYNASOR_HOST_H */
