/*
 * dynasor_gpu.h -- C-ABI of the MI355X sparse CP-ALS engine.
 *
 * Plain C types only (pointers, sizes, fixed-width integers); no HIP or
 * torch types cross this boundary.  Every entry point returns an int status
 * (DYN_OK == 0) and never calls exit(); the last error text of a context is
 * available from dyn_gpu_last_error().
 *
 * Conventions
 *   - One context = one device + one HIP stream.  Calls on a context are
 *     serialised by the caller (one host thread); contexts are not shared
 *     across threads.
 *   - Host buffers are caller-owned and copied; device buffers are owned by
 *     the context.  Entry points with a `_dev` suffix take device pointers
 *     and run asynchronously on the context stream (see dyn_gpu_set_stream).
 *   - Factor matrices on the host are row-major I_n x R with a leading
 *     dimension `ld >= R` (the reference pads rows to Rp, include/
 *     dynasor_common.h:604-615 and src/dynasor_io.cpp:157-160, so callers can
 *     pass their padded buffers unchanged).
 *   - Indices are 0-based uint32 (include/dynasor_common.h:28), values are
 *     float32 (include/dynasor_common.h:30).  nnz must be < 2^32.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   dyn_gpu_upload_coo   <- Tensor SoA slabs filled by load_coo_tensor
 *                           (include/dynasor_common.h:322-600,
 *                            src/dynasor_io.cpp:1184-1261)
 *   dyn_gpu_remap        <- build_flycoo + NCopy fiber sort + CSR rowptr
 *                           (src/dynasor_preprocess.cpp:2013-2163, 1609-2008,
 *                            1467-1607)
 *   dyn_gpu_mttkrp       <- SpmttkrpSweep::process / dynasor_process_shard_
 *                           fiber_csr and the JIT C-ABI dyn_jit_fiber_csr
 *                           (src/dynasor.cpp:352-539, src/dynasor_kernel.cpp:
 *                            597-710, include/dynasor_jit.h:50-57)
 *   dyn_gpu_mttkrp_all   <- spmttkrp_all_modes_dynasor
 *                           (src/dynasor.cpp:592-817, src/dynasor_kernel.cpp:
 *                            830-1039)
 *   dyn_gpu_gram         <- la::gram (src/dynasor_linalg.cpp:47-99)
 *   dyn_gpu_cpals        <- cpals() (include/dynasor_cpals.h:121,
 *                           src/dynasor_cpals.cpp:120-396)
 *   dyn_gpu_forward / dyn_gpu_adjoint: no reference counterpart (sparse CP
 *                           completion); the adjoint is the reference MTTKRP
 *                           (src/dynasor_reference.cpp:18-50) applied to the
 *                           loss-gradient tensor.
 */
#ifndef DYNASOR_GPU_H
#define DYNASOR_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DYN_GPU_ABI_VERSION 1
#define DYN_GPU_MAX_MODES   8      /* DYN_MAX_MODES, include/dynasor_common.h:24 */
#define DYN_GPU_MAX_RANK    128    /* padded rank held in LDS by the R x R kernels */

/* status codes */
#define DYN_OK                 0
#define DYN_ERR_INVALID        1   /* bad argument (sizes, mode, rank, null) */
#define DYN_ERR_STATE          2   /* call order (e.g. mttkrp before remap) */
#define DYN_ERR_OOM            3   /* device allocation failed / over cap */
#define DYN_ERR_HIP            4   /* HIP runtime error */
#define DYN_ERR_UNSUPPORTED    5   /* shape outside the compiled grid */

/* gathered-operand precision of the spMTTKRP family */
#define DYN_PREC_FP32          0   /* FP32 factors, FP32 accumulate */
#define DYN_PREC_BF16          1   /* BF16 factor shadow gathered, FP32 accumulate */

/* CP-ALS sweep order (CpKernel, include/dynasor_cpals.h:77-80) */
#define DYN_CP_GAUSS_SEIDEL    0
#define DYN_CP_JACOBI          1

/* sparse-completion losses */
#define DYN_LOSS_LS            0   /* 0.5 (m - y)^2 */
#define DYN_LOSS_BERNOULLI     1   /* log(1 + e^m) - y m */

typedef struct dyn_gpu_ctx dyn_gpu_ctx;

/* ---- context --------------------------------------------------------- */
/* Create a context on `device`.  Environment DYN_GPU_MEM_CAP_MB (optional)
 * caps the context's device allocations so tests can force DYN_ERR_OOM. */
int         dyn_gpu_create(int device, dyn_gpu_ctx** out);
int         dyn_gpu_destroy(dyn_gpu_ctx* ctx);
const char* dyn_gpu_last_error(const dyn_gpu_ctx* ctx);
int         dyn_gpu_abi_version(void);
/* Run subsequent work on `stream`, a hipStream_t; NULL selects the device's
 * default (null) stream, which is what torch's default stream is.  Callers
 * that share buffers with another stream must pass that stream here. */
int         dyn_gpu_set_stream(dyn_gpu_ctx* ctx, void* stream);
/* Return to the context's own (non-blocking) stream. */
int         dyn_gpu_reset_stream(dyn_gpu_ctx* ctx);
int         dyn_gpu_synchronize(dyn_gpu_ctx* ctx);
/* Bytes of device memory currently held by the context. */
int         dyn_gpu_mem_used(const dyn_gpu_ctx* ctx, uint64_t* bytes);
/* Free / total device memory of `device` (no context needed); used by the
 * planner (dyn_build_compute_plan) the way the reference's planner queries
 * host RAM (src/dynasor_preprocess.cpp:951-1104). */
int         dyn_gpu_device_memory(int device, uint64_t* free_bytes, uint64_t* total_bytes);

/* ---- tensor ---------------------------------------------------------- */
/* Copy a COO tensor (SoA: vals[nnz], idx[n][nnz]) to the device.  Replaces
 * any previous tensor and drops per-mode copies and factors.  Indices must
 * satisfy idx[n][k] < dims[n] (checked on the device; DYN_ERR_INVALID).
 * nnz must be below 2^32 - 1 (positions are 32-bit, as in the reference's
 * .dnb format), else DYN_ERR_INVALID; columns of 4 GiB or more (2^30
 * nonzeros or more) are read through per-piece buffer ranges. */
int dyn_gpu_upload_coo(dyn_gpu_ctx* ctx, int num_modes, uint64_t nnz,
                       const uint32_t* dims, const float* vals,
                       const uint32_t* const* idx);

/* Build the mode-`mode` copy: a stable device radix sort of the nonzeros by
 * idx[mode] (ties keep input order), the gathered value / other-index
 * columns, the row pointer (rowptr[r]..rowptr[r+1] = nonzeros of row r) and
 * the work tables (heavy-row pieces, light-row segments).  Optional host
 * outputs: rowptr_out[dims[mode]+1], perm_out[nnz] (sorted position -> input
 * position).  `secondary` >= 0 additionally orders each row by
 * idx[secondary] (composite key); -1 keeps input order within rows.
 * spMTTKRPs on 3-mode tensors (FP32 or BF16 operands) later reorder the
 * nonzeros inside each row (hot split: those whose next-mode factor row is
 * held in LDS first, both classes stable), and BF16 spMTTKRPs on 4 to 7
 * modes those of rows of 4096+ nonzeros (the nonzeros whose every gathered
 * row is held in LDS first); the row pointer is unchanged. */
int dyn_gpu_remap(dyn_gpu_ctx* ctx, int mode, int secondary,
                  uint64_t* rowptr_out, uint32_t* perm_out);
/* Build every mode's copy that is not built yet, secondary mode = the first
 * other mode.  Unlike dyn_gpu_remap, whose secondary key is the input index,
 * these copies key the secondary sort by that mode's frequency rank (the
 * memory order of its factor rows) when it has more than 2048 rows, with its
 * 1024 most frequent ranks hashed over the range; rows and results are the
 * same, the order within rows differs. */
int dyn_gpu_remap_all(dyn_gpu_ctx* ctx);
/* Copy the mode-`mode` sorted copy back, in its current order (see the hot
 * split above): vals_out[nnz] and, for every mode w != mode, idx_out[w][nnz]
 * (idx_out[mode] is ignored and may be NULL). */
int dyn_gpu_get_copy(dyn_gpu_ctx* ctx, int mode, float* vals_out,
                     uint32_t* const* idx_out);

/* ---- out-of-HBM tensors (replaces the reference's out-of-core streaming,
 *      include/dynasor_ooc.h:25-61, src/dynasor_ooc.cpp:340-535) ---------- */
/* Attach a COO tensor that stays in caller host memory (vals[nnz],
 * idx[n][nnz], valid until the next upload/attach or destroy); it passes
 * through device chunk buffers of `chunk_nnz` nonzeros (0: 2^27) once per
 * sweep, the next chunk's upload overlapping the current chunk's work (a
 * loader thread and a copy stream).  One pass at attach checks every index
 * (DYN_ERR_INVALID), counts the
 * rows for the frequency relabel and sums ||T||^2.  Afterwards
 * dyn_gpu_set_factors, dyn_gpu_mttkrp, dyn_gpu_mttkrp_all, dyn_gpu_gram and
 * dyn_gpu_cpals (Jacobi; a Gauss-Seidel request runs Jacobi with a warning,
 * as the reference's out-of-core path does) work; each chunk is sorted per
 * mode on the device and its results are added in chunk order.  Entry
 * points that need the whole tensor on the device return DYN_ERR_STATE. */
int dyn_gpu_attach_host_coo(dyn_gpu_ctx* ctx, int num_modes, uint64_t nnz, const uint32_t* dims,
                            const float* vals, const uint32_t* const* idx, uint64_t chunk_nnz);
/* The same with the tensor left in a .dnb file (src/dynasor_io.cpp:747-774
 * layout; replaces OocStream::open / read_chunk, include/dynasor_ooc.h:25-61):
 * every pass reads each chunk's value and index ranges from the file into
 * pinned staging.  Only the header is read here before the attach pass; a
 * bad header or a size that does not match it returns DYN_ERR_INVALID. */
int dyn_gpu_attach_dnb(dyn_gpu_ctx* ctx, const char* path, uint64_t chunk_nnz);

/* ---- processed state (replaces save_dnp_tensor / load_dnp_tensor,
 *      include/dynasor_dnp.h:110-145) ------------------------------------ */
/* Write the uploaded tensor and the permutation of every built copy to
 * `path` ("DGP1" format: header, COO in input order, one u32 permutation
 * per copy), through `path`.tmp and a rename. */
int dyn_gpu_save_processed(dyn_gpu_ctx* ctx, const char* path);
/* Replace the context's tensor with the file's and rebuild the stored copies
 * by gathering through their permutations (no sort).  Every index is checked
 * against its mode size and every permutation must be a permutation of
 * [0, nnz) that orders its mode's indices, else DYN_ERR_INVALID.  Factors are
 * dropped, as by dyn_gpu_upload_coo. */
int dyn_gpu_load_processed(dyn_gpu_ctx* ctx, const char* path);
/* Header of a processed-state file (no device needed): mode count, nnz,
 * dims[num_modes] and the mask of stored copies; DYN_ERR_INVALID for a file
 * that is not one, or whose size does not match its header. */
int dyn_gpu_processed_info(const char* path, int* num_modes, uint64_t* nnz, uint32_t* dims,
                           uint32_t* copies_mask);

/* ---- factors --------------------------------------------------------- */
/* Allocate factor storage for rank R (device rows padded to a power of two
 * >= 16) and copy Y[n] (I_n x R, leading dimension ld) in; Y may be NULL to
 * only allocate (factors zeroed). */
int dyn_gpu_set_factors(dyn_gpu_ctx* ctx, int rank, const float* const* Y,
                        int ld);
int dyn_gpu_get_factors(dyn_gpu_ctx* ctx, int mode, float* Y, int ld);
int dyn_gpu_get_lambda(dyn_gpu_ctx* ctx, double* lambda);

/* ---- kernels --------------------------------------------------------- */
/* Yhat_mode = MTTKRP of the current factors (definition:
 * src/dynasor_reference.cpp:18-50).  Yhat_out (host, I_mode x R, ld) may be
 * NULL to leave the result on the device only. */
int dyn_gpu_mttkrp(dyn_gpu_ctx* ctx, int mode, int precision, float* Yhat_out,
                   int ld);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): dyn_gpu_mttkrp into the caller's device buffer Yhat_dev
 * (I_mode x ld floats, ld equal to the context's padded rank), with the
 * heavy-row partial sums in a fixed per-GPU device array of
 * DYN_GPU_NOALLOC_SCRATCH_BYTES instead of context scratch.  Allocates
 * nothing: the mode's copy must already be built and in launch order (an
 * earlier dyn_gpu_remap and dyn_gpu_mttkrp of that mode at this rank), else
 * DYN_ERR_STATE; a copy whose partials exceed the array returns
 * DYN_ERR_UNSUPPORTED.  Nothing in the array needs clearing between calls
 * (every partial is written before it is read), but all calls on one GPU
 * share it: calls that overlap on one GPU (several contexts or streams)
 * race.  Runs asynchronously on the context stream. */
#define DYN_GPU_NOALLOC_SCRATCH_BYTES (64u << 20)
int dyn_gpu_mttkrp_noalloc(dyn_gpu_ctx* ctx, int mode, int precision, float* Yhat_dev,
                           int ld);
/* Every Yhat_n from one snapshot of the factors (Jacobi all-modes
 * semantics).  With the per-mode copies (the default layout): one segmented,
 * atomic-free spMTTKRP launch per mode over that mode's sorted copy, all
 * reading the same factors.  With the linearised layout
 * (dyn_gpu_set_layout): one pass over the single Z-order copy that scatters
 * every mode's rows (FP64 atomics outside the hottest rows).  Yhat_out may be
 * NULL, or an array of N host pointers (entries may be NULL). */
int dyn_gpu_mttkrp_all(dyn_gpu_ctx* ctx, int precision, float* const* Yhat_out,
                       int ld);

/* ---- linearised (Z-order) layout (replaces the Morton / ALTO layout and its
 *      all-modes kernel: build_morton_layout, spmttkrp_all_modes_morton,
 *      morton_build_plan, include/dynasor_morton.h, src/dynasor_morton.cpp:
 *      103-138, 143-360, 398-579) ------------------------------------------- */
#define DYN_LAYOUT_COPIES      0   /* one sorted copy per mode (default) */
#define DYN_LAYOUT_LINEAR      1   /* one Z-order copy, single-pass all-modes */
/* Key plan of the Z-order copy, no device needed: bits[n] = ceil(log2 dims[n])
 * (at least 1); mask[n] = the key bit positions of mode n's rank, assigned
 * LSB-first round robin over the modes with bits left (the reference's
 * morton_build_plan); shift[n] = low rank bits left out of the key, nonzero
 * only when the bits exceed 64 (the largest modes give up bits until they
 * fit; the reference refuses such tensors).  Returns the key's bit count
 * (1..64), or -1 for invalid arguments (N outside 1..8, a zero dimension or
 * a null pointer). */
int dyn_linear_plan(int num_modes, const uint32_t* dims, uint64_t* mask, uint8_t* bits,
                    uint8_t* shift);
/* Build the Z-order copy of the uploaded tensor: the nonzeros sorted by the
 * key of their frequency ranks (a stable device radix sort), with one rank
 * column per mode.  perm_out (optional, host nnz): Z-order position -> input
 * position. */
int dyn_gpu_build_linear(dyn_gpu_ctx* ctx, uint32_t* perm_out);
/* Select the layout the spMTTKRPs and CP-ALS run on.  DYN_LAYOUT_LINEAR
 * builds the Z-order copy if needed and releases the per-mode copies (they
 * are rebuilt on demand after switching back).  On the linear layout
 * dyn_gpu_mttkrp and dyn_gpu_mttkrp_all run the all-modes pass (restricted
 * to the requested modes) with FP32 gathers whatever the precision argument
 * (the pass is bound by its scattered additions, not by its gathers),
 * and dyn_gpu_cpals runs Jacobi (a Gauss-Seidel request runs Jacobi with a
 * warning, as the reference's Morton path does, src/dynasor_cpals.cpp:
 * 232-240).  Results vary in the last bits from run to run (atomic
 * summation order).  Not available for streamed tensors (DYN_ERR_STATE). */
int dyn_gpu_set_layout(dyn_gpu_ctx* ctx, int layout);
/* G = Y_mode^T Y_mode (R x R, row-major, FP64 result of FP32 MFMA partials). */
int dyn_gpu_gram(dyn_gpu_ctx* ctx, int mode, double* G);

/* ---- CP-ALS ---------------------------------------------------------- */
typedef struct dyn_cp_opts {
    int      max_iters;     /* CpOptions::max_iters */
    double   tol;           /* stop when |fit_i - fit_{i-1}| < tol (i > 0) */
    double   ridge;         /* relative Cholesky ridge, CpOptions::ridge */
    int      kernel;        /* DYN_CP_GAUSS_SEIDEL | DYN_CP_JACOBI */
    int      precision;     /* DYN_PREC_FP32 | DYN_PREC_BF16 */
    int      verbose;       /* print "[cpals] iter ..." lines to stdout */
    int      reserved[5];
} dyn_cp_opts;

typedef struct dyn_cp_result {
    int      iters_done;
    double   final_fit;
    double   final_residual;
    double   tensor_norm;
    double   total_time_s;
    double   mttkrp_time_s;  /* sum of spMTTKRP kernel time (HIP events) */
    double   linalg_time_s;  /* the rest of the iteration */
    double   remap_time_s;   /* one-time per-mode copies built inside cpals */
    /* caller-provided arrays of length >= max_iters, or NULL */
    double*  fit_history;
    double*  iter_time_s;
} dyn_cp_result;

/* Run CP-ALS from the factors currently set (dyn_gpu_set_factors).  Builds
 * missing per-mode copies first.  On return the device factors are column-
 * normalised and lambda holds the column weights (dyn_gpu_get_lambda). */
int dyn_gpu_cpals(dyn_gpu_ctx* ctx, const dyn_cp_opts* opts,
                  dyn_cp_result* res);

/* ---- sparse completion: forward / adjoint ----------------------------- */
/* m[k] = sum_r lambda_r prod_n A_n[idx_n[k], r] at every stored coordinate,
 * in upload order, from the current factors (lambda: the context's lambda if
 * use_lambda, else ones).  m_out: host nnz floats. */
int dyn_gpu_forward(dyn_gpu_ctx* ctx, int use_lambda, float* m_out);
/* grad_out (host, I_mode x R, ld) = MTTKRP_mode of the sparse tensor with
 * values g[k] (upload order) at the stored coordinates, from the current
 * factors: the adjoint of dyn_gpu_forward with respect to A_mode. */
int dyn_gpu_adjoint(dyn_gpu_ctx* ctx, const float* g, int mode,
                    float* grad_out, int ld);

/* Device-pointer variants used by the autograd harness; they read factors
 * from A[n] (device, I_n x ld floats) rather than the context's factors. */
int dyn_gpu_forward_dev(dyn_gpu_ctx* ctx, const float* const* A, int ld,
                        float* m);
int dyn_gpu_adjoint_dev(dyn_gpu_ctx* ctx, const float* g, int mode,
                        const float* const* A, int ld, float* grad);
/* Fused loss + gradient for every mode over the per-mode copies, with the
 * stored tensor values as labels y: loss = sum_k l(m_k, y_k) and
 * grad[n] = MTTKRP_n(l'(m, y)).  A, grad: device, I_n x ld floats. */
int dyn_gpu_completion_grad_dev(dyn_gpu_ctx* ctx, int loss, const float* const* A,
                                int ld, float* const* grad, double* loss_out);

/* ---- timing ---------------------------------------------------------- */
/* When enabled, every spMTTKRP kernel launch records a pair of HIP events
 * at its start and end (as part of the dispatch, on the launch stream);
 * dyn_gpu_kernel_stats returns (and resets) the number of launches and
 * their summed duration in milliseconds. */
int dyn_gpu_enable_timing(dyn_gpu_ctx* ctx, int on);
int dyn_gpu_kernel_stats(dyn_gpu_ctx* ctx, uint64_t* launches, double* total_ms);

#ifdef __cplusplus
}
#endif

#endif /* DYNASOR_GPU_H */
