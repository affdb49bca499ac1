// Internal state of a dyn_gpu_ctx and helpers shared by the HIP translation
// units of the engine.  Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/dynasor_gpu.h"

namespace dyn {

constexpr int kMaxModes = DYN_GPU_MAX_MODES;
constexpr int kBlock = 256;          // threads per workgroup for the streaming kernels
constexpr int kMaxRp = DYN_GPU_MAX_RANK;

typedef __hip_bfloat16 bf16;

// Smallest power of two >= max(R, 16): the device row stride of every factor.
inline int padded_rank(int R) {
    int p = 16;
    while (p < R) p <<= 1;
    return p;
}

// One mode's sorted copy: the nonzeros ordered by idx[mode] (ties in input
// order, or by a secondary key), with that mode's index column dropped and
// replaced by a row pointer.  The other modes' index columns hold their
// frequency ranks (relab[w][idx]) so hot factor rows are a contiguous prefix.
//
// Work tables (built on the host from the row pointer):
//   * a heavy row (> kHeavyMin nonzeros) is cut into pieces of kPiece
//     nonzeros; piece p covers [pc_k[p], pc_e[p]) of row pc_row[p] and writes
//     a partial row to slot p; the pieces of one heavy row are consecutive
//     (heavy_first / heavy_count) and are summed in slot order;
//   * light rows are never split: consecutive light rows whose first nonzero
//     falls in the same kLightWindow-aligned window form one segment
//     [seg_k[i], seg_e[i]) starting at row seg_row[i], stored directly;
//   * rows without nonzeros are listed in empty_rows and zeroed by the
//     spMTTKRP kernel itself, so no output row needs clearing beforehand.
struct ModeCopy {
    bool      built = false;
    int       secondary = -1;
    bool      rank_key = false;                 // secondary key: rank (true) or input index
    float*    vals = nullptr;                   // nnz
    uint32_t* idx[kMaxModes] = {nullptr};      // relab[w][idx[w]] for w != mode
    uint32_t* rowptr = nullptr;                 // I_mode + 1
    uint32_t* perm = nullptr;                   // nnz : sorted position -> input position
    uint32_t  nseg = 0;                         // light segments
    uint32_t* seg_k = nullptr;
    uint32_t* seg_e = nullptr;
    uint32_t* seg_row = nullptr;
    uint32_t  npieces = 0;                      // heavy-row pieces
    uint32_t* pc_k = nullptr;
    uint32_t* pc_e = nullptr;
    uint32_t* pc_row = nullptr;
    uint32_t* pc_hot = nullptr;                 // null, or 1 when every nonzero of the
                                                // piece has hot_w rank < hot_h, else 0
    // hot split: inside every row, the nonzeros whose mode-hot_w rank is below
    // hot_h come first (stable), and heavy pieces never straddle the split
    int       hot_w = -1;
    uint32_t  hot_h = 0;
    // all-hot class split (4+ modes): inside every row of at least kClassMin
    // nonzeros, the nonzeros whose every other-mode rank w is below cls_h[w]
    // (in operand order) come first; pc_hot marks their pieces.  All zero: none.
    uint32_t  cls_h[kMaxModes] = {0};
    uint32_t  nheavy = 0;
    uint32_t* heavy_row = nullptr;              // nheavy
    uint32_t* heavy_first = nullptr;            // nheavy : first piece
    uint32_t* heavy_count = nullptr;            // nheavy : pieces
    uint32_t  nempty = 0;
    uint32_t* empty_rows = nullptr;             // nempty
};

// The linearised (Z-order) copy: every nonzero once, ordered by a 64-bit key
// that interleaves the bits of the modes' frequency ranks, with explicit rank
// columns for every mode.  One pass over it produces every mode's spMTTKRP
// (dyn_linear.hip).  The key only orders the copy; coordinates stay exact
// even when the ranks have more than 64 bits in total (the key then uses the
// leading bits of the largest modes' ranks).
struct LinearCopy {
    bool      built = false;
    float*    vals = nullptr;                   // nnz, Z-order
    uint32_t* idx[kMaxModes] = {nullptr};      // relab[n][idx[n]], Z-order
    uint32_t* perm = nullptr;                   // Z-order position -> input position
    uint64_t  mask[kMaxModes] = {0};           // key bits of mode n
    uint8_t   bits[kMaxModes] = {0};           // bits of mode n's rank (>= 1)
    uint8_t   shift[kMaxModes] = {0};          // low rank bits left out of the key
    double*   acc = nullptr;                    // rank-order FP64 accumulators, sum_n I_n x Rp
    size_t    acc_count = 0;
};

constexpr uint32_t kPieceLane = 4;                   // nonzeros per lane of a piece
constexpr uint32_t kPiece = 64 * kPieceLane;         // nonzeros per heavy piece (one wave)
constexpr uint32_t kHeavyMin = 256;                  // longer rows are heavy
constexpr uint32_t kLightWindow = 128;
constexpr uint32_t kClassMin = 4096;                 // shortest row given an all-hot class
constexpr uint64_t kMaxNnz = 0xFFFFFFFFull;          // nonzeros per device (32-bit positions)

struct TimedPair {
    hipEvent_t a, b;
};

struct Streamer;   // out-of-HBM source, chunk buffers and pass state (dyn_stream.hip)

}  // namespace dyn

struct dyn_gpu_ctx {
    int          device = 0;
    hipStream_t  own_stream = nullptr;
    hipStream_t  stream = nullptr;
    hipStream_t  aux_stream = nullptr;   // CP-ALS: R x R work overlapped with spMTTKRP
    std::string  err;

    // memory accounting
    std::unordered_map<void*, size_t> allocs;
    uint64_t     mem_used = 0;
    uint64_t     mem_cap = 0;      // 0 = unlimited

    // tensor (input order)
    int          N = 0;
    uint64_t     nnz = 0;
    uint32_t     dims[dyn::kMaxModes] = {0};
    float*       vals = nullptr;
    uint32_t*    idx[dyn::kMaxModes] = {nullptr};
    double       tnorm2 = -1.0;    // ||T||^2 (FP64), computed lazily
    dyn::ModeCopy copy[dyn::kMaxModes];

    // factors
    int          R = 0;
    int          Rp = 0;
    float*       Y[dyn::kMaxModes] = {nullptr};
    // gathered operands, rows in frequency-rank order: Yg[w][relab[w][i]] = Y[w][i]
    float*       Yg[dyn::kMaxModes] = {nullptr};
    dyn::bf16*   Ygb[dyn::kMaxModes] = {nullptr};
    float*       Yhat[dyn::kMaxModes] = {nullptr};
    double*      G[dyn::kMaxModes] = {nullptr};   // Rp x Rp FP64 Grams
    double*      lambda = nullptr;                // Rp
    double*      inv_lambda = nullptr;            // Rp
    double*      L = nullptr;                     // Rp x Rp inverse of the Hadamard Gram
    double*      fitbuf = nullptr;                // device scalars of the fit
    double*      fitbuf_host = nullptr;           // pinned mirror
    bool         gather_valid = false;            // Yg / Ygb mirror Y

    // frequency relabelling of every mode (rank 0 = most frequent index)
    bool         relabeled = false;
    uint32_t*    relab[dyn::kMaxModes] = {nullptr};
    std::vector<uint32_t> relab_inv_host[dyn::kMaxModes];
    std::vector<uint32_t> rank_count_host[dyn::kMaxModes];   // nonzeros of each rank

    // linearised layout: when use_linear is set, spMTTKRPs run as one pass
    // over the Z-order copy instead of the per-mode copies
    bool         use_linear = false;
    dyn::LinearCopy lin;

    // device properties
    int          num_cus = 256;
    size_t       max_lds = 65536;

    // out-of-HBM streaming (dyn_gpu_attach_host_coo / dyn_gpu_attach_dnb):
    // the tensor stays in host memory or in a .dnb file and passes through
    // two device chunk buffers once per sweep; nnz is the whole tensor's
    // count except inside a pass, where it is the current chunk's and vals /
    // idx point at the chunk's buffers
    bool         streamed = false;
    dyn::Streamer* streamer = nullptr;            // owned (dyn_stream.hip)

    // allocation cache: while set, dfree parks blocks for dalloc to hand out
    // again at the same size (no hipFree, which waits for the whole device,
    // inside a streamed pass); flush_alloc_cache releases them
    bool         cache_allocs = false;
    std::multimap<size_t, void*> free_cache;

    // device address of this device's fixed partials array (no-allocation
    // entry point), from hipGetSymbolAddress at context creation
    void*        fixed_scratch = nullptr;

    // scratch (grown on demand)
    void*        scratch = nullptr;
    size_t       scratch_bytes = 0;
    void*        scratch2 = nullptr;
    size_t       scratch2_bytes = 0;

    // timing of spMTTKRP launches
    bool         timing = false;
    std::vector<dyn::TimedPair> timed;
    std::vector<dyn::TimedPair> event_pool;
    // events of the CP-ALS loop that bracket one segment kernel in place of a
    // timing pair (one record each side instead of two); the loop adds their
    // times below when timing is on
    const hipEvent_t* seg_pair = nullptr;
    // events the next segment-kernel launch records itself, at its start
    // and end (part of the dispatch, so no separate marker on the stream)
    hipEvent_t   launch_ev[2] = {nullptr, nullptr};
    double       timed_extra_ms = 0.0;
    uint64_t     timed_extra_n = 0;
};

namespace dyn {

// error helpers -------------------------------------------------------------
int set_err(dyn_gpu_ctx* c, int code, const char* fmt, ...);

#define DYN_HIP(ctx, call)                                                    \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess)                                                 \
            return ::dyn::set_err((ctx), DYN_ERR_HIP, "%s failed: %s (%s:%d)", \
                                  #call, hipGetErrorString(e_), __FILE__,     \
                                  __LINE__);                                  \
    } while (0)

#define DYN_CHECK(expr)                 \
    do {                                \
        int rc_ = (expr);               \
        if (rc_ != DYN_OK) return rc_;  \
    } while (0)

// device memory owned by the context ---------------------------------------
int dalloc(dyn_gpu_ctx* c, void** p, size_t bytes);
void dfree(dyn_gpu_ctx* c, void* p);
template <typename T>
inline int dalloc_t(dyn_gpu_ctx* c, T** p, size_t count) {
    return dalloc(c, reinterpret_cast<void**>(p), count * sizeof(T));
}
// scratch buffer of at least `bytes` (contents not preserved across calls)
int scratch(dyn_gpu_ctx* c, size_t bytes, void** p);
int scratch2(dyn_gpu_ctx* c, size_t bytes, void** p);

int check_launch(dyn_gpu_ctx* c, const char* what);
// Raise a kernel's dynamic-LDS limit once per device (the attribute is set
// for the current device); `done` is the kernel's own flag word, one bit per
// device id below 64.
void lds_limit_once(dyn_gpu_ctx* c, std::atomic<uint64_t>* done, const void* fn, int bytes);

// building blocks implemented in the .hip units -----------------------------
// Copy `mode` sorted by (index of `mode`, secondary key), where the secondary
// key is the secondary mode's input index (the reference's order) unless
// `engine_order` is set and that mode has more than 2048 rows: then it is the
// secondary mode's frequency rank (the order in which its factor rows sit in
// memory, so consecutive nonzeros of a row gather neighbouring rows), with
// the 1024 most frequent ranks hashed over the range so that the LDS-resident
// rows spread evenly along each row.  secondary < 0: stable by input position.
int remap_mode(dyn_gpu_ctx* c, int mode, int secondary, bool engine_order = true);
// Columns and row pointer of copy `mode` from cp.perm and the sorted keys.
int materialize_copy(dyn_gpu_ctx* c, int mode, const uint32_t* keys);
// Work tables from the row pointer (checked monotone and spanning nnz).
int finish_copy(dyn_gpu_ctx* c, int mode);
// Copy `mode` from a stored permutation (host), checked to be a permutation
// that sorts idx[mode]; no sort runs.
int restore_copy(dyn_gpu_ctx* c, int mode, int secondary, const uint32_t* perm_host);
void free_copy(dyn_gpu_ctx* c, int mode);
// drop the tensor (copies, relabel, streaming state) / the factor storage
void free_tensor(dyn_gpu_ctx* c);
void free_factors(dyn_gpu_ctx* c);
// Reorders copy `mode` so that in every row the nonzeros with relabelled
// mode-w index < H precede the others (stable within both classes) and
// rebuilds its work tables with hot/cold pieces.  No-op when already split
// for (w, H).
int ensure_hot_split(dyn_gpu_ctx* c, int mode, int w, uint32_t H);
// Reorders copy `mode` so that in every row of at least kClassMin nonzeros the
// nonzeros whose operand-j rank is below H[j] for every other mode j (operands
// in mode order) precede the others, stable within both, and rebuilds its work
// tables with those pieces flagged in pc_hot.  No-op when already split at H.
int ensure_class_split(dyn_gpu_ctx* c, int mode, const uint32_t* H);

// Segmented spMTTKRP of mode `mode` from the per-mode copy into `out`
// (I_mode x Rp floats, device).  `Yin[w]` are the gathered factors (FP32 or
// BF16 according to `prec`; entry `mode` ignored), `vals_override` replaces
// the copy's values with vals_override[perm[k]] when non-null.
// `part_buf` (nullable): device memory for the heavy-row piece partials, at
// least seg_part_bytes(c, mode, Rp) bytes; the call then allocates nothing and
// requires split_ready(c, mode, Rp).
int mttkrp_mode(dyn_gpu_ctx* c, int mode, int prec, const void* const* Yin,
                const float* vals_override, float* out, int Rp, float* part_buf = nullptr);
size_t seg_part_bytes(const dyn_gpu_ctx* c, int mode, int Rp);
// Whether copy `mode` is already in the order a launch at padded rank Rp
// uses (hot split on 3-mode tensors, plain order otherwise).
bool split_ready(const dyn_gpu_ctx* c, int mode, int Rp);
// Adjoint of the forward model for mode `mode`: the spMTTKRP of the
// per-nonzero vector g (input order) with caller factors A (input order,
// Rp-strided device rows), into out (I_mode x Rp).
int adjoint_mode(dyn_gpu_ctx* c, int mode, const float* const* A, int Rp, const float* g,
                 float* out);

// Streamed sweep: Yhat[n] for every mode n in `mask` over all chunks of an
// attached out-of-HBM tensor, mode n gathering at precision prec[n]; chunk
// results are added in chunk order.
int stream_sweep(dyn_gpu_ctx* c, uint32_t mask, const int* prec);
// DYN_ERR_STATE for entry points that need the whole tensor on the device
int reject_streamed(dyn_gpu_ctx* c, const char* what);
// Releases the streamer (loader state, chunk buffers, staging, copy stream)
void stream_release(dyn_gpu_ctx* c);
// hipFree every block parked by the allocation cache
void flush_alloc_cache(dyn_gpu_ctx* c);

// Sorting and gather building blocks (dyn_remap.hip).  sort_pairs: stable
// LSD radix sort of (k0, v0) over the low `bits` bits, ping-ponging with
// (k1, v1); *which (in/out) names the buffer pair holding the data.
int sort_pairs(dyn_gpu_ctx* c, uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, uint32_t n,
               int bits, int* which);
int iota_u32(dyn_gpu_ctx* c, uint32_t* p, uint32_t n);
int gather_u32(dyn_gpu_ctx* c, const uint32_t* src, const uint32_t* perm, uint32_t* dst,
               uint32_t n);
int gather_f32(dyn_gpu_ctx* c, const float* src, const uint32_t* perm, float* dst, uint32_t n);
// dst[p] = relab[src[perm[p]]]
int gather_relabel(dyn_gpu_ctx* c, const uint32_t* src, const uint32_t* perm,
                   const uint32_t* relab, uint32_t* dst, uint32_t n);
unsigned grid_for(uint64_t n);

// Linearised layout (dyn_linear.hip): build the Z-order copy; the all-modes
// pass writes Yhat[n] (input row order) for every mode n in `mask`.
int build_linear(dyn_gpu_ctx* c);
void free_linear(dyn_gpu_ctx* c);
int linear_sweep(dyn_gpu_ctx* c, uint32_t mask);

int forward_eval(dyn_gpu_ctx* c, const float* const* A, int Rp,
                 const double* lambda_or_null, float* m_out, const float* vals_dot,
                 double* dot_out_dev);
int completion_grad(dyn_gpu_ctx* c, int loss, const float* const* A, int Rp,
                    float* const* grad, double* loss_out);

int tensor_norm2(dyn_gpu_ctx* c, double* out);
int gram_mode(dyn_gpu_ctx* c, int mode);
int ensure_relabel(dyn_gpu_ctx* c);
// histogram of ix[0, n) over [0, dim) added into cnt (device)
int count_rows(dyn_gpu_ctx* c, const uint32_t* ix, uint64_t n, uint32_t dim, uint32_t* cnt);
int zero_u32(dyn_gpu_ctx* c, uint32_t* p, size_t n);
// relab[w] (allocated if needed) and its host inverse from per-row counts
int relabel_from_counts(dyn_gpu_ctx* c, int w, const uint32_t* cnt);
// Yg / Ygb of every mode from Y (after set_factors)
int ensure_gather(dyn_gpu_ctx* c);
// dst32[relab[i]] = src[i] (and dstbf, if non-null), rows x Rp
int relabel_rows(dyn_gpu_ctx* c, const float* src, const uint32_t* relab, float* dst32,
                 bf16* dstbf, size_t rows, int Rp);
void free_relabel(dyn_gpu_ctx* c);
int fill_f32(dyn_gpu_ctx* c, float* p, size_t n, float v);

bool timing_acquire(dyn_gpu_ctx* c, TimedPair* tp);   // a pool pair, not yet recorded

}  // namespace dyn
