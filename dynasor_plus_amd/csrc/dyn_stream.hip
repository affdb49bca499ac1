// Out-of-HBM tensors: the COO tensor stays in caller host memory and is
// streamed through device chunk buffers once per all-modes sweep, the
// device counterpart of the reference's out-of-core Jacobi driver
// (include/dynasor_ooc.h:25-61, src/dynasor_ooc.cpp:340-535), which streams
// .dnb chunks from disk through RAM.
//
// Per chunk and mode the chunk is sorted by that mode on the device (the
// in-core remap: radix sort, copy, row pointer, work tables), its segmented
// spMTTKRP is written to a scratch row block, and the block is added into
// Yhat_n; chunks are added in order, so a sweep is deterministic.  The
// frequency relabel (for the LDS hot rows) and ||T||^2 come from one pass
// over the chunks at attach time, which also checks every index.
#include "dyn_internal.hpp"

#include <algorithm>
#include <vector>

namespace dyn {

namespace {

__global__ void k_add_rows(float* __restrict__ dst, const float* __restrict__ src, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] += src[i];
}

__global__ void k_count_oob_s(const uint32_t* __restrict__ ix, uint64_t n, uint32_t dim,
                              unsigned long long* bad) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long local = 0;
    for (; i < n; i += stride) local += (ix[i] >= dim) ? 1ull : 0ull;
    if (local) atomicAdd(bad, local);
}

unsigned grid_of(uint64_t n) {
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + kBlock - 1) / kBlock, 16384));
}

// copy chunk [b, b + n) of the host tensor into the device chunk buffers
int load_chunk(dyn_gpu_ctx* c, uint64_t b, uint64_t n) {
    DYN_HIP(c, hipMemcpyAsync(c->vals, c->host_vals + b, n * sizeof(float), hipMemcpyHostToDevice,
                              c->stream));
    for (int w = 0; w < c->N; ++w)
        DYN_HIP(c, hipMemcpyAsync(c->idx[w], c->host_idx[w] + b, n * sizeof(uint32_t),
                                  hipMemcpyHostToDevice, c->stream));
    return DYN_OK;
}

}  // namespace

int reject_streamed(dyn_gpu_ctx* c, const char* what) {
    if (!c->streamed) return DYN_OK;
    return set_err(c, DYN_ERR_STATE,
                   "%s needs the tensor in device memory; this context streams it from host "
                   "memory (only all-modes spMTTKRP, per-mode spMTTKRP and Jacobi CP-ALS run)",
                   what);
}

int stream_sweep(dyn_gpu_ctx* c, uint32_t mask, const int* prec) {
    const int Rp = c->Rp;
    size_t need = 0;
    for (int n = 0; n < c->N; ++n) need = std::max<size_t>(need, (size_t)c->dims[n] * Rp);
    if (c->stream_tmp_elems < need) {
        dfree(c, c->stream_tmp);
        c->stream_tmp = nullptr;
        c->stream_tmp_elems = 0;
        DYN_CHECK(dalloc_t(c, &c->stream_tmp, need));
        c->stream_tmp_elems = need;
    }
    for (int n = 0; n < c->N; ++n)
        if ((mask >> n) & 1u) DYN_CHECK(fill_f32(c, c->Yhat[n], (size_t)c->dims[n] * Rp, 0.0f));
    DYN_CHECK(ensure_gather(c));
    const void* Yb[kMaxModes];
    const void* Y32[kMaxModes];
    for (int w = 0; w < c->N; ++w) {
        Yb[w] = c->Ygb[w];
        Y32[w] = c->Yg[w];
    }
    const uint64_t total = c->total_nnz;
    int rc = DYN_OK;
    for (uint64_t b = 0; b < total && rc == DYN_OK; b += c->chunk_nnz) {
        const uint64_t nb = std::min<uint64_t>(c->chunk_nnz, total - b);
        if ((rc = load_chunk(c, b, nb)) != DYN_OK) break;
        c->nnz = nb;
        for (int n = 0; n < c->N && rc == DYN_OK; ++n) {
            if (!((mask >> n) & 1u)) continue;
            rc = remap_mode(c, n, n == 0 ? 1 : 0);
            if (rc == DYN_OK)
                rc = mttkrp_mode(c, n, prec[n], prec[n] == DYN_PREC_BF16 ? Yb : Y32, nullptr,
                                 c->stream_tmp, Rp);
            if (rc == DYN_OK) {
                const size_t e = (size_t)c->dims[n] * Rp;
                hipLaunchKernelGGL(k_add_rows, dim3(grid_of(e)), dim3(kBlock), 0, c->stream,
                                   c->Yhat[n], c->stream_tmp, e);
                rc = check_launch(c, "add_rows");
            }
            free_copy(c, n);
        }
        // the next chunk overwrites the buffers the copies were gathered from
        if (rc == DYN_OK && hipStreamSynchronize(c->stream) != hipSuccess)
            rc = set_err(c, DYN_ERR_HIP, "stream sweep failed");
    }
    c->nnz = total;
    return rc;
}

}  // namespace dyn

using namespace dyn;

extern "C" int dyn_gpu_attach_host_coo(dyn_gpu_ctx* c, int N, uint64_t nnz, const uint32_t* dims,
                                       const float* vals, const uint32_t* const* idx,
                                       uint64_t chunk_nnz) {
    if (!c) return DYN_ERR_INVALID;
    if (N < 2 || N > kMaxModes)
        return set_err(c, DYN_ERR_INVALID, "num_modes must be in [2, %d], got %d", kMaxModes, N);
    if (nnz == 0) return set_err(c, DYN_ERR_INVALID, "a streamed tensor needs nonzeros");
    if (!dims || !vals || !idx) return set_err(c, DYN_ERR_INVALID, "dims/vals/idx is null");
    for (int n = 0; n < N; ++n) {
        if (dims[n] == 0) return set_err(c, DYN_ERR_INVALID, "mode %d has size 0", n);
        if (!idx[n]) return set_err(c, DYN_ERR_INVALID, "idx[%d] is null", n);
    }
    uint64_t chunk = chunk_nnz ? chunk_nnz : (1ull << 27);
    chunk = std::min<uint64_t>(std::min<uint64_t>(chunk, nnz), kMaxNnz - 1);
    DYN_HIP(c, hipSetDevice(c->device));
    DYN_HIP(c, hipStreamSynchronize(c->stream));
    free_tensor(c);
    free_factors(c);
    DYN_CHECK(dalloc_t(c, &c->vals, chunk));
    for (int n = 0; n < N; ++n) DYN_CHECK(dalloc_t(c, &c->idx[n], chunk));
    c->N = N;
    for (int n = 0; n < N; ++n) c->dims[n] = dims[n];
    c->streamed = true;
    c->host_vals = vals;
    for (int n = 0; n < N; ++n) c->host_idx[n] = idx[n];
    c->total_nnz = nnz;
    c->chunk_nnz = chunk;

    // one pass over the chunks: index check, per-row counts (the frequency
    // relabel) and ||T||^2, summed in chunk order
    size_t cnt_total = 0;
    for (int n = 0; n < N; ++n) cnt_total += ((size_t)dims[n] + 63) & ~size_t(63);
    uint32_t* cnt = nullptr;
    unsigned long long* bad = nullptr;
    int rc = dalloc_t(c, &cnt, cnt_total);
    if (rc == DYN_OK) rc = dalloc_t(c, &bad, (size_t)kMaxModes);
    if (rc == DYN_OK) rc = zero_u32(c, cnt, cnt_total);
    if (rc == DYN_OK) rc = zero_u32(c, (uint32_t*)bad, 2 * (size_t)kMaxModes);
    double tn2 = 0.0;
    for (uint64_t b = 0; b < nnz && rc == DYN_OK; b += chunk) {
        const uint64_t nb = std::min<uint64_t>(chunk, nnz - b);
        if ((rc = load_chunk(c, b, nb)) != DYN_OK) break;
        size_t off = 0;
        for (int n = 0; n < N && rc == DYN_OK; ++n) {
            hipLaunchKernelGGL(k_count_oob_s, dim3(grid_of(nb)), dim3(kBlock), 0, c->stream,
                               c->idx[n], nb, dims[n], bad + n);
            rc = check_launch(c, "count_oob");
            // out-of-range entries fall outside every histogram bin
            if (rc == DYN_OK) rc = count_rows(c, c->idx[n], nb, dims[n], cnt + off);
            off += ((size_t)dims[n] + 63) & ~size_t(63);
        }
        if (rc != DYN_OK) break;
        c->nnz = nb;
        c->tnorm2 = -1.0;
        double part = 0.0;
        rc = tensor_norm2(c, &part);   // synchronises: the chunk buffers are free again
        tn2 += part;
    }
    unsigned long long hb[kMaxModes] = {0};
    if (rc == DYN_OK) {
        const hipError_t e = hipMemcpy(hb, bad, sizeof(hb), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = set_err(c, DYN_ERR_HIP, "attach: %s", hipGetErrorString(e));
    }
    for (int n = 0; n < N && rc == DYN_OK; ++n)
        if (hb[n])
            rc = set_err(c, DYN_ERR_INVALID, "%llu indices of mode %d are >= its size %u", hb[n],
                         n, dims[n]);
    size_t off = 0;
    for (int n = 0; n < N && rc == DYN_OK; ++n) {
        rc = relabel_from_counts(c, n, cnt + off);
        off += ((size_t)dims[n] + 63) & ~size_t(63);
    }
    dfree(c, cnt);
    dfree(c, bad);
    if (rc != DYN_OK) {
        free_tensor(c);
        return rc;
    }
    c->relabeled = true;
    c->nnz = nnz;
    c->tnorm2 = tn2;
    return DYN_OK;
}
