// Context, memory, tensor/factor transfer and the thin C-ABI wrappers.
#include "dyn_internal.hpp"

#include <algorithm>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <mutex>

namespace dyn {

// ---------------------------------------------------------------------------
//  Fixed device scratch of dyn_gpu_mttkrp_noalloc: the heavy-row piece
//  partials live in this file-scope array instead of context scratch, so
//  that entry point allocates nothing.  Every GPU the library runs on holds
//  its own copy until the process ends.  Host code never passes the array's
//  name to a kernel: the name is a host-side stand-in, not a device address.
//  Each device's address comes from hipGetSymbolAddress (once per device, at
//  context creation, with that device current) and is what kernels receive.
// ---------------------------------------------------------------------------
namespace fixed {
constexpr size_t kBytes = DYN_GPU_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) float scratch[kBytes / sizeof(float)];

constexpr int kMaxDevices = 64;
static std::mutex g_mu;
static void* g_addr[kMaxDevices] = {nullptr};

// Device address of `device`'s array; `device` must be current.
static int address(int device, void** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return DYN_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_addr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return DYN_ERR_HIP;
        }
        g_addr[device] = p;
    }
    *out = g_addr[device];
    return DYN_OK;
}
}  // namespace fixed

int set_err(dyn_gpu_ctx* c, int code, const char* fmt, ...) {
    if (!c) return code;
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    c->err = buf;
    return code;
}

void flush_alloc_cache(dyn_gpu_ctx* c) {
    if (c->free_cache.empty()) return;
    // parked blocks may still be read by queued work on the context stream
    (void)hipStreamSynchronize(c->stream);
    for (auto& kv : c->free_cache) {
        auto it = c->allocs.find(kv.second);
        if (it != c->allocs.end()) {
            c->mem_used -= it->second;
            c->allocs.erase(it);
        }
        (void)hipFree(kv.second);
    }
    c->free_cache.clear();
}

int dalloc(dyn_gpu_ctx* c, void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 256;
    bytes = (bytes + 255) & ~size_t(255);
    if (c->cache_allocs) {
        // a parked block of this size: queued work that used it precedes
        // the new user on the same stream
        auto it = c->free_cache.find(bytes);
        if (it != c->free_cache.end()) {
            *p = it->second;
            c->free_cache.erase(it);
            return DYN_OK;
        }
    }
    if (c->mem_cap && c->mem_used + bytes > c->mem_cap) flush_alloc_cache(c);
    if (c->mem_cap && c->mem_used + bytes > c->mem_cap)
        return set_err(c, DYN_ERR_OOM,
                       "device allocation of %zu bytes exceeds the context cap "
                       "(%llu of %llu bytes in use)",
                       bytes, (unsigned long long)c->mem_used,
                       (unsigned long long)c->mem_cap);
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess && !c->free_cache.empty()) {
        (void)hipGetLastError();
        flush_alloc_cache(c);
        e = hipMalloc(p, bytes);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return set_err(c, DYN_ERR_OOM, "hipMalloc(%zu) failed: %s", bytes,
                       hipGetErrorString(e));
    }
    c->allocs[*p] = bytes;
    c->mem_used += bytes;
    return DYN_OK;
}

void dfree(dyn_gpu_ctx* c, void* p) {
    if (!p) return;
    auto it = c->allocs.find(p);
    if (it != c->allocs.end()) {
        if (c->cache_allocs) {
            c->free_cache.emplace(it->second, p);
            return;
        }
        c->mem_used -= it->second;
        c->allocs.erase(it);
    }
    (void)hipFree(p);
}

static int grow(dyn_gpu_ctx* c, void** buf, size_t* have, size_t bytes, void** p) {
    if (*have < bytes) {
        if (*buf) {
            (void)hipStreamSynchronize(c->stream);
            dfree(c, *buf);
            *buf = nullptr;
            *have = 0;
        }
        size_t want = std::max(bytes, size_t(1) << 20);
        DYN_CHECK(dalloc(c, buf, want));
        *have = want;
    }
    *p = *buf;
    return DYN_OK;
}

int scratch(dyn_gpu_ctx* c, size_t bytes, void** p) {
    return grow(c, &c->scratch, &c->scratch_bytes, bytes, p);
}
int scratch2(dyn_gpu_ctx* c, size_t bytes, void** p) {
    return grow(c, &c->scratch2, &c->scratch2_bytes, bytes, p);
}

int check_launch(dyn_gpu_ctx* c, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return set_err(c, DYN_ERR_HIP, "launch of %s failed: %s", what,
                       hipGetErrorString(e));
    return DYN_OK;
}

void lds_limit_once(dyn_gpu_ctx* c, std::atomic<uint64_t>* done, const void* fn, int bytes) {
    const uint64_t bit = (c->device >= 0 && c->device < 64) ? (1ull << c->device) : 0ull;
    if (bit && (done->load(std::memory_order_acquire) & bit)) return;
    (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (bit) done->fetch_or(bit, std::memory_order_acq_rel);
}

bool timing_acquire(dyn_gpu_ctx* c, TimedPair* tp) {
    if (!c->timing) return false;
    if (c->event_pool.empty()) {
        TimedPair t;
        if (hipEventCreate(&t.a) != hipSuccess) return false;
        if (hipEventCreate(&t.b) != hipSuccess) {
            (void)hipEventDestroy(t.a);
            return false;
        }
        c->event_pool.push_back(t);
    }
    *tp = c->event_pool.back();
    c->event_pool.pop_back();
    return true;
}

// ---------------------------------------------------------------------------
__global__ void k_fill_f32(float* __restrict__ p, size_t n, float v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

int fill_f32(dyn_gpu_ctx* c, float* p, size_t n, float v) {
    if (n == 0) return DYN_OK;
    size_t blocks = std::min<size_t>((n + kBlock - 1) / kBlock, 8192);
    hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)blocks), dim3(kBlock), 0, c->stream,
                       p, n, v);
    return check_launch(c, "fill_f32");
}

// Counts indices outside [0, dim) in one column; used to reject bad input.
__global__ void k_count_oob(const uint32_t* __restrict__ ix, uint64_t n, uint32_t dim,
                            unsigned long long* bad) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long local = 0;
    for (; i < n; i += stride) local += (ix[i] >= dim) ? 1ull : 0ull;
    if (local) atomicAdd(bad, local);
}

void free_factors(dyn_gpu_ctx* c) {
    for (int n = 0; n < kMaxModes; ++n) {
        dfree(c, c->Y[n]);
        dfree(c, c->Yg[n]);
        dfree(c, c->Ygb[n]);
        dfree(c, c->Yhat[n]);
        dfree(c, c->G[n]);
        c->Y[n] = nullptr;
        c->Yg[n] = nullptr;
        c->Ygb[n] = nullptr;
        c->Yhat[n] = nullptr;
        c->G[n] = nullptr;
    }
    dfree(c, c->lambda);
    dfree(c, c->inv_lambda);
    dfree(c, c->L);
    dfree(c, c->fitbuf);
    c->lambda = c->inv_lambda = c->L = c->fitbuf = nullptr;
    c->R = c->Rp = 0;
    c->gather_valid = false;
}

void free_copy(dyn_gpu_ctx* c, int m) {
    ModeCopy& cp = c->copy[m];
    dfree(c, cp.vals);
    for (int w = 0; w < kMaxModes; ++w) dfree(c, cp.idx[w]);
    dfree(c, cp.rowptr);
    dfree(c, cp.perm);
    dfree(c, cp.seg_k);
    dfree(c, cp.seg_e);
    dfree(c, cp.seg_row);
    dfree(c, cp.pc_k);
    dfree(c, cp.pc_e);
    dfree(c, cp.pc_row);
    dfree(c, cp.pc_hot);
    dfree(c, cp.heavy_row);
    dfree(c, cp.heavy_first);
    dfree(c, cp.heavy_count);
    dfree(c, cp.empty_rows);
    cp = ModeCopy();
}

void free_relabel(dyn_gpu_ctx* c) {
    for (int n = 0; n < kMaxModes; ++n) {
        dfree(c, c->relab[n]);
        c->relab[n] = nullptr;
        c->relab_inv_host[n].clear();
        c->rank_count_host[n].clear();
    }
    c->relabeled = false;
    c->gather_valid = false;
}

void free_tensor(dyn_gpu_ctx* c) {
    if (c->streamed) {
        // vals / idx alias the streamer's chunk buffers only inside a pass
        c->vals = nullptr;
        for (int n = 0; n < kMaxModes; ++n) c->idx[n] = nullptr;
    }
    for (int m = 0; m < kMaxModes; ++m) free_copy(c, m);
    free_linear(c);
    c->use_linear = false;
    free_relabel(c);
    dfree(c, c->vals);
    c->vals = nullptr;
    for (int n = 0; n < kMaxModes; ++n) {
        dfree(c, c->idx[n]);
        c->idx[n] = nullptr;
        c->dims[n] = 0;
    }
    c->N = 0;
    c->nnz = 0;
    c->tnorm2 = -1.0;
    stream_release(c);
    c->streamed = false;
}

}  // namespace dyn

using namespace dyn;

// ===========================================================================
//  C-ABI: context
// ===========================================================================
extern "C" int dyn_gpu_abi_version(void) { return DYN_GPU_ABI_VERSION; }

extern "C" int dyn_gpu_create(int device, dyn_gpu_ctx** out) {
    if (!out) return DYN_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return DYN_ERR_HIP;
    }
    if (device < 0 || device >= count) return DYN_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return DYN_ERR_HIP;
    dyn_gpu_ctx* c = new dyn_gpu_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return DYN_ERR_HIP;
    }
    c->stream = c->own_stream;
    if (hipStreamCreateWithFlags(&c->aux_stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipStreamDestroy(c->own_stream);
        delete c;
        return DYN_ERR_HIP;
    }
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
            if (prop.multiProcessorCount > 0) c->num_cus = prop.multiProcessorCount;
            if (prop.sharedMemPerBlock > 0) c->max_lds = prop.sharedMemPerBlock;
        }
    }
    if (const char* cap = std::getenv("DYN_GPU_MEM_CAP_MB")) {
        long long mb = std::atoll(cap);
        if (mb > 0) c->mem_cap = (uint64_t)mb << 20;
    }
    // the fixed partials array of this device (dyn_gpu_mttkrp_noalloc); on
    // failure that entry point reports DYN_ERR_STATE, the rest is unaffected
    if (fixed::address(device, &c->fixed_scratch) != DYN_OK) c->fixed_scratch = nullptr;
    if (hipHostMalloc((void**)&c->fitbuf_host, 16 * sizeof(double), hipHostMallocDefault) !=
        hipSuccess) {
        (void)hipStreamDestroy(c->aux_stream);
        (void)hipStreamDestroy(c->own_stream);
        delete c;
        return DYN_ERR_HIP;
    }
    *out = c;
    return DYN_OK;
}

extern "C" int dyn_gpu_destroy(dyn_gpu_ctx* c) {
    if (!c) return DYN_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_tensor(c);
    free_factors(c);
    dfree(c, c->scratch);
    dfree(c, c->scratch2);
    for (auto& t : c->timed) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    for (auto& t : c->event_pool) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    c->cache_allocs = false;
    flush_alloc_cache(c);
    for (auto& kv : c->allocs) (void)hipFree(kv.first);
    (void)hipHostFree(c->fitbuf_host);
    (void)hipStreamSynchronize(c->aux_stream);
    (void)hipStreamDestroy(c->aux_stream);
    (void)hipStreamDestroy(c->own_stream);
    delete c;
    return DYN_OK;
}

extern "C" const char* dyn_gpu_last_error(const dyn_gpu_ctx* c) {
    return c ? c->err.c_str() : "null context";
}

extern "C" int dyn_gpu_set_stream(dyn_gpu_ctx* c, void* stream) {
    if (!c) return DYN_ERR_INVALID;
    c->stream = (hipStream_t)stream;   // NULL: the device's default stream
    return DYN_OK;
}

extern "C" int dyn_gpu_reset_stream(dyn_gpu_ctx* c) {
    if (!c) return DYN_ERR_INVALID;
    c->stream = c->own_stream;
    return DYN_OK;
}

extern "C" int dyn_gpu_synchronize(dyn_gpu_ctx* c) {
    if (!c) return DYN_ERR_INVALID;
    DYN_HIP(c, hipStreamSynchronize(c->stream));
    return DYN_OK;
}

extern "C" int dyn_gpu_device_memory(int device, uint64_t* free_bytes, uint64_t* total_bytes) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return DYN_ERR_HIP;
    }
    if (device < 0 || device >= count) return DYN_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return DYN_ERR_HIP;
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) != hipSuccess) return DYN_ERR_HIP;
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return DYN_OK;
}

extern "C" int dyn_gpu_mem_used(const dyn_gpu_ctx* c, uint64_t* bytes) {
    if (!c || !bytes) return DYN_ERR_INVALID;
    *bytes = c->mem_used;
    return DYN_OK;
}

// ===========================================================================
//  C-ABI: tensor upload
// ===========================================================================
extern "C" int dyn_gpu_upload_coo(dyn_gpu_ctx* c, int N, uint64_t nnz,
                                  const uint32_t* dims, const float* vals,
                                  const uint32_t* const* idx) {
    if (!c) return DYN_ERR_INVALID;
    if (N < 2 || N > kMaxModes)
        return set_err(c, DYN_ERR_INVALID, "num_modes must be in [2, %d], got %d",
                       kMaxModes, N);
    if (nnz >= 0xFFFFFFFFull)
        return set_err(c, DYN_ERR_INVALID, "nnz must be < 2^32 - 1, got %llu",
                       (unsigned long long)nnz);
    if (!dims) return set_err(c, DYN_ERR_INVALID, "dims is null");
    if (nnz > 0 && (!vals || !idx))
        return set_err(c, DYN_ERR_INVALID, "vals/idx is null");
    for (int n = 0; n < N; ++n) {
        if (dims[n] == 0)
            return set_err(c, DYN_ERR_INVALID, "mode %d has size 0", n);
        if (nnz > 0 && !idx[n])
            return set_err(c, DYN_ERR_INVALID, "idx[%d] is null", n);
    }
    DYN_HIP(c, hipSetDevice(c->device));
    DYN_HIP(c, hipStreamSynchronize(c->stream));
    free_tensor(c);
    free_factors(c);

    const size_t cnt = nnz ? nnz : 1;
    DYN_CHECK(dalloc_t(c, &c->vals, cnt));
    for (int n = 0; n < N; ++n) DYN_CHECK(dalloc_t(c, &c->idx[n], cnt));
    if (nnz) {
        DYN_HIP(c, hipMemcpyAsync(c->vals, vals, nnz * sizeof(float),
                                  hipMemcpyHostToDevice, c->stream));
        for (int n = 0; n < N; ++n)
            DYN_HIP(c, hipMemcpyAsync(c->idx[n], idx[n], nnz * sizeof(uint32_t),
                                      hipMemcpyHostToDevice, c->stream));
    }
    c->N = N;
    c->nnz = nnz;
    for (int n = 0; n < N; ++n) c->dims[n] = dims[n];

    if (nnz) {
        unsigned long long* bad = nullptr;
        void* sp = nullptr;
        DYN_CHECK(scratch(c, sizeof(unsigned long long) * kMaxModes, &sp));
        bad = (unsigned long long*)sp;
        DYN_HIP(c, hipMemsetAsync(bad, 0, sizeof(unsigned long long) * kMaxModes, c->stream));
        unsigned blocks = (unsigned)std::min<uint64_t>((nnz + kBlock - 1) / kBlock, 4096);
        for (int n = 0; n < N; ++n)
            hipLaunchKernelGGL(k_count_oob, dim3(blocks), dim3(kBlock), 0, c->stream,
                               c->idx[n], nnz, dims[n], bad + n);
        DYN_CHECK(check_launch(c, "count_oob"));
        unsigned long long hb[kMaxModes] = {0};
        DYN_HIP(c, hipMemcpyAsync(hb, bad, sizeof(hb), hipMemcpyDeviceToHost, c->stream));
        DYN_HIP(c, hipStreamSynchronize(c->stream));
        for (int n = 0; n < N; ++n) {
            if (hb[n]) {
                int code = set_err(c, DYN_ERR_INVALID,
                                   "%llu indices of mode %d are >= its size %u",
                                   hb[n], n, dims[n]);
                free_tensor(c);
                return code;
            }
        }
    }
    return DYN_OK;
}

extern "C" int dyn_gpu_remap(dyn_gpu_ctx* c, int mode, int secondary,
                             uint64_t* rowptr_out, uint32_t* perm_out) {
    if (!c) return DYN_ERR_INVALID;
    if (c->N == 0) return set_err(c, DYN_ERR_STATE, "no tensor uploaded");
    DYN_CHECK(reject_streamed(c, "dyn_gpu_remap"));
    if (mode < 0 || mode >= c->N)
        return set_err(c, DYN_ERR_INVALID, "mode %d out of range", mode);
    if (secondary >= c->N || secondary == mode)
        return set_err(c, DYN_ERR_INVALID, "bad secondary mode %d", secondary);
    DYN_HIP(c, hipSetDevice(c->device));
    // the reference's order (secondary by input index), for comparison
    DYN_CHECK(remap_mode(c, mode, secondary, false));
    const ModeCopy& cp = c->copy[mode];
    if (rowptr_out) {
        std::vector<uint32_t> tmp(c->dims[mode] + 1);
        DYN_HIP(c, hipMemcpyAsync(tmp.data(), cp.rowptr, tmp.size() * sizeof(uint32_t),
                                  hipMemcpyDeviceToHost, c->stream));
        DYN_HIP(c, hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < tmp.size(); ++i) rowptr_out[i] = tmp[i];
    }
    if (perm_out && c->nnz) {
        DYN_HIP(c, hipMemcpyAsync(perm_out, cp.perm, c->nnz * sizeof(uint32_t),
                                  hipMemcpyDeviceToHost, c->stream));
    }
    DYN_HIP(c, hipStreamSynchronize(c->stream));
    return DYN_OK;
}

extern "C" int dyn_gpu_remap_all(dyn_gpu_ctx* c) {
    if (!c) return DYN_ERR_INVALID;
    if (c->N == 0) return set_err(c, DYN_ERR_STATE, "no tensor uploaded");
    DYN_CHECK(reject_streamed(c, "dyn_gpu_remap_all"));
    DYN_HIP(c, hipSetDevice(c->device));
    for (int m = 0; m < c->N; ++m) {
        if (c->copy[m].built) continue;
        DYN_CHECK(remap_mode(c, m, m == 0 ? 1 : 0));
    }
    DYN_HIP(c, hipStreamSynchronize(c->stream));
    return DYN_OK;
}

extern "C" int dyn_gpu_get_copy(dyn_gpu_ctx* c, int mode, float* vals_out,
                                uint32_t* const* idx_out) {
    if (!c) return DYN_ERR_INVALID;
    if (mode < 0 || mode >= c->N) return set_err(c, DYN_ERR_INVALID, "bad mode %d", mode);
    DYN_CHECK(reject_streamed(c, "dyn_gpu_get_copy"));
    const ModeCopy& cp = c->copy[mode];
    if (!cp.built) return set_err(c, DYN_ERR_STATE, "mode %d copy not built", mode);
    if (c->nnz == 0) return DYN_OK;
    if (vals_out)
        DYN_HIP(c, hipMemcpyAsync(vals_out, cp.vals, c->nnz * sizeof(float),
                                  hipMemcpyDeviceToHost, c->stream));
    if (idx_out) {
        for (int w = 0; w < c->N; ++w) {
            if (w == mode || !idx_out[w]) continue;
            DYN_HIP(c, hipMemcpyAsync(idx_out[w], cp.idx[w], c->nnz * sizeof(uint32_t),
                                      hipMemcpyDeviceToHost, c->stream));
        }
    }
    DYN_HIP(c, hipStreamSynchronize(c->stream));
    if (idx_out) {
        // stored columns hold frequency ranks: back to the input indices
        for (int w = 0; w < c->N; ++w) {
            if (w == mode || !idx_out[w]) continue;
            const std::vector<uint32_t>& inv = c->relab_inv_host[w];
            for (uint64_t k = 0; k < c->nnz; ++k) idx_out[w][k] = inv[idx_out[w][k]];
        }
    }
    return DYN_OK;
}

// ===========================================================================
//  C-ABI: factors
// ===========================================================================
extern "C" int dyn_gpu_set_factors(dyn_gpu_ctx* c, int rank, const float* const* Yh,
                                   int ld) {
    if (!c) return DYN_ERR_INVALID;
    if (c->N == 0) return set_err(c, DYN_ERR_STATE, "no tensor uploaded");
    if (rank <= 0) return set_err(c, DYN_ERR_INVALID, "rank must be >= 1, got %d", rank);
    if (padded_rank(rank) > kMaxRp)
        return set_err(c, DYN_ERR_UNSUPPORTED, "rank %d exceeds the supported maximum %d",
                       rank, kMaxRp);
    if (Yh && ld < rank) return set_err(c, DYN_ERR_INVALID, "ld %d < rank %d", ld, rank);
    DYN_HIP(c, hipSetDevice(c->device));
    DYN_HIP(c, hipStreamSynchronize(c->stream));
    const int Rp = padded_rank(rank);
    if (c->R != rank || c->Rp != Rp || !c->Y[0]) {
        free_factors(c);
        for (int n = 0; n < c->N; ++n) {
            const size_t cnt = (size_t)c->dims[n] * Rp;
            DYN_CHECK(dalloc_t(c, &c->Y[n], cnt));
            DYN_CHECK(dalloc_t(c, &c->Yg[n], cnt));
            DYN_CHECK(dalloc_t(c, &c->Ygb[n], cnt));
            DYN_CHECK(dalloc_t(c, &c->Yhat[n], cnt));
            DYN_CHECK(dalloc_t(c, &c->G[n], (size_t)Rp * Rp));
        }
        DYN_CHECK(dalloc_t(c, &c->lambda, Rp));
        DYN_CHECK(dalloc_t(c, &c->inv_lambda, Rp));
        DYN_CHECK(dalloc_t(c, &c->L, (size_t)Rp * Rp));
        DYN_CHECK(dalloc_t(c, &c->fitbuf, 16));
        c->R = rank;
        c->Rp = Rp;
    }
    std::vector<float> pad;
    for (int n = 0; n < c->N; ++n) {
        const size_t rows = c->dims[n];
        if (!Yh || !Yh[n]) {
            DYN_CHECK(fill_f32(c, c->Y[n], rows * Rp, 0.0f));
            continue;
        }
        pad.assign(rows * Rp, 0.0f);
        for (size_t i = 0; i < rows; ++i)
            std::memcpy(&pad[i * Rp], Yh[n] + i * (size_t)ld, sizeof(float) * rank);
        DYN_HIP(c, hipMemcpyAsync(c->Y[n], pad.data(), pad.size() * sizeof(float),
                                  hipMemcpyHostToDevice, c->stream));
        DYN_HIP(c, hipStreamSynchronize(c->stream));
    }
    {
        std::vector<double> ones(Rp, 0.0);
        for (int r = 0; r < rank; ++r) ones[r] = 1.0;
        DYN_HIP(c, hipMemcpyAsync(c->lambda, ones.data(), Rp * sizeof(double),
                                  hipMemcpyHostToDevice, c->stream));
        DYN_HIP(c, hipMemcpyAsync(c->inv_lambda, ones.data(), Rp * sizeof(double),
                                  hipMemcpyHostToDevice, c->stream));
    }
    c->gather_valid = false;
    DYN_HIP(c, hipStreamSynchronize(c->stream));
    return DYN_OK;
}

extern "C" int dyn_gpu_get_factors(dyn_gpu_ctx* c, int mode, float* Yo, int ld) {
    if (!c || !Yo) return DYN_ERR_INVALID;
    if (!c->Y[0]) return set_err(c, DYN_ERR_STATE, "factors not set");
    if (mode < 0 || mode >= c->N) return set_err(c, DYN_ERR_INVALID, "bad mode %d", mode);
    if (ld < c->R) return set_err(c, DYN_ERR_INVALID, "ld %d < rank %d", ld, c->R);
    const size_t rows = c->dims[mode];
    std::vector<float> pad(rows * c->Rp);
    DYN_HIP(c, hipMemcpyAsync(pad.data(), c->Y[mode], pad.size() * sizeof(float),
                              hipMemcpyDeviceToHost, c->stream));
    DYN_HIP(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < rows; ++i)
        std::memcpy(Yo + i * (size_t)ld, &pad[i * c->Rp], sizeof(float) * c->R);
    return DYN_OK;
}

extern "C" int dyn_gpu_get_lambda(dyn_gpu_ctx* c, double* lam) {
    if (!c || !lam) return DYN_ERR_INVALID;
    if (!c->lambda) return set_err(c, DYN_ERR_STATE, "factors not set");
    DYN_HIP(c, hipMemcpyAsync(lam, c->lambda, c->R * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
    DYN_HIP(c, hipStreamSynchronize(c->stream));
    return DYN_OK;
}

// ===========================================================================
//  C-ABI: spMTTKRP and Gram
// ===========================================================================
static int copy_out_rows(dyn_gpu_ctx* c, const float* dsrc, size_t rows, int Rp, int R,
                         float* host, int ld) {
    std::vector<float> pad(rows * Rp);
    DYN_HIP(c, hipMemcpyAsync(pad.data(), dsrc, pad.size() * sizeof(float),
                              hipMemcpyDeviceToHost, c->stream));
    DYN_HIP(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < rows; ++i)
        std::memcpy(host + i * (size_t)ld, &pad[i * Rp], sizeof(float) * R);
    return DYN_OK;
}

static int gathered_operands(dyn_gpu_ctx* c, int prec, const void** Yin) {
    DYN_CHECK(ensure_gather(c));
    for (int w = 0; w < c->N; ++w)
        Yin[w] = (prec == DYN_PREC_BF16) ? (const void*)c->Ygb[w] : (const void*)c->Yg[w];
    return DYN_OK;
}

extern "C" int dyn_gpu_mttkrp(dyn_gpu_ctx* c, int mode, int prec, float* out, int ld) {
    if (!c) return DYN_ERR_INVALID;
    if (c->N == 0) return set_err(c, DYN_ERR_STATE, "no tensor uploaded");
    if (!c->Y[0]) return set_err(c, DYN_ERR_STATE, "factors not set");
    if (mode < 0 || mode >= c->N) return set_err(c, DYN_ERR_INVALID, "bad mode %d", mode);
    if (prec != DYN_PREC_FP32 && prec != DYN_PREC_BF16)
        return set_err(c, DYN_ERR_INVALID, "bad precision %d", prec);
    if (out && ld < c->R) return set_err(c, DYN_ERR_INVALID, "ld %d < rank %d", ld, c->R);
    DYN_HIP(c, hipSetDevice(c->device));
    if (c->streamed) {
        int precs[kMaxModes];
        for (int m = 0; m < c->N; ++m) precs[m] = prec;
        DYN_CHECK(stream_sweep(c, 1u << mode, precs));
    } else if (c->use_linear) {
        DYN_CHECK(linear_sweep(c, 1u << mode));
    } else {
        if (!c->copy[mode].built) DYN_CHECK(remap_mode(c, mode, mode == 0 ? 1 : 0));
        const void* Yin[kMaxModes];
        DYN_CHECK(gathered_operands(c, prec, Yin));
        DYN_CHECK(mttkrp_mode(c, mode, prec, Yin, nullptr, c->Yhat[mode], c->Rp));
    }
    if (out) DYN_CHECK(copy_out_rows(c, c->Yhat[mode], c->dims[mode], c->Rp, c->R, out, ld));
    return DYN_OK;
}

extern "C" int dyn_gpu_mttkrp_noalloc(dyn_gpu_ctx* c, int mode, int prec, float* Yhat_dev,
                                      int ld) {
    if (!c || !Yhat_dev) return DYN_ERR_INVALID;
    if (c->N == 0) return set_err(c, DYN_ERR_STATE, "no tensor uploaded");
    DYN_CHECK(reject_streamed(c, "dyn_gpu_mttkrp_noalloc"));
    if (!c->Y[0]) return set_err(c, DYN_ERR_STATE, "factors not set");
    if (mode < 0 || mode >= c->N) return set_err(c, DYN_ERR_INVALID, "bad mode %d", mode);
    if (prec != DYN_PREC_FP32 && prec != DYN_PREC_BF16)
        return set_err(c, DYN_ERR_INVALID, "bad precision %d", prec);
    if (ld != c->Rp)
        return set_err(c, DYN_ERR_INVALID, "ld must equal the padded rank %d (got %d)", c->Rp, ld);
    if (!c->fixed_scratch)
        return set_err(c, DYN_ERR_STATE, "no fixed scratch address on device %d", c->device);
    DYN_HIP(c, hipSetDevice(c->device));
    // everything this call needs must exist already: building or re-sorting
    // the copy, or the frequency relabel, would allocate
    if (!c->copy[mode].built || !c->relabeled)
        return set_err(c, DYN_ERR_STATE, "mode %d copy not built (call dyn_gpu_remap first)", mode);
    if (!split_ready(c, mode, c->Rp))
        return set_err(c, DYN_ERR_STATE,
                       "mode %d copy is not in launch order (run dyn_gpu_mttkrp on it once)", mode);
    const size_t need = seg_part_bytes(c, mode, c->Rp);
    if (need > fixed::kBytes)
        return set_err(c, DYN_ERR_UNSUPPORTED,
                       "mode %d needs %zu bytes of partials; the fixed array holds %zu", mode, need,
                       fixed::kBytes);
    const void* Yin[kMaxModes];
    DYN_CHECK(gathered_operands(c, prec, Yin));
    return mttkrp_mode(c, mode, prec, Yin, nullptr, Yhat_dev, c->Rp, (float*)c->fixed_scratch);
}

extern "C" int dyn_gpu_mttkrp_all(dyn_gpu_ctx* c, int prec, float* const* out, int ld) {
    if (!c) return DYN_ERR_INVALID;
    if (c->N == 0) return set_err(c, DYN_ERR_STATE, "no tensor uploaded");
    if (!c->Y[0]) return set_err(c, DYN_ERR_STATE, "factors not set");
    if (prec != DYN_PREC_FP32 && prec != DYN_PREC_BF16)
        return set_err(c, DYN_ERR_INVALID, "bad precision %d", prec);
    if (out && ld < c->R) return set_err(c, DYN_ERR_INVALID, "ld %d < rank %d", ld, c->R);
    DYN_HIP(c, hipSetDevice(c->device));
    if (c->streamed) {
        int precs[kMaxModes];
        for (int m = 0; m < c->N; ++m) precs[m] = prec;
        DYN_CHECK(stream_sweep(c, (1u << c->N) - 1u, precs));
    } else if (c->use_linear) {
        DYN_CHECK(linear_sweep(c, (1u << c->N) - 1u));
    } else {
        for (int m = 0; m < c->N; ++m)
            if (!c->copy[m].built) DYN_CHECK(remap_mode(c, m, m == 0 ? 1 : 0));
        const void* Yin[kMaxModes];
        DYN_CHECK(gathered_operands(c, prec, Yin));
        // one snapshot of the factors feeds every mode (Jacobi semantics)
        for (int m = 0; m < c->N; ++m)
            DYN_CHECK(mttkrp_mode(c, m, prec, Yin, nullptr, c->Yhat[m], c->Rp));
    }
    if (out)
        for (int m = 0; m < c->N; ++m)
            if (out[m])
                DYN_CHECK(copy_out_rows(c, c->Yhat[m], c->dims[m], c->Rp, c->R, out[m], ld));
    DYN_HIP(c, hipStreamSynchronize(c->stream));
    return DYN_OK;
}

extern "C" int dyn_gpu_gram(dyn_gpu_ctx* c, int mode, double* Gout) {
    if (!c || !Gout) return DYN_ERR_INVALID;
    if (!c->Y[0]) return set_err(c, DYN_ERR_STATE, "factors not set");
    if (mode < 0 || mode >= c->N) return set_err(c, DYN_ERR_INVALID, "bad mode %d", mode);
    DYN_HIP(c, hipSetDevice(c->device));
    DYN_CHECK(gram_mode(c, mode));
    std::vector<double> g((size_t)c->Rp * c->Rp);
    DYN_HIP(c, hipMemcpyAsync(g.data(), c->G[mode], g.size() * sizeof(double),
                              hipMemcpyDeviceToHost, c->stream));
    DYN_HIP(c, hipStreamSynchronize(c->stream));
    for (int r = 0; r < c->R; ++r)
        for (int q = 0; q < c->R; ++q) Gout[(size_t)r * c->R + q] = g[(size_t)r * c->Rp + q];
    return DYN_OK;
}

// ===========================================================================
//  C-ABI: timing
// ===========================================================================
extern "C" int dyn_gpu_enable_timing(dyn_gpu_ctx* c, int on) {
    if (!c) return DYN_ERR_INVALID;
    c->timing = on != 0;
    return DYN_OK;
}

extern "C" int dyn_gpu_kernel_stats(dyn_gpu_ctx* c, uint64_t* launches, double* total_ms) {
    if (!c) return DYN_ERR_INVALID;
    DYN_HIP(c, hipStreamSynchronize(c->stream));
    double tot = 0.0;
    for (auto& t : c->timed) {
        float ms = 0.f;
        DYN_HIP(c, hipEventElapsedTime(&ms, t.a, t.b));
        tot += ms;
        c->event_pool.push_back(t);
    }
    if (launches) *launches = c->timed.size() + c->timed_extra_n;
    if (total_ms) *total_ms = tot + c->timed_extra_ms;
    c->timed.clear();
    c->timed_extra_ms = 0.0;
    c->timed_extra_n = 0;
    return DYN_OK;
}
