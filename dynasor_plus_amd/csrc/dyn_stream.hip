// Out-of-HBM tensors: the COO tensor stays in host memory -- the caller's
// arrays or a .dnb file -- and is streamed through device chunk buffers once
// per all-modes sweep: the device counterpart of the reference's out-of-core
// Jacobi driver (include/dynasor_ooc.h:25-61, src/dynasor_ooc.cpp:340-535),
// which streams .dnb chunks from disk through RAM.
//
// Pipeline.  Two device chunk buffers.  A loader thread, alive while the
// tensor is attached, walks the chunks cyclically (0, 1, ..., K-1, 0, 1, ...;
// step g holds chunk g % K in buffer g % 2) and brings each into its buffer:
// column by column, up to 16 host threads copy (host arrays) or pread (.dnb
// file) the column's range into pinned staging, and its DMA on a copy stream
// runs while the next column is filled -- all while the compute stream works
// on the previous step in the other buffer -- so the first chunks of the next sweep upload during the
// end of the current one, and a chunk still in its buffer (K <= 2) is not
// uploaded again.  Events order the two streams (a buffer is refilled only
// after the compute stream is done with the step that used it, and a step is
// used only after its upload); a mutex / condition variable orders the two
// host threads' event records.
//
// Per chunk and mode the chunk is sorted by that mode on the device (the
// in-core remap without the secondary key: radix sort, copy, row pointer,
// work tables), its segmented
// spMTTKRP is written to a scratch row block, and the block is added into
// Yhat_n; chunks are added in order, so a sweep is deterministic.  Device
// blocks freed inside a pass are parked and handed out again (no hipFree,
// which waits for the whole device, between chunks).  The frequency relabel
// (for the LDS hot rows) and ||T||^2 come from one pass over the chunks at
// attach time, which also checks every index.
#include "dyn_internal.hpp"

#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace dyn {

struct Streamer {
    // source: caller arrays, or a .dnb file (fd >= 0)
    const float*    host_vals = nullptr;
    const uint32_t* host_idx[kMaxModes] = {nullptr};
    int             fd = -1;
    uint64_t        total = 0;        // nonzeros of the tensor
    uint64_t        chunk = 0;        // nonzeros per chunk
    // device chunk buffers, pinned staging (file source), copy stream, events
    float*          dvals[2] = {nullptr, nullptr};
    uint32_t*       didx[2][kMaxModes] = {{nullptr}, {nullptr}};
    unsigned char*  pin[2] = {nullptr, nullptr};
    hipStream_t     copy = nullptr;
    hipEvent_t      uploaded[2] = {nullptr, nullptr};   // chunk in buffer j is resident
    hipEvent_t      released[2] = {nullptr, nullptr};   // compute is done with buffer j
    float*          tmp = nullptr;    // one mode's chunk spMTTKRP
    size_t          tmp_elems = 0;
    // loader state (steps count from 0 at each loader start)
    std::thread             th;
    std::mutex              mu;
    std::condition_variable cv;
    uint64_t                loaded = 0;     // steps whose upload is enqueued
    uint64_t                consumed = 0;   // steps whose work is enqueued
    uint64_t                resident[2] = {~0ull, ~0ull};   // chunk held by buffer j
    size_t                  fill_threads = 1;   // host threads filling a staging column
    bool                    stop = false;
    int                     err = DYN_OK;
    std::string             msg;
};

namespace {

constexpr uint64_t kDnbHeaderBytes = 88;   // src/dynasor_io.cpp:747-774
constexpr uint64_t kDefaultChunk = 1ull << 27;

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

uint64_t num_chunks(const Streamer& s) { return (s.total + s.chunk - 1) / s.chunk; }

// all `bytes` of the file at `off` into dst (pread may return short counts)
bool pread_all(int fd, unsigned char* dst, size_t bytes, uint64_t off) {
    while (bytes) {
        const ssize_t got = ::pread(fd, dst, std::min<size_t>(bytes, size_t(1) << 30), (off_t)off);
        if (got <= 0) return false;
        dst += got;
        bytes -= (size_t)got;
        off += (uint64_t)got;
    }
    return true;
}

// Loader thread body: step g uploads chunk g % K into buffer g % 2 once the
// compute stream has enqueued step g - 2's work (which recorded
// released[g % 2]), until stopped.
void loader(dyn_gpu_ctx* c, int N) {
    Streamer& s = *c->streamer;
    auto fail = [&](const char* what, hipError_t e) {
        std::lock_guard<std::mutex> lk(s.mu);
        if (s.err == DYN_OK) {
            s.err = e == hipSuccess ? DYN_ERR_INVALID : DYN_ERR_HIP;
            s.msg = what;
            if (e != hipSuccess) s.msg += std::string(": ") + hipGetErrorString(e);
        }
        s.cv.notify_all();
    };
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return fail("loader: hipSetDevice", e);
    const uint64_t K = num_chunks(s);
    for (uint64_t g = 0;; ++g) {
        const int j = (int)(g & 1);
        const uint64_t b = g % K;
        {
            std::unique_lock<std::mutex> lk(s.mu);
            s.cv.wait(lk, [&] { return s.stop || s.consumed + 2 > g; });
            if (s.stop) return;
        }
        if (s.resident[j] != b) {
            const uint64_t off = b * s.chunk;
            const uint64_t nb = std::min<uint64_t>(s.chunk, s.total - off);
            s.resident[j] = ~0ull;
            if ((e = hipStreamWaitEvent(s.copy, s.released[j], 0)) != hipSuccess)
                return fail("loader: wait for a free chunk buffer", e);
            // staging j was last read by the previous upload into buffer j
            if ((e = hipEventSynchronize(s.uploaded[j])) != hipSuccess)
                return fail("loader: staging reuse", e);
            // column by column (values, then each mode's indices): fill the
            // staging range with s.fill_threads host threads, then DMA it
            // while the next column is filled
            const size_t cb = (size_t)nb * 4;
            for (int col = 0; col <= N && e == hipSuccess; ++col) {
                unsigned char* dst = s.pin[j] + (size_t)col * s.chunk * 4;
                const unsigned char* src =
                    s.fd >= 0 ? nullptr
                              : reinterpret_cast<const unsigned char*>(
                                    col == 0 ? (const void*)(s.host_vals + off)
                                             : (const void*)(s.host_idx[col - 1] + off));
                const uint64_t at = kDnbHeaderBytes + ((uint64_t)col * s.total + off) * 4;
                const size_t parts = std::max<size_t>(
                    1, std::min<size_t>(s.fill_threads, cb / (size_t(16) << 20)));
                const size_t step = (cb / parts + 4095) & ~size_t(4095);
                std::vector<std::thread> workers;
                std::vector<char> ok(parts, 1);
                for (size_t p = 0; p < parts; ++p)
                    workers.emplace_back([&, p] {
                        const size_t lo = std::min(cb, p * step), hi = std::min(cb, lo + step);
                        if (src)
                            std::memcpy(dst + lo, src + lo, hi - lo);
                        else
                            ok[p] = pread_all(s.fd, dst + lo, hi - lo, at + lo) ? 1 : 0;
                    });
                for (auto& t : workers) t.join();
                for (size_t p = 0; p < parts; ++p)
                    if (!ok[p]) return fail("loader: short read from the .dnb file", hipSuccess);
                void* dcol = col == 0 ? (void*)s.dvals[j] : (void*)s.didx[j][col - 1];
                e = hipMemcpyAsync(dcol, dst, cb, hipMemcpyHostToDevice, s.copy);
            }
            if (e == hipSuccess) e = hipEventRecord(s.uploaded[j], s.copy);
            if (e != hipSuccess) return fail("loader: chunk upload", e);
            s.resident[j] = b;
        }
        {
            std::lock_guard<std::mutex> lk(s.mu);
            s.loaded = g + 1;
        }
        s.cv.notify_all();
    }
}

void stop_loader(Streamer& s) {
    {
        std::lock_guard<std::mutex> lk(s.mu);
        s.stop = true;
    }
    s.cv.notify_all();
    if (s.th.joinable()) s.th.join();
}

// (Re)starts the loader at step 0 after waiting for every upload it
// enqueued; buffer contents stay valid (resident).
void start_loader(dyn_gpu_ctx* c) {
    Streamer& s = *c->streamer;
    stop_loader(s);
    (void)hipStreamSynchronize(s.copy);
    s.loaded = s.consumed = 0;
    s.stop = false;
    s.err = DYN_OK;
    s.msg.clear();
    s.th = std::thread(loader, c, c->N);
}

// Runs body(b) on every chunk b in order on the compute stream, with the
// chunk bound as c->vals / c->idx / c->nnz, the loader's next uploads
// overlapping.  body may synchronise the compute stream.  On return c->nnz
// is the tensor's count again and vals / idx are unbound.
template <class Body>
int stream_pass(dyn_gpu_ctx* c, Body&& body) {
    Streamer& s = *c->streamer;
    const int N = c->N;
    const uint64_t K = num_chunks(s);
    bool restart;
    {
        std::lock_guard<std::mutex> lk(s.mu);
        // a pass starts at chunk 0: after an error or an interrupted pass the
        // loader restarts
        restart = !s.th.joinable() || s.err != DYN_OK || s.consumed % K != 0;
    }
    if (restart) start_loader(c);
    const uint64_t g0 = s.consumed;
    c->cache_allocs = true;
    int rc = DYN_OK;
    for (uint64_t b = 0; b < K && rc == DYN_OK; ++b) {
        const uint64_t g = g0 + b;
        const int j = (int)(g & 1);
        {
            std::unique_lock<std::mutex> lk(s.mu);
            s.cv.wait(lk, [&] { return s.loaded > g || s.err != DYN_OK; });
            if (s.err != DYN_OK) {
                rc = set_err(c, s.err, "stream pass: %s", s.msg.c_str());
                break;
            }
        }
        hipError_t e = hipStreamWaitEvent(c->stream, s.uploaded[j], 0);
        if (e != hipSuccess) {
            rc = set_err(c, DYN_ERR_HIP, "stream pass: %s", hipGetErrorString(e));
            break;
        }
        c->vals = s.dvals[j];
        for (int w = 0; w < N; ++w) c->idx[w] = s.didx[j][w];
        c->nnz = std::min<uint64_t>(s.chunk, s.total - b * s.chunk);
        rc = body(b);
        // every read of buffer j by this step's work precedes this record
        e = hipEventRecord(s.released[j], c->stream);
        if (rc == DYN_OK && e != hipSuccess)
            rc = set_err(c, DYN_ERR_HIP, "stream pass: %s", hipGetErrorString(e));
        {
            std::lock_guard<std::mutex> lk(s.mu);
            s.consumed = g + 1;
        }
        s.cv.notify_all();
    }
    const hipError_t e2 = hipStreamSynchronize(c->stream);
    if (rc == DYN_OK && e2 != hipSuccess)
        rc = set_err(c, DYN_ERR_HIP, "stream pass: %s", hipGetErrorString(e2));
    c->vals = nullptr;
    for (int w = 0; w < kMaxModes; ++w) c->idx[w] = nullptr;
    c->nnz = s.total;
    c->cache_allocs = false;
    flush_alloc_cache(c);
    return rc;
}

// Buffers, staging, copy stream and events of a streamer over `nnz`
// nonzeros in chunks of `chunk` (the source fields are set by the caller).
int stream_setup(dyn_gpu_ctx* c, int N, uint64_t nnz, uint64_t chunk) {
    Streamer& s = *c->streamer;
    s.total = nnz;
    s.chunk = chunk;
    for (int j = 0; j < 2; ++j) {
        DYN_CHECK(dalloc_t(c, &s.dvals[j], chunk));
        for (int w = 0; w < N; ++w) DYN_CHECK(dalloc_t(c, &s.didx[j][w], chunk));
    }
    s.fill_threads = std::max<size_t>(1, std::min<size_t>(16, std::thread::hardware_concurrency()));
    // pinned staging: one chunk per buffer, filled by host threads, read by DMA
    for (int j = 0; j < 2; ++j) {
        void* p = nullptr;
        const hipError_t e = hipHostMalloc(&p, (size_t)chunk * 4 * (1 + N));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return set_err(c, DYN_ERR_OOM, "pinned staging of %llu bytes: %s",
                           (unsigned long long)chunk * 4 * (1 + N), hipGetErrorString(e));
        }
        s.pin[j] = (unsigned char*)p;
    }
    DYN_HIP(c, hipStreamCreateWithFlags(&s.copy, hipStreamNonBlocking));
    for (int j = 0; j < 2; ++j) {
        DYN_HIP(c, hipEventCreateWithFlags(&s.uploaded[j], hipEventDisableTiming));
        DYN_HIP(c, hipEventCreateWithFlags(&s.released[j], hipEventDisableTiming));
        // recorded once so that the first waits on them pass
        DYN_HIP(c, hipEventRecord(s.uploaded[j], s.copy));
        DYN_HIP(c, hipEventRecord(s.released[j], c->stream));
    }
    return DYN_OK;
}

// The attach pass: index check, per-row counts (the frequency relabel) and
// ||T||^2 summed in chunk order.
int attach_pass(dyn_gpu_ctx* c) {
    const int N = c->N;
    size_t cnt_total = 0;
    for (int n = 0; n < N; ++n) cnt_total += ((size_t)c->dims[n] + 63) & ~size_t(63);
    uint32_t* cnt = nullptr;
    unsigned long long* bad = nullptr;
    int rc = dalloc_t(c, &cnt, cnt_total);
    if (rc == DYN_OK) rc = dalloc_t(c, &bad, (size_t)kMaxModes);
    if (rc == DYN_OK) rc = zero_u32(c, cnt, cnt_total);
    if (rc == DYN_OK) rc = zero_u32(c, (uint32_t*)bad, 2 * (size_t)kMaxModes);
    double tn2 = 0.0;
    if (rc == DYN_OK)
        rc = stream_pass(c, [&](uint64_t) {
            const uint64_t nb = c->nnz;
            size_t off = 0;
            for (int n = 0; n < N; ++n) {
                hipLaunchKernelGGL(k_count_oob_s, dim3(grid_of(nb)), dim3(kBlock), 0, c->stream,
                                   c->idx[n], nb, c->dims[n], bad + n);
                DYN_CHECK(check_launch(c, "count_oob"));
                // out-of-range entries fall outside every histogram bin
                DYN_CHECK(count_rows(c, c->idx[n], nb, c->dims[n], cnt + off));
                off += ((size_t)c->dims[n] + 63) & ~size_t(63);
            }
            c->tnorm2 = -1.0;
            double part = 0.0;
            DYN_CHECK(tensor_norm2(c, &part));
            tn2 += part;
            return DYN_OK;
        });
    unsigned long long hb[kMaxModes] = {0};
    if (rc == DYN_OK) {
        const hipError_t e = hipMemcpy(hb, bad, sizeof(hb), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = set_err(c, DYN_ERR_HIP, "attach: %s", hipGetErrorString(e));
    }
    for (int n = 0; n < N && rc == DYN_OK; ++n)
        if (hb[n])
            rc = set_err(c, DYN_ERR_INVALID, "%llu indices of mode %d are >= its size %u", hb[n],
                         n, c->dims[n]);
    size_t off = 0;
    for (int n = 0; n < N && rc == DYN_OK; ++n) {
        rc = relabel_from_counts(c, n, cnt + off);
        off += ((size_t)c->dims[n] + 63) & ~size_t(63);
    }
    dfree(c, cnt);
    dfree(c, bad);
    if (rc == DYN_OK) {
        c->relabeled = true;
        c->tnorm2 = tn2;
    }
    return rc;
}

int check_shape(dyn_gpu_ctx* c, int N, uint64_t nnz, const uint32_t* dims) {
    if (N < 2 || N > kMaxModes)
        return set_err(c, DYN_ERR_INVALID, "num_modes must be in [2, %d], got %d", kMaxModes, N);
    if (nnz == 0) return set_err(c, DYN_ERR_INVALID, "a streamed tensor needs nonzeros");
    for (int n = 0; n < N; ++n)
        if (dims[n] == 0) return set_err(c, DYN_ERR_INVALID, "mode %d has size 0", n);
    return DYN_OK;
}

uint64_t clamp_chunk(uint64_t chunk_nnz, uint64_t nnz) {
    uint64_t chunk = chunk_nnz ? chunk_nnz : kDefaultChunk;
    return std::min<uint64_t>(std::min<uint64_t>(chunk, nnz), kMaxNnz - 1);
}

// common part of the attach entry points once the source is set
int attach_finish(dyn_gpu_ctx* c, int N, uint64_t nnz, const uint32_t* dims, uint64_t chunk) {
    c->N = N;
    for (int n = 0; n < N; ++n) c->dims[n] = dims[n];
    c->streamed = true;
    c->nnz = nnz;
    int rc = stream_setup(c, N, nnz, chunk);
    if (rc == DYN_OK) rc = attach_pass(c);
    if (rc != DYN_OK) free_tensor(c);
    return rc;
}

}  // namespace

void stream_release(dyn_gpu_ctx* c) {
    Streamer* s = c->streamer;
    if (!s) return;
    stop_loader(*s);
    if (s->copy) (void)hipStreamSynchronize(s->copy);
    (void)hipStreamSynchronize(c->stream);
    for (int j = 0; j < 2; ++j) {
        dfree(c, s->dvals[j]);
        for (int w = 0; w < kMaxModes; ++w) dfree(c, s->didx[j][w]);
        if (s->pin[j]) (void)hipHostFree(s->pin[j]);
        if (s->uploaded[j]) (void)hipEventDestroy(s->uploaded[j]);
        if (s->released[j]) (void)hipEventDestroy(s->released[j]);
    }
    dfree(c, s->tmp);
    if (s->copy) (void)hipStreamDestroy(s->copy);
    if (s->fd >= 0) ::close(s->fd);
    delete s;
    c->streamer = nullptr;
}

int reject_streamed(dyn_gpu_ctx* c, const char* what) {
    if (!c->streamed) return DYN_OK;
    return set_err(c, DYN_ERR_STATE,
                   "%s needs the tensor in device memory; this context streams it from host "
                   "memory (only all-modes spMTTKRP, per-mode spMTTKRP and Jacobi CP-ALS run)",
                   what);
}

int stream_sweep(dyn_gpu_ctx* c, uint32_t mask, const int* prec) {
    Streamer& s = *c->streamer;
    const int Rp = c->Rp;
    size_t need = 0;
    for (int n = 0; n < c->N; ++n) need = std::max<size_t>(need, (size_t)c->dims[n] * Rp);
    if (s.tmp_elems < need) {
        dfree(c, s.tmp);
        s.tmp = nullptr;
        s.tmp_elems = 0;
        DYN_CHECK(dalloc_t(c, &s.tmp, need));
        s.tmp_elems = need;
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
    return stream_pass(c, [&](uint64_t) {
        for (int n = 0; n < c->N; ++n) {
            if (!((mask >> n) & 1u)) continue;
            // no secondary sort: on chunks it does not speed the spMTTKRP
            // and costs a third of the chunk's sort
            int rc = remap_mode(c, n, -1);
            if (rc == DYN_OK)
                rc = mttkrp_mode(c, n, prec[n], prec[n] == DYN_PREC_BF16 ? Yb : Y32, nullptr,
                                 s.tmp, Rp);
            if (rc == DYN_OK) {
                const size_t e = (size_t)c->dims[n] * Rp;
                hipLaunchKernelGGL(k_add_rows, dim3(grid_of(e)), dim3(kBlock), 0, c->stream,
                                   c->Yhat[n], s.tmp, e);
                rc = check_launch(c, "add_rows");
            }
            free_copy(c, n);   // parked: the next chunk's copy reuses the blocks
            if (rc != DYN_OK) return rc;
        }
        return DYN_OK;
    });
}

}  // namespace dyn

using namespace dyn;

extern "C" int dyn_gpu_attach_host_coo(dyn_gpu_ctx* c, int N, uint64_t nnz, const uint32_t* dims,
                                       const float* vals, const uint32_t* const* idx,
                                       uint64_t chunk_nnz) {
    if (!c) return DYN_ERR_INVALID;
    if (!dims || !vals || !idx) return set_err(c, DYN_ERR_INVALID, "dims/vals/idx is null");
    DYN_CHECK(check_shape(c, N, nnz, dims));
    for (int n = 0; n < N; ++n)
        if (!idx[n]) return set_err(c, DYN_ERR_INVALID, "idx[%d] is null", n);
    DYN_HIP(c, hipSetDevice(c->device));
    DYN_HIP(c, hipStreamSynchronize(c->stream));
    free_tensor(c);
    free_factors(c);
    c->streamer = new Streamer();
    c->streamer->host_vals = vals;
    for (int n = 0; n < N; ++n) c->streamer->host_idx[n] = idx[n];
    return attach_finish(c, N, nnz, dims, clamp_chunk(chunk_nnz, nnz));
}

extern "C" int dyn_gpu_attach_dnb(dyn_gpu_ctx* c, const char* path, uint64_t chunk_nnz) {
    if (!c) return DYN_ERR_INVALID;
    if (!path) return set_err(c, DYN_ERR_INVALID, "path is null");
    const int fd = ::open(path, O_RDONLY);
    if (fd < 0) return set_err(c, DYN_ERR_INVALID, "cannot open '%s'", path);
    // header: "DNB1", version 1, value and index widths 4, N, nnz, dims[8]
    unsigned char h[kDnbHeaderBytes];
    uint32_t version = 0, vb = 0, ib = 0, N = 0, dims[8] = {0};
    uint64_t nnz = 0;
    const bool got = pread_all(fd, h, sizeof(h), 0);
    if (got) {
        std::memcpy(&version, h + 4, 4);
        std::memcpy(&vb, h + 8, 4);
        std::memcpy(&ib, h + 12, 4);
        std::memcpy(&N, h + 16, 4);
        std::memcpy(&nnz, h + 20, 8);
        std::memcpy(dims, h + 28, sizeof(dims));
    }
    const off_t size = got ? ::lseek(fd, 0, SEEK_END) : (off_t)-1;
    const char* why = !got                               ? "truncated header"
                      : std::memcmp(h, "DNB1", 4) != 0   ? "not a .dnb file (bad magic)"
                      : version != 1 || vb != 4 || ib != 4 ? "unsupported version or widths"
                      : N < 2 || N > (uint32_t)kMaxModes  ? "mode count out of range"
                      : nnz == 0                          ? "no nonzeros"
                      : size < 0 || (uint64_t)size != kDnbHeaderBytes + nnz * 4ull * (1 + N)
                          ? "file size does not match its header"
                          : nullptr;
    if (why) {
        ::close(fd);
        return set_err(c, DYN_ERR_INVALID, "'%s': %s", path, why);
    }
    if (const int rc = check_shape(c, (int)N, nnz, dims); rc != DYN_OK) {
        ::close(fd);
        return rc;
    }
    const hipError_t e = hipSetDevice(c->device);
    if (e == hipSuccess) (void)hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        ::close(fd);
        return set_err(c, DYN_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(e));
    }
    free_tensor(c);
    free_factors(c);
    c->streamer = new Streamer();
    c->streamer->fd = fd;
    return attach_finish(c, (int)N, nnz, dims, clamp_chunk(chunk_nnz, nnz));
}
