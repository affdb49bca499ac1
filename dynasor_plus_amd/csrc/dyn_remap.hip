// Per-mode remap on the device: stable LSD radix sort of the nonzeros by the
// output-mode coordinate (optionally with a secondary coordinate as the
// minor key), gather of the value / other-index columns into the mode copy
// (other-mode indices replaced by their frequency ranks), the row pointer and
// the row-segment table.
//
// Semantics follow the reference's NCopy + fiber sort + CSR compaction
// (src/dynasor_preprocess.cpp:471-521 counting sort per shard, :1467-1607
// rowptr and compaction): after the remap, rowptr[r]..rowptr[r+1] holds
// exactly the nonzeros whose mode coordinate is r, and the mode's own index
// column is not stored.  The reference orders rows inside power-of-two
// super-shards scheduled over threads; the device copy is one global order,
// so the per-row multiset of nonzeros is the same but there is no shard
// structure.
#include "dyn_internal.hpp"

#include <algorithm>
#include <numeric>
#include <vector>

namespace dyn {

constexpr int kRadixItems = 16;                       // rows of 256 keys per tile
constexpr int kRadixTile = kBlock * kRadixItems;      // 4096 keys per workgroup
constexpr int kScanItems = 16;
constexpr int kScanTile = kBlock * kScanItems;

// ---------------------------------------------------------------------------
//  digit histogram of one tile -> counts[digit * nblocks + block]
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_radix_hist(const uint32_t* __restrict__ keys,
                                                       uint32_t n, int shift,
                                                       uint32_t* __restrict__ counts,
                                                       uint32_t nblocks) {
    __shared__ uint32_t h[256];
    const int tid = threadIdx.x;
    h[tid] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kRadixTile;
#pragma unroll 4
    for (int j = 0; j < kRadixItems; ++j) {
        const uint64_t i = base + (uint64_t)j * kBlock + tid;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    counts[(uint64_t)tid * nblocks + blockIdx.x] = h[tid];
}

// ---------------------------------------------------------------------------
//  exclusive scan of a u32 array (three phases)
// ---------------------------------------------------------------------------
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
    __shared__ uint32_t wsum[kBlock / 64];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    uint32_t woff = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        if (w < wave) woff += wsum[w];
        tot += wsum[w];
    }
    __syncthreads();
    *total = tot;
    return woff + x - v;
}

__global__ __launch_bounds__(kBlock) void k_scan_reduce(const uint32_t* __restrict__ a, uint64_t n,
                                                        uint32_t* __restrict__ bsum) {
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    uint32_t s = 0;
    for (int j = 0; j < kScanItems; ++j) {
        uint64_t i = base + (uint64_t)j * kBlock + threadIdx.x;
        if (i < n) s += a[i];
    }
    uint32_t tot;
    (void)block_exclusive_scan(s, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kBlock) void k_scan_top(uint32_t* __restrict__ bsum, uint32_t nb) {
    // single workgroup: exclusive scan of nb block sums, in chunks of 256
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < nb; c0 += kBlock) {
        uint32_t i = c0 + threadIdx.x;
        uint32_t v = (i < nb) ? bsum[i] : 0u;
        uint32_t tot;
        uint32_t ex = block_exclusive_scan(v, &tot);
        if (i < nb) bsum[i] = carry + ex;
        carry += tot;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlock) void k_scan_down(uint32_t* __restrict__ a, uint64_t n,
                                                      const uint32_t* __restrict__ bsum) {
    // thread t scans kScanItems consecutive elements of the tile
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint32_t loc[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        uint64_t i = base + j;
        loc[j] = (i < n) ? a[i] : 0u;
        s += loc[j];
    }
    uint32_t tot;
    uint32_t off = block_exclusive_scan(s, &tot) + bsum[blockIdx.x];
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        uint64_t i = base + j;
        if (i < n) a[i] = off;
        off += loc[j];
    }
}

static int exclusive_scan(dyn_gpu_ctx* c, uint32_t* a, uint64_t n, uint32_t* bsum_buf) {
    const uint32_t nb = (uint32_t)((n + kScanTile - 1) / kScanTile);
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kBlock), 0, c->stream, a, n, bsum_buf);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kBlock), 0, c->stream, bsum_buf, nb);
    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(kBlock), 0, c->stream, a, n, bsum_buf);
    return check_launch(c, "exclusive_scan");
}

// ---------------------------------------------------------------------------
//  stable scatter of one tile.  Tile order is row-major over (row j, thread):
//  inside a row the rank of a key among equal digits comes from eight wave
//  ballots (bit-sliced match), across waves from per-wave digit counts in
//  LDS, across rows from a running per-digit count.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_radix_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, uint32_t n, int shift,
    const uint32_t* __restrict__ offsets, uint32_t nblocks) {
    __shared__ uint32_t base_off[256];
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[kBlock / 64][256];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    base_off[tid] = offsets[(uint64_t)tid * nblocks + blockIdx.x];
    run[tid] = 0;
    const uint64_t lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    const uint64_t base = (uint64_t)blockIdx.x * kRadixTile;
    for (int j = 0; j < kRadixItems; ++j) {
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) wcnt[w][tid] = 0;
        __syncthreads();
        const uint64_t i = base + (uint64_t)j * kBlock + tid;
        const bool valid = i < n;
        const uint32_t key = valid ? kin[i] : 0u;
        const uint32_t d = (key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
        if (valid && rank == 0) wcnt[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t prior = run[d];
            for (int w = 0; w < wave; ++w) prior += wcnt[w][d];
            const uint32_t pos = base_off[d] + prior + rank;
            kout[pos] = key;
            vout[pos] = vin[i];
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) add += wcnt[w][tid];
        run[tid] += add;
        __syncthreads();
    }
}

__global__ void k_iota(uint32_t* __restrict__ p, uint32_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = (uint32_t)i;
}

__global__ void k_gather_u32(const uint32_t* __restrict__ src, const uint32_t* __restrict__ perm,
                             uint32_t* __restrict__ dst, uint32_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[perm[i]];
}

__global__ void k_gather_f32(const float* __restrict__ src, const uint32_t* __restrict__ perm,
                             float* __restrict__ dst, uint32_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[perm[i]];
}

// rowptr[r] = first sorted position whose key >= r, for r in [0, dim]
__global__ void k_rowptr(const uint32_t* __restrict__ keys, uint32_t n, uint32_t dim,
                         uint32_t* __restrict__ rowptr) {
    // 64-bit positions: a 32-bit one would wrap past 2^32 - stride
    uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; p <= n; p += stride) {
        const uint32_t lo = (p == 0) ? 0u : keys[p - 1] + 1u;   // rows (keys[p-1], ...]
        const uint32_t hi = (p == n) ? dim : keys[p];           // ... keys[p]]
        for (uint64_t r = lo; r <= hi; ++r) rowptr[r] = (uint32_t)p;
    }
}

// dst[p] = relab[src[perm[p]]]
__global__ void k_gather_relabel(const uint32_t* __restrict__ src, const uint32_t* __restrict__ perm,
                                 const uint32_t* __restrict__ relab, uint32_t* __restrict__ dst,
                                 uint32_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = relab[src[perm[i]]];
}

// Histogram of ix restricted to bins [lo, lo + nb): per-workgroup LDS counts
// (contention on hot bins stays inside the CU), then one global add per
// non-empty bin and workgroup.
constexpr int kHistBlock = 1024;
constexpr uint32_t kHistBins = 32768;

__global__ __launch_bounds__(kHistBlock) void k_bincount(const uint32_t* __restrict__ ix,
                                                         uint32_t n, uint32_t lo, uint32_t nb,
                                                         uint32_t* __restrict__ cnt) {
    extern __shared__ uint32_t hist[];
    for (uint32_t b = threadIdx.x; b < nb; b += kHistBlock) hist[b] = 0u;
    __syncthreads();
    for (uint64_t i = (uint64_t)blockIdx.x * kHistBlock + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * kHistBlock) {
        const uint32_t v = ix[i] - lo;
        if (v < nb) atomicAdd(&hist[v], 1u);
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < nb; b += kHistBlock)
        if (hist[b]) atomicAdd(&cnt[lo + b], hist[b]);
}

__global__ void k_zero_u32(uint32_t* __restrict__ p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0u;
}

// dst32[relab[r]] = src[r] (and the BF16 rounding into dstbf), rows x RP
__global__ void k_relabel_rows(const float* __restrict__ src, const uint32_t* __restrict__ relab,
                               float* __restrict__ dst32, bf16* __restrict__ dstbf, size_t rows,
                               int RP) {
    const size_t n = rows * (size_t)RP;
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; e < n; e += stride) {
        const size_t r = e / RP, col = e % RP;
        const size_t o = (size_t)relab[r] * RP + col;
        const float x = src[e];
        if (dst32) dst32[o] = x;
        if (dstbf) dstbf[o] = __float2bfloat16(x);
    }
}

// bad[0] += entries of perm that are out of range or repeated
__global__ void k_perm_check(const uint32_t* __restrict__ perm, uint32_t n,
                             uint32_t* __restrict__ bits, unsigned long long* bad) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long local = 0;
    for (; i < n; i += stride) {
        const uint32_t p = perm[i];
        if (p >= n) {
            ++local;
            continue;
        }
        const uint32_t m = 1u << (p & 31u);
        if (atomicOr(&bits[p >> 5], m) & m) ++local;
    }
    if (local) atomicAdd(bad, local);
}

// *bad += positions k with keys[k] < keys[k - 1]
__global__ void k_count_descents(const uint32_t* __restrict__ keys, uint32_t n,
                                 unsigned long long* bad) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x + 1;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long local = 0;
    for (; i < n; i += stride) local += (keys[i] < keys[i - 1]) ? 1ull : 0ull;
    if (local) atomicAdd(bad, local);
}

static int bits_for(uint32_t dim) {
    int b = 0;
    while (b < 32 && ((uint64_t)1 << b) < dim) ++b;
    return b;
}

unsigned grid_for(uint64_t n) {
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + kBlock - 1) / kBlock, 16384));
}

// LSD passes over `bits` bits of (k0, v0); returns which buffer (0/1) holds
// the result.
static int radix_passes(dyn_gpu_ctx* c, uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1,
                        uint32_t n, int bits, uint32_t* counts, uint32_t* bsum, int* which) {
    const uint32_t nblocks = (n + kRadixTile - 1) / kRadixTile;
    uint32_t* ks[2] = {k0, k1};
    uint32_t* vs[2] = {v0, v1};
    int cur = *which;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3(nblocks), dim3(kBlock), 0, c->stream, ks[cur], n,
                           shift, counts, nblocks);
        DYN_CHECK(check_launch(c, "radix_hist"));
        DYN_CHECK(exclusive_scan(c, counts, (uint64_t)nblocks * 256, bsum));
        hipLaunchKernelGGL(k_radix_scatter, dim3(nblocks), dim3(kBlock), 0, c->stream, ks[cur],
                           vs[cur], ks[cur ^ 1], vs[cur ^ 1], n, shift, counts, nblocks);
        DYN_CHECK(check_launch(c, "radix_scatter"));
        cur ^= 1;
    }
    *which = cur;
    return DYN_OK;
}

int sort_pairs(dyn_gpu_ctx* c, uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, uint32_t n,
               int bits, int* which) {
    if (n == 0 || bits <= 0) return DYN_OK;
    const uint32_t nblocks = (n + kRadixTile - 1) / kRadixTile;
    const uint64_t ncounts = (uint64_t)std::max<uint32_t>(nblocks, 1) * 256;
    void* sp2 = nullptr;
    DYN_CHECK(scratch2(c, (ncounts + ncounts / kScanTile + 16) * sizeof(uint32_t), &sp2));
    uint32_t* counts = (uint32_t*)sp2;
    return radix_passes(c, k0, v0, k1, v1, n, bits, counts, counts + ncounts, which);
}

int iota_u32(dyn_gpu_ctx* c, uint32_t* p, uint32_t n) {
    if (n == 0) return DYN_OK;
    hipLaunchKernelGGL(k_iota, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, p, n);
    return check_launch(c, "iota");
}

int gather_u32(dyn_gpu_ctx* c, const uint32_t* src, const uint32_t* perm, uint32_t* dst,
               uint32_t n) {
    if (n == 0) return DYN_OK;
    hipLaunchKernelGGL(k_gather_u32, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, src, perm,
                       dst, n);
    return check_launch(c, "gather_u32");
}

int gather_f32(dyn_gpu_ctx* c, const float* src, const uint32_t* perm, float* dst, uint32_t n) {
    if (n == 0) return DYN_OK;
    hipLaunchKernelGGL(k_gather_f32, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, src, perm,
                       dst, n);
    return check_launch(c, "gather_f32");
}

int gather_relabel(dyn_gpu_ctx* c, const uint32_t* src, const uint32_t* perm,
                   const uint32_t* relab, uint32_t* dst, uint32_t n) {
    if (n == 0) return DYN_OK;
    hipLaunchKernelGGL(k_gather_relabel, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, src, perm,
                       relab, dst, n);
    return check_launch(c, "gather_relabel");
}

int count_rows(dyn_gpu_ctx* c, const uint32_t* ix, uint64_t n, uint32_t dim, uint32_t* cnt) {
    if (n == 0) return DYN_OK;
    const uint32_t nblk = (uint32_t)std::min<uint64_t>(std::max<uint64_t>((n + 65535) / 65536, 1),
                                                       (uint64_t)c->num_cus * 4);
    static std::atomic<uint64_t> lds_set{0};
    lds_limit_once(c, &lds_set, (const void*)k_bincount, (int)(kHistBins * sizeof(uint32_t)));
    for (uint32_t lo = 0; lo < dim; lo += kHistBins) {
        const uint32_t nb = std::min(kHistBins, dim - lo);
        hipLaunchKernelGGL(k_bincount, dim3(nblk), dim3(kHistBlock), nb * sizeof(uint32_t),
                           c->stream, ix, (uint32_t)n, lo, nb, cnt);
        DYN_CHECK(check_launch(c, "bincount"));
    }
    return DYN_OK;
}

int zero_u32(dyn_gpu_ctx* c, uint32_t* p, size_t n) {
    if (n == 0) return DYN_OK;
    hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, p, n);
    return check_launch(c, "zero_u32");
}

int relabel_from_counts(dyn_gpu_ctx* c, int w, const uint32_t* cnt) {
    const uint32_t dim = c->dims[w];
    if (!c->relab[w]) DYN_CHECK(dalloc_t(c, &c->relab[w], dim));
    std::vector<uint32_t> h(dim);
    DYN_HIP(c, hipMemcpyAsync(h.data(), cnt, (size_t)dim * 4, hipMemcpyDeviceToHost, c->stream));
    DYN_HIP(c, hipStreamSynchronize(c->stream));
    std::vector<uint32_t>& inv = c->relab_inv_host[w];
    inv.resize(dim);
    std::iota(inv.begin(), inv.end(), 0u);
    std::stable_sort(inv.begin(), inv.end(), [&](uint32_t a, uint32_t b) { return h[a] > h[b]; });
    std::vector<uint32_t> rank(dim);
    for (uint32_t r = 0; r < dim; ++r) rank[inv[r]] = r;
    std::vector<uint32_t>& rc = c->rank_count_host[w];
    rc.resize(dim);
    for (uint32_t r = 0; r < dim; ++r) rc[r] = h[inv[r]];
    DYN_HIP(c, hipMemcpyAsync(c->relab[w], rank.data(), (size_t)dim * 4, hipMemcpyHostToDevice,
                              c->stream));
    DYN_HIP(c, hipStreamSynchronize(c->stream));
    return DYN_OK;
}

// Frequency ranks of every mode's indices: rank 0 is the index with the most
// nonzeros (ties by index).  The ranking only orders factor rows for the
// gathers; results do not depend on it.
int ensure_relabel(dyn_gpu_ctx* c) {
    if (c->relabeled) return DYN_OK;
    for (int w = 0; w < c->N; ++w) {
        const uint32_t dim = c->dims[w];
        void* sp = nullptr;
        DYN_CHECK(scratch(c, (size_t)dim * sizeof(uint32_t), &sp));
        uint32_t* cnt = (uint32_t*)sp;
        DYN_CHECK(zero_u32(c, cnt, dim));
        DYN_CHECK(count_rows(c, c->idx[w], c->nnz, dim, cnt));
        DYN_CHECK(relabel_from_counts(c, w, cnt));
    }
    c->relabeled = true;
    return DYN_OK;
}

int relabel_rows(dyn_gpu_ctx* c, const float* src, const uint32_t* relab, float* dst32,
                 bf16* dstbf, size_t rows, int Rp) {
    const size_t n = rows * (size_t)Rp;
    if (n == 0) return DYN_OK;
    hipLaunchKernelGGL(k_relabel_rows, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, src, relab,
                       dst32, dstbf, rows, Rp);
    return check_launch(c, "relabel_rows");
}

int ensure_gather(dyn_gpu_ctx* c) {
    if (c->gather_valid) return DYN_OK;
    DYN_CHECK(ensure_relabel(c));
    for (int w = 0; w < c->N; ++w)
        DYN_CHECK(relabel_rows(c, c->Y[w], c->relab[w], c->Yg[w], c->Ygb[w], c->dims[w], c->Rp));
    c->gather_valid = true;
    return DYN_OK;
}

// Work tables from the row pointer (host): heavy rows cut into kPiece
// pieces, light rows grouped by the kLightWindow window of their first
// nonzero.  With per-row hot counts `nh` (a hot split), a heavy row's pieces
// are cut separately in its hot prefix and its cold rest, each piece records
// which class it holds, and the two classes are interleaved in proportion
// within the row so that every wave's contiguous range of pieces gets about
// the same mix of fast (LDS) and slow (memory) pieces.
static int build_segments(dyn_gpu_ctx* c, ModeCopy& cp, const std::vector<uint32_t>& rp,
                          const std::vector<uint32_t>* nh = nullptr) {
    const uint32_t dim = (uint32_t)rp.size() - 1;
    std::vector<uint32_t> sk, se, sr, pk, pe, pr, ph, hr, hf, hc, er;
    bool open = false;
    uint32_t window = 0;
    auto cut = [&](uint32_t r, uint32_t b, uint32_t e, uint32_t hot) {
        for (uint32_t k = b; k < e; k += kPiece) {
            pk.push_back(k);
            pe.push_back(std::min(k + kPiece, e));
            pr.push_back(r);
            ph.push_back(hot);
        }
    };
    for (uint32_t r = 0; r < dim; ++r) {
        const uint32_t L = rp[r + 1] - rp[r];
        if (L == 0) {
            er.push_back(r);
            continue;
        }
        if (L > kHeavyMin) {
            if (open) {
                se.push_back(rp[r]);
                open = false;
            }
            hr.push_back(r);
            hf.push_back((uint32_t)pk.size());
            if (nh) {
                const uint32_t split = rp[r] + (*nh)[r];
                const uint32_t nH = (split - rp[r] + kPiece - 1) / kPiece;
                const uint32_t nC = (rp[r + 1] - split + kPiece - 1) / kPiece;
                const uint64_t tot = (uint64_t)nH + nC;
                uint32_t ih = 0, ic = 0;
                for (uint64_t t = 0; t < tot; ++t) {
                    const bool hot = ic >= nC || (ih < nH && (uint64_t)ih * tot < (t + 1) * nH);
                    const uint32_t b = hot ? rp[r] + ih * kPiece : split + ic * kPiece;
                    const uint32_t e = hot ? std::min(b + kPiece, split) : std::min(b + kPiece, rp[r + 1]);
                    pk.push_back(b);
                    pe.push_back(e);
                    pr.push_back(r);
                    ph.push_back(hot ? 1u : 0u);
                    if (hot) ++ih; else ++ic;
                }
            } else {
                cut(r, rp[r], rp[r + 1], 0u);
            }
            hc.push_back((uint32_t)pk.size() - hf.back());
        } else {
            const uint32_t wdw = rp[r] / kLightWindow;
            if (!open || wdw != window) {
                if (open) se.push_back(rp[r]);
                sk.push_back(rp[r]);
                sr.push_back(r);
                window = wdw;
                open = true;
            }
        }
    }
    if (open) se.push_back(rp[dim]);
    for (uint32_t** t : {&cp.seg_k, &cp.seg_e, &cp.seg_row, &cp.pc_k, &cp.pc_e, &cp.pc_row,
                         &cp.pc_hot, &cp.heavy_row, &cp.heavy_first, &cp.heavy_count,
                         &cp.empty_rows}) {
        dfree(c, *t);
        *t = nullptr;
    }
    cp.nseg = (uint32_t)sk.size();
    cp.npieces = (uint32_t)pk.size();
    cp.nheavy = (uint32_t)hr.size();
    cp.nempty = (uint32_t)er.size();
    auto up = [&](uint32_t** d, const std::vector<uint32_t>& v) -> int {
        DYN_CHECK(dalloc_t(c, d, std::max<size_t>(v.size(), 1)));
        if (!v.empty())
            DYN_HIP(c, hipMemcpyAsync(*d, v.data(), v.size() * 4, hipMemcpyHostToDevice, c->stream));
        return DYN_OK;
    };
    DYN_CHECK(up(&cp.seg_k, sk));
    DYN_CHECK(up(&cp.seg_e, se));
    DYN_CHECK(up(&cp.seg_row, sr));
    DYN_CHECK(up(&cp.pc_k, pk));
    DYN_CHECK(up(&cp.pc_e, pe));
    DYN_CHECK(up(&cp.pc_row, pr));
    if (nh) DYN_CHECK(up(&cp.pc_hot, ph));
    DYN_CHECK(up(&cp.heavy_row, hr));
    DYN_CHECK(up(&cp.heavy_first, hf));
    DYN_CHECK(up(&cp.heavy_count, hc));
    DYN_CHECK(up(&cp.empty_rows, er));
    DYN_HIP(c, hipStreamSynchronize(c->stream));
    return DYN_OK;
}

// ---------------------------------------------------------------------------
//  hot split of a sorted copy (see ensure_hot_split)
// ---------------------------------------------------------------------------
__global__ void k_hot_flag(const uint32_t* __restrict__ ix, uint32_t n, uint32_t H,
                           uint32_t* __restrict__ f) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k <= n; k += stride)
        f[k] = (k < n && ix[k] < H) ? 1u : 0u;
}

// per row: hot count = hs[rp[r+1]] - hs[rp[r]] (hs: exclusive scan of the flags)
__global__ void k_row_hot(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ hs,
                          uint32_t dim, uint32_t* __restrict__ nh) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < dim; r += stride)
        nh[r] = hs[rp[r + 1]] - hs[rp[r]];
}

struct HotScatter {
    const uint32_t* rp;
    const uint32_t* hs;                     // exclusive scan of the hot flags (n + 1)
    uint32_t        dim;
    uint32_t        n;
    const float*    vals_in;
    float*          vals_out;
    const uint32_t* perm_in;
    uint32_t*       perm_out;
    const uint32_t* idx_in[kMaxModes];
    uint32_t*       idx_out[kMaxModes];
    int             ncols;
};

__global__ void k_hot_scatter(HotScatter a) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k64 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k64 < a.n; k64 += stride) {
        const uint32_t k = (uint32_t)k64;
        // row of k: the last r with rp[r] <= k
        uint32_t lo = 0, hi = a.dim;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (a.rp[mid] <= k) lo = mid;
            else hi = mid;
        }
        const uint32_t b = a.rp[lo], e = a.rp[lo + 1];
        const uint32_t hb = a.hs[b];
        const uint32_t nh = a.hs[e] - hb;
        const uint32_t hk = a.hs[k] - hb;
        const uint32_t pos = (a.hs[k + 1] != a.hs[k]) ? b + hk : b + nh + (k - b - hk);
        a.vals_out[pos] = a.vals_in[k];
        a.perm_out[pos] = a.perm_in[k];
        for (int j = 0; j < a.ncols; ++j) a.idx_out[j][pos] = a.idx_in[j][k];
    }
}

// all-hot classes: flag a nonzero of a row of at least kClassMin nonzeros
// when every other mode's rank is below its threshold
struct ClassFlag {
    const uint32_t* rp;
    uint32_t        dim;
    uint32_t        n;
    const uint32_t* idx[kMaxModes];
    uint32_t        H[kMaxModes];
    int             nops;
    uint32_t*       f;
};

__global__ void k_class_flag(ClassFlag a) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k64 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k64 <= a.n; k64 += stride) {
        const uint32_t k = (uint32_t)k64;
        uint32_t hot = 0;
        if (k < a.n) {
            uint32_t lo = 0, hi = a.dim;   // the last r with rp[r] <= k
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (a.rp[mid] <= k) lo = mid;
                else hi = mid;
            }
            if (a.rp[lo + 1] - a.rp[lo] >= kClassMin) {
                hot = 1;
                for (int j = 0; j < a.nops; ++j) hot &= (a.idx[j][k] < a.H[j]) ? 1u : 0u;
            }
        }
        a.f[k] = hot;
    }
}

// Regroups copy `mode` inside every row: the nonzeros flagged in hs (n + 1
// entries, flags on entry) first, in their order, then the rest; heavy
// pieces are cut in both groups and interleaved in proportion (pc_hot: 1 for
// a piece of flagged nonzeros).  Frees hs.
static int regroup_rows(dyn_gpu_ctx* c, int mode, uint32_t* hs) {
    ModeCopy& cp = c->copy[mode];
    const uint32_t n = (uint32_t)c->nnz;
    const uint32_t dim = c->dims[mode];
    uint32_t* nh_dev = nullptr;
    DYN_CHECK(dalloc_t(c, &nh_dev, std::max<uint32_t>(dim, 1)));
    void* sp2 = nullptr;
    const uint64_t nb = ((uint64_t)n + 1 + kScanTile - 1) / kScanTile;
    DYN_CHECK(scratch2(c, (nb + 16) * sizeof(uint32_t), &sp2));
    DYN_CHECK(exclusive_scan(c, hs, (uint64_t)n + 1, (uint32_t*)sp2));
    // scatter every column into fresh arrays
    HotScatter a;
    a.rp = cp.rowptr;
    a.hs = hs;
    a.dim = dim;
    a.n = n;
    a.vals_in = cp.vals;
    a.perm_in = cp.perm;
    float* nv = nullptr;
    uint32_t* np = nullptr;
    uint32_t* ni[kMaxModes] = {nullptr};
    DYN_CHECK(dalloc_t(c, &nv, std::max<uint32_t>(n, 1)));
    DYN_CHECK(dalloc_t(c, &np, std::max<uint32_t>(n, 1)));
    a.vals_out = nv;
    a.perm_out = np;
    a.ncols = 0;
    for (int v = 0; v < c->N; ++v) {
        if (v == mode) continue;
        DYN_CHECK(dalloc_t(c, &ni[v], std::max<uint32_t>(n, 1)));
        a.idx_in[a.ncols] = cp.idx[v];
        a.idx_out[a.ncols] = ni[v];
        ++a.ncols;
    }
    if (n) {
        hipLaunchKernelGGL(k_hot_scatter, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, a);
        DYN_CHECK(check_launch(c, "hot_scatter"));
    }
    if (dim) {
        hipLaunchKernelGGL(k_row_hot, dim3(grid_for(dim)), dim3(kBlock), 0, c->stream, cp.rowptr,
                           hs, dim, nh_dev);
        DYN_CHECK(check_launch(c, "row_hot"));
    }
    std::vector<uint32_t> rp((size_t)dim + 1), nh(dim);
    DYN_HIP(c, hipMemcpyAsync(rp.data(), cp.rowptr, rp.size() * 4, hipMemcpyDeviceToHost,
                              c->stream));
    if (dim)
        DYN_HIP(c, hipMemcpyAsync(nh.data(), nh_dev, (size_t)dim * 4, hipMemcpyDeviceToHost,
                                  c->stream));
    DYN_HIP(c, hipStreamSynchronize(c->stream));
    dfree(c, cp.vals);
    dfree(c, cp.perm);
    cp.vals = nv;
    cp.perm = np;
    for (int v = 0; v < c->N; ++v) {
        if (v == mode) continue;
        dfree(c, cp.idx[v]);
        cp.idx[v] = ni[v];
    }
    dfree(c, hs);
    dfree(c, nh_dev);
    return build_segments(c, cp, rp, &nh);
}

int ensure_class_split(dyn_gpu_ctx* c, int mode, const uint32_t* H) {
    ModeCopy& cp = c->copy[mode];
    if (!cp.built) return set_err(c, DYN_ERR_STATE, "mode %d copy not built", mode);
    int nops = 0;
    bool same = cp.hot_w < 0;
    for (int v = 0; v < c->N; ++v) {
        if (v == mode) continue;
        same = same && cp.cls_h[nops] == H[nops];
        ++nops;
    }
    if (same) return DYN_OK;
    const uint32_t n = (uint32_t)c->nnz;
    const uint32_t dim = c->dims[mode];
    uint32_t* hs = nullptr;
    DYN_CHECK(dalloc_t(c, &hs, (size_t)n + 1));
    ClassFlag a;
    a.rp = cp.rowptr;
    a.dim = dim;
    a.n = n;
    a.nops = nops;
    a.f = hs;
    {
        int j = 0;
        for (int v = 0; v < c->N; ++v) {
            if (v == mode) continue;
            a.idx[j] = cp.idx[v];
            a.H[j] = H[j];
            ++j;
        }
    }
    hipLaunchKernelGGL(k_class_flag, dim3(grid_for((uint64_t)n + 1)), dim3(kBlock), 0, c->stream, a);
    DYN_CHECK(check_launch(c, "class_flag"));
    DYN_CHECK(regroup_rows(c, mode, hs));
    for (int j = 0; j < kMaxModes; ++j) cp.cls_h[j] = j < nops ? H[j] : 0u;
    cp.hot_w = -1;
    cp.hot_h = 0;
    return DYN_OK;
}

int ensure_hot_split(dyn_gpu_ctx* c, int mode, int w, uint32_t H) {
    ModeCopy& cp = c->copy[mode];
    if (!cp.built) return set_err(c, DYN_ERR_STATE, "mode %d copy not built", mode);
    if (cp.hot_w == w && cp.hot_h == H) return DYN_OK;
    const uint32_t n = (uint32_t)c->nnz;
    // flags over n + 1 entries (entry n: 0, so the scan's last entry is the total)
    uint32_t* hs = nullptr;
    DYN_CHECK(dalloc_t(c, &hs, (size_t)n + 1));
    hipLaunchKernelGGL(k_hot_flag, dim3(grid_for((uint64_t)n + 1)), dim3(kBlock), 0, c->stream,
                       cp.idx[w], n, H, hs);
    DYN_CHECK(check_launch(c, "hot_flag"));
    DYN_CHECK(regroup_rows(c, mode, hs));
    cp.hot_w = w;
    cp.hot_h = H;
    for (int j = 0; j < kMaxModes; ++j) cp.cls_h[j] = 0;
    return DYN_OK;
}

// The engine's copies key the secondary sort by frequency rank (the memory
// order of the secondary operand's rows, so consecutive nonzeros of a row
// gather neighbouring rows), except that the 1024 most frequent ranks, whose
// rows the launches hold in LDS, are hashed over the rank range: sorted
// first they would leave some waves' pieces all-LDS and others all-memory.
// Measured on one box against the input-index key: 5-mode Zipf(1.0) BF16
// Jacobi 1.3 % faster, 4-mode Zipf(1.2) 0.7 %, uniform completion 3 %
// (plain rank keys: 1.6 % slower on the Zipf(1.2) tensor).  Modes of up to
// 2048 rows keep the input-index key.
static bool secondary_rank_key(const dyn_gpu_ctx* c, int secondary) {
    return c->dims[secondary] > 2048;
}

// key = rank for ranks >= 1024, else 1024 + a multiplicative hash of the rank
// scaled to [0, dim - 1024): always below dim
__global__ void k_spread_key(const uint32_t* __restrict__ relab, const uint32_t* __restrict__ ix,
                             uint32_t n, uint32_t span, uint32_t* __restrict__ key) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        const uint32_t r = relab[ix[k]];
        key[k] = r >= 1024u ? r
                            : 1024u + (uint32_t)(((uint64_t)(r * 0x9E3779B1u) * span) >> 32);
    }
}

int remap_mode(dyn_gpu_ctx* c, int mode, int secondary, bool engine_order) {
    DYN_CHECK(ensure_relabel(c));
    const bool rank_key = engine_order && secondary >= 0 && secondary_rank_key(c, secondary);
    free_copy(c, mode);
    ModeCopy& cp = c->copy[mode];
    const uint32_t n = (uint32_t)c->nnz;
    const uint32_t dim = c->dims[mode];

    DYN_CHECK(dalloc_t(c, &cp.rowptr, (size_t)dim + 1));
    DYN_CHECK(dalloc_t(c, &cp.perm, std::max<uint32_t>(n, 1)));
    DYN_CHECK(dalloc_t(c, &cp.vals, std::max<uint32_t>(n, 1)));
    for (int w = 0; w < c->N; ++w)
        if (w != mode) DYN_CHECK(dalloc_t(c, &cp.idx[w], std::max<uint32_t>(n, 1)));

    // scratch: keys x2, payload x1 (the second payload buffer is cp.perm)
    void* sp = nullptr;
    const size_t nn = std::max<uint32_t>(n, 1);
    DYN_CHECK(scratch(c, nn * 3 * sizeof(uint32_t), &sp));
    uint32_t* kA = (uint32_t*)sp;
    uint32_t* kB = kA + nn;
    uint32_t* vA = kB + nn;
    uint32_t* vB = cp.perm;
    const uint32_t nblocks = (n + kRadixTile - 1) / kRadixTile;
    const uint64_t ncounts = (uint64_t)std::max<uint32_t>(nblocks, 1) * 256;
    void* sp2 = nullptr;
    DYN_CHECK(scratch2(c, (ncounts + ncounts / kScanTile + 16) * sizeof(uint32_t), &sp2));
    uint32_t* counts = (uint32_t*)sp2;
    uint32_t* bsum = counts + ncounts;

    const unsigned g = grid_for(n);
    int which = 0;   // 0: (kA, vA) current, 1: (kB, vB)
    if (n) {
        hipLaunchKernelGGL(k_iota, dim3(g), dim3(kBlock), 0, c->stream, vA, n);
        if (secondary >= 0) {
            if (rank_key) {
                hipLaunchKernelGGL(k_spread_key, dim3(g), dim3(kBlock), 0, c->stream,
                                   c->relab[secondary], c->idx[secondary], n,
                                   c->dims[secondary] - 1024u, kA);
                DYN_CHECK(check_launch(c, "remap rank key"));
            } else {
                DYN_HIP(c, hipMemcpyAsync(kA, c->idx[secondary], (size_t)n * 4,
                                          hipMemcpyDeviceToDevice, c->stream));
            }
            DYN_CHECK(radix_passes(c, kA, vA, kB, vB, n, bits_for(c->dims[secondary]), counts,
                                   bsum, &which));
            uint32_t* kcur = which ? kB : kA;
            uint32_t* vcur = which ? vB : vA;
            hipLaunchKernelGGL(k_gather_u32, dim3(g), dim3(kBlock), 0, c->stream, c->idx[mode],
                               vcur, kcur, n);
        } else {
            DYN_HIP(c, hipMemcpyAsync(kA, c->idx[mode], (size_t)n * 4, hipMemcpyDeviceToDevice,
                                      c->stream));
        }
        DYN_CHECK(radix_passes(c, kA, vA, kB, vB, n, bits_for(dim), counts, bsum, &which));
        uint32_t* keys = which ? kB : kA;
        uint32_t* perm = which ? vB : vA;
        if (perm != cp.perm)
            DYN_HIP(c, hipMemcpyAsync(cp.perm, perm, (size_t)n * 4, hipMemcpyDeviceToDevice,
                                      c->stream));
        DYN_CHECK(materialize_copy(c, mode, keys));
    } else {
        hipLaunchKernelGGL(k_rowptr, dim3(grid_for((uint64_t)dim + 1)), dim3(kBlock), 0,
                           c->stream, kA, 0u, dim, cp.rowptr);
        DYN_CHECK(check_launch(c, "remap empty"));
    }
    DYN_CHECK(finish_copy(c, mode));
    cp.secondary = secondary;
    cp.rank_key = rank_key;
    return DYN_OK;
}

// The columns and row pointer of copy `mode` from its permutation (cp.perm)
// and the sorted keys (idx[mode] in sorted order).
int materialize_copy(dyn_gpu_ctx* c, int mode, const uint32_t* keys) {
    ModeCopy& cp = c->copy[mode];
    const uint32_t n = (uint32_t)c->nnz;
    const unsigned g = grid_for(n);
    hipLaunchKernelGGL(k_gather_f32, dim3(g), dim3(kBlock), 0, c->stream, c->vals, cp.perm,
                       cp.vals, n);
    for (int w = 0; w < c->N; ++w)
        if (w != mode)
            hipLaunchKernelGGL(k_gather_relabel, dim3(g), dim3(kBlock), 0, c->stream, c->idx[w],
                               cp.perm, c->relab[w], cp.idx[w], n);
    hipLaunchKernelGGL(k_rowptr, dim3(grid_for((uint64_t)n + 1)), dim3(kBlock), 0, c->stream,
                       keys, n, c->dims[mode], cp.rowptr);
    return check_launch(c, "remap gather/rowptr");
}

// Work tables of copy `mode` from its row pointer; the copy is then usable.
int finish_copy(dyn_gpu_ctx* c, int mode) {
    ModeCopy& cp = c->copy[mode];
    std::vector<uint32_t> rp((size_t)c->dims[mode] + 1);
    DYN_HIP(c, hipMemcpyAsync(rp.data(), cp.rowptr, rp.size() * 4, hipMemcpyDeviceToHost,
                              c->stream));
    // the scratch buffers are reused by the next call: finish before returning
    DYN_HIP(c, hipStreamSynchronize(c->stream));
    for (size_t r = 0; r + 1 < rp.size(); ++r)
        if (rp[r] > rp[r + 1] || rp[r + 1] > c->nnz)
            return set_err(c, DYN_ERR_INVALID, "mode %d row pointer is not monotone", mode);
    if (rp[0] != 0 || rp.back() != c->nnz)
        return set_err(c, DYN_ERR_INVALID, "mode %d row pointer does not span the tensor", mode);
    DYN_CHECK(build_segments(c, cp, rp));
    cp.built = true;
    return DYN_OK;
}

// Copy `mode` rebuilt from a stored permutation (perm[k] = input position of
// sorted position k): checks that perm is a permutation whose keys are in
// non-decreasing order, then gathers the columns as remap_mode does.
int restore_copy(dyn_gpu_ctx* c, int mode, int secondary, const uint32_t* perm_host) {
    DYN_CHECK(ensure_relabel(c));
    free_copy(c, mode);
    ModeCopy& cp = c->copy[mode];
    const uint32_t n = (uint32_t)c->nnz;
    const uint32_t dim = c->dims[mode];
    DYN_CHECK(dalloc_t(c, &cp.rowptr, (size_t)dim + 1));
    DYN_CHECK(dalloc_t(c, &cp.perm, std::max<uint32_t>(n, 1)));
    DYN_CHECK(dalloc_t(c, &cp.vals, std::max<uint32_t>(n, 1)));
    for (int w = 0; w < c->N; ++w)
        if (w != mode) DYN_CHECK(dalloc_t(c, &cp.idx[w], std::max<uint32_t>(n, 1)));
    if (n) {
        DYN_HIP(c, hipMemcpyAsync(cp.perm, perm_host, (size_t)n * 4, hipMemcpyHostToDevice,
                                  c->stream));
        const size_t words = ((size_t)n + 31) / 32;
        void* sp = nullptr;
        DYN_CHECK(scratch(c, (size_t)n * 4 + words * 4 + 256, &sp));
        uint32_t* keys = (uint32_t*)sp;
        uint32_t* bits = keys + (((size_t)n + 63) & ~size_t(63));
        void* sp2 = nullptr;
        DYN_CHECK(scratch2(c, 2 * sizeof(unsigned long long), &sp2));
        unsigned long long* bad = (unsigned long long*)sp2;
        const unsigned g = grid_for(n);
        hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(words)), dim3(kBlock), 0, c->stream, bits,
                           words);
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(kBlock), 0, c->stream, (uint32_t*)bad,
                           (size_t)4);
        // every entry in range and marked once: a permutation
        hipLaunchKernelGGL(k_perm_check, dim3(g), dim3(kBlock), 0, c->stream, cp.perm, n, bits,
                           bad);
        DYN_CHECK(check_launch(c, "perm_check"));
        unsigned long long hb[2] = {0, 0};
        DYN_HIP(c, hipMemcpyAsync(hb, bad, sizeof(hb), hipMemcpyDeviceToHost, c->stream));
        DYN_HIP(c, hipStreamSynchronize(c->stream));
        if (hb[0])
            return set_err(c, DYN_ERR_INVALID, "stored mode-%d permutation is not a permutation",
                           mode);
        hipLaunchKernelGGL(k_gather_u32, dim3(g), dim3(kBlock), 0, c->stream, c->idx[mode],
                           cp.perm, keys, n);
        hipLaunchKernelGGL(k_count_descents, dim3(g), dim3(kBlock), 0, c->stream, keys, n,
                           bad + 1);
        DYN_CHECK(check_launch(c, "count_descents"));
        DYN_HIP(c, hipMemcpyAsync(hb, bad, sizeof(hb), hipMemcpyDeviceToHost, c->stream));
        DYN_HIP(c, hipStreamSynchronize(c->stream));
        if (hb[1])
            return set_err(c, DYN_ERR_INVALID,
                           "stored mode-%d permutation does not sort the mode's indices", mode);
        DYN_CHECK(materialize_copy(c, mode, keys));
    } else {
        hipLaunchKernelGGL(k_zero_u32, dim3(grid_for((uint64_t)dim + 1)), dim3(kBlock), 0,
                           c->stream, cp.rowptr, (size_t)dim + 1);
        DYN_CHECK(check_launch(c, "restore empty"));
    }
    DYN_CHECK(finish_copy(c, mode));
    cp.secondary = secondary;
    return DYN_OK;
}

}  // namespace dyn
