// Per-mode remap on the device: stable LSD radix sort of the nonzeros by the
// output-mode coordinate (optionally with a secondary coordinate as the
// minor key), gather of the value / other-index columns into the mode copy,
// and the row pointer plus the kSeg-granular segment table.
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
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t stride = gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = i;
}

__global__ void k_gather_u32(const uint32_t* __restrict__ src, const uint32_t* __restrict__ perm,
                             uint32_t* __restrict__ dst, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t stride = gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[perm[i]];
}

__global__ void k_gather_f32(const float* __restrict__ src, const uint32_t* __restrict__ perm,
                             float* __restrict__ dst, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t stride = gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[perm[i]];
}

// rowptr[r] = first sorted position whose key >= r, for r in [0, dim]
__global__ void k_rowptr(const uint32_t* __restrict__ keys, uint32_t n, uint32_t dim,
                         uint32_t* __restrict__ rowptr) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t stride = gridDim.x * blockDim.x;
    for (; p <= n; p += stride) {
        const uint32_t lo = (p == 0) ? 0u : keys[p - 1] + 1u;   // rows (keys[p-1], ...]
        const uint32_t hi = (p == n) ? dim : keys[p];           // ... keys[p]]
        for (uint32_t r = lo; r <= hi; ++r) rowptr[r] = p;
    }
}

// seg[s] = row of sorted position s*kSeg; seg[nseg] = dim
__global__ void k_segtable(const uint32_t* __restrict__ keys, uint32_t n, uint32_t nseg,
                           uint32_t dim, uint32_t* __restrict__ seg) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t stride = gridDim.x * blockDim.x;
    for (; s <= nseg; s += stride) {
        const uint64_t p = (uint64_t)s * kSeg;
        seg[s] = (s == nseg || p >= n) ? dim : keys[p];
    }
}

static int bits_for(uint32_t dim) {
    int b = 0;
    while (b < 32 && ((uint64_t)1 << b) < dim) ++b;
    return b;
}

static unsigned grid_for(uint64_t n) {
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

int remap_mode(dyn_gpu_ctx* c, int mode, int secondary) {
    free_copy(c, mode);
    ModeCopy& cp = c->copy[mode];
    const uint32_t n = (uint32_t)c->nnz;
    const uint32_t dim = c->dims[mode];
    const uint32_t nseg = (uint32_t)((c->nnz + kSeg - 1) / kSeg);

    DYN_CHECK(dalloc_t(c, &cp.rowptr, (size_t)dim + 1));
    DYN_CHECK(dalloc_t(c, &cp.seg, (size_t)nseg + 1));
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
            DYN_HIP(c, hipMemcpyAsync(kA, c->idx[secondary], (size_t)n * 4,
                                      hipMemcpyDeviceToDevice, c->stream));
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
        hipLaunchKernelGGL(k_gather_f32, dim3(g), dim3(kBlock), 0, c->stream, c->vals, cp.perm,
                           cp.vals, n);
        for (int w = 0; w < c->N; ++w)
            if (w != mode)
                hipLaunchKernelGGL(k_gather_u32, dim3(g), dim3(kBlock), 0, c->stream, c->idx[w],
                                   cp.perm, cp.idx[w], n);
        hipLaunchKernelGGL(k_rowptr, dim3(grid_for((uint64_t)n + 1)), dim3(kBlock), 0, c->stream,
                           keys, n, dim, cp.rowptr);
        hipLaunchKernelGGL(k_segtable, dim3(grid_for((uint64_t)nseg + 1)), dim3(kBlock), 0,
                           c->stream, keys, n, nseg, dim, cp.seg);
        DYN_CHECK(check_launch(c, "remap gather/rowptr"));
    } else {
        hipLaunchKernelGGL(k_rowptr, dim3(grid_for((uint64_t)dim + 1)), dim3(kBlock), 0,
                           c->stream, kA, 0u, dim, cp.rowptr);
        hipLaunchKernelGGL(k_segtable, dim3(1), dim3(kBlock), 0, c->stream, kA, 0u, 0u, dim,
                           cp.seg);
        DYN_CHECK(check_launch(c, "remap empty"));
    }
    // the scratch buffers are reused by the next call: finish before returning
    DYN_HIP(c, hipStreamSynchronize(c->stream));
    cp.built = true;
    cp.secondary = secondary;
    return DYN_OK;
}

}  // namespace dyn
