// Linearised (Z-order) layout and the single-pass all-modes spMTTKRP.
//
// Reference: the Morton / ALTO layout and its all-modes kernel
// (src/dynasor_morton.cpp:103-138 bit plan, :143-360 build, :398-579 kernel;
// include/dynasor_morton.h) and the all-modes Jacobi kernel it shares with the
// SoA path (src/dynasor_kernel.cpp:830-1039, driver src/dynasor.cpp:592-817).
// The reference stores a 64-bit key per nonzero and extracts indices with
// pext; it refuses tensors whose index bits exceed 64.  Here the copy keeps
// one explicit rank column per mode (coalesced loads, no decode), so the key
// is only the sort order and may drop the lowest bits of the largest modes'
// ranks when they do not fit: every tensor can be linearised.
//
// The pass: every nonzero once, every mode's output from one factor
// snapshot (Jacobi semantics).  A wave takes a contiguous Z-order range;
// RP/F lanes (F floats each) hold one nonzero's rows, so a wave instruction
// covers 64/LPR nonzeros.  Per nonzero: N gathered rows, prefix/suffix
// products (the leave-one-out Khatri-Rao rows, as the reference's N >= 4
// path), and N scattered row updates.  A scattered update goes to the wave's
// private LDS tile when its rank is one of the mode's hottest (plain
// read-add-write: the tile belongs to one wave, and the nonzeros of one
// instruction are applied one after the other), else to an FP64 global
// atomic.  Each wave adds its tiles into the rank-order FP64 accumulators at
// the end.  FP64 because a row that takes n scattered additions in FP32 drifts
// by about sqrt(n) units in the last place: rows of ~180k updates were 4e-5
// off at the 400M-nonzero 5-mode workload, against 1e-7 for the per-mode
// copies' tree-shaped sums.  The atomics make the summation order vary from
// run to run; the per-mode copies stay the deterministic path.  On this GPU
// the atomics bound the pass (about 10 G FP32 or 5.8 G FP64 row updates/s
// against about 130 G row gathers/s), so the per-mode segmented passes are
// faster whenever their N copies fit; the linear copy needs one.
#include "dyn_internal.hpp"

#include <hip/hip_ext.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

namespace dyn {

constexpr int kLinThreads = 512;   // 8 waves per workgroup, LDS split evenly among them

// ---------------------------------------------------------------------------
//  key bit plan (host)
// ---------------------------------------------------------------------------
static int ceil_log2_min1(uint32_t v) {
    int b = 0;
    while (b < 32 && ((uint64_t)1 << b) < v) ++b;
    return std::max(b, 1);
}

// Bits of every mode's rank and their key positions: LSB-first round robin
// over the modes that still have bits (src/dynasor_morton.cpp:103-138).  When
// the bits exceed 64, the largest modes give up their lowest bits (shift)
// until the rest fits.  Returns the key's bit count.
static int plan_bits(int N, const uint32_t* dims, uint64_t* mask, uint8_t* bits, uint8_t* shift) {
    int b[kMaxModes] = {0}, kb[kMaxModes] = {0};
    int total = 0;
    for (int n = 0; n < N; ++n) {
        b[n] = ceil_log2_min1(dims[n]);
        kb[n] = b[n];
        total += b[n];
    }
    while (total > 64) {
        int m = 0;
        for (int n = 1; n < N; ++n)
            if (kb[n] > kb[m]) m = n;
        --kb[m];
        --total;
    }
    int rem[kMaxModes];
    for (int n = 0; n < N; ++n) {
        rem[n] = kb[n];
        mask[n] = 0;
        bits[n] = (uint8_t)b[n];
        shift[n] = (uint8_t)(b[n] - kb[n]);
    }
    int pos = 0;
    bool progressed = true;
    while (pos < 64 && progressed) {
        progressed = false;
        for (int n = 0; n < N && pos < 64; ++n) {
            if (rem[n] > 0) {
                mask[n] |= (uint64_t)1 << pos;
                --rem[n];
                ++pos;
                progressed = true;
            }
        }
    }
    return total;
}

// ---------------------------------------------------------------------------
//  key of every input position: the interleaved bits of the (shifted) ranks
// ---------------------------------------------------------------------------
struct KeyArgs {
    const uint32_t* idx[kMaxModes];
    const uint32_t* relab[kMaxModes];
    uint64_t mask[kMaxModes];
    uint32_t shift[kMaxModes];
    int N;
};

__device__ inline uint64_t deposit_bits(uint32_t x, uint64_t m) {
    uint64_t out = 0;
    while (m) {
        const uint64_t low = m & (~m + 1ull);
        if (x & 1u) out |= low;
        x >>= 1;
        m ^= low;
    }
    return out;
}

__global__ void k_zkey(KeyArgs a, uint32_t n, uint32_t* __restrict__ klo,
                       uint32_t* __restrict__ khi) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        uint64_t key = 0;
        for (int m = 0; m < a.N; ++m) {
            const uint32_t r = a.relab[m][a.idx[m][i]] >> a.shift[m];
            key |= deposit_bits(r, a.mask[m]);
        }
        klo[i] = (uint32_t)key;
        khi[i] = (uint32_t)(key >> 32);
    }
}

// out[i][col] = acc[relab[i]][col] for the I rows of one mode (input row order)
__global__ void k_unrelabel(const double* __restrict__ acc, const uint32_t* __restrict__ relab,
                            float* __restrict__ out, size_t rows, int RP) {
    const size_t n = rows * (size_t)RP;
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; e < n; e += stride) {
        const size_t r = e / RP, col = e % RP;
        out[e] = (float)acc[(size_t)relab[r] * RP + col];
    }
}

// ---------------------------------------------------------------------------
//  the all-modes pass
// ---------------------------------------------------------------------------
struct LinArgs {
    const float* vals;
    const uint32_t* idx[kMaxModes];   // ranks
    const float* Y[kMaxModes];        // rank-order FP32 factor rows
    double* acc[kMaxModes];           // rank-order FP64 outputs; null: mode not computed
    uint32_t hot[kMaxModes];          // ranks below hot[n] go to the wave's LDS tile
    uint32_t toff[kMaxModes];         // float offset of mode n's tile in the wave's slice
    uint32_t wave_floats;             // floats of one wave's LDS slice
    uint64_t nnz;
    uint64_t per_wave;                // nonzeros per wave (multiple of 64)
};

template <int RP>
struct LinShape {
    static constexpr int LPR = RP < 64 ? RP : 64;   // lanes holding one nonzero's row
    static constexpr int F = RP / LPR;              // floats per lane
    static constexpr int NPI = 64 / LPR;            // nonzeros per wave instruction
};

// lane j*NPI + sub's value to every lane of sub-group sub (uniform j)
template <int NPI>
__device__ inline uint32_t bcast_u32(uint32_t x, int j, int sub) {
    uint32_t r = __builtin_amdgcn_readlane(x, j * NPI);
#pragma unroll
    for (int s = 1; s < NPI; ++s) {
        const uint32_t t = __builtin_amdgcn_readlane(x, j * NPI + s);
        r = (sub == s) ? t : r;
    }
    return r;
}

template <int RP, int N>
struct LinGroup {
    float v;
    uint32_t r[N];
    float y[N][LinShape<RP>::F];
};

template <int RP, int N>
__device__ inline void lin_load(const LinArgs& a, float vv, const uint32_t (&ii)[N], int j,
                                int sub, int col, LinGroup<RP, N>& g) {
    constexpr int NPI = LinShape<RP>::NPI;
    constexpr int F = LinShape<RP>::F;
    g.v = __uint_as_float(bcast_u32<NPI>(__float_as_uint(vv), j, sub));
#pragma unroll
    for (int n = 0; n < N; ++n) {
        g.r[n] = bcast_u32<NPI>(ii[n], j, sub);
        const float* p = a.Y[n] + (size_t)g.r[n] * RP + col;
        if constexpr (F == 1) {
            g.y[n][0] = p[0];
        } else {
            const float2 t = *reinterpret_cast<const float2*>(p);
            g.y[n][0] = t.x;
            g.y[n][1] = t.y;
        }
    }
}

template <int RP, int N>
__device__ inline void lin_apply(const LinArgs& a, const LinGroup<RP, N>& g, bool valid, int sub,
                                 int col, float* tile) {
    constexpr int NPI = LinShape<RP>::NPI;
    constexpr int F = LinShape<RP>::F;
    // leave-one-out products: out[n] = v * prod_{w != n} y[w]
    float pre[N][F], out[N][F];
#pragma unroll
    for (int f = 0; f < F; ++f) {
        pre[0][f] = g.v;
#pragma unroll
        for (int n = 1; n < N; ++n) pre[n][f] = pre[n - 1][f] * g.y[n - 1][f];
        float suf = 1.0f;
#pragma unroll
        for (int n = N - 1; n >= 0; --n) {
            out[n][f] = pre[n][f] * suf;
            suf *= g.y[n][f];
        }
    }
#pragma unroll
    for (int n = 0; n < N; ++n) {
        if (!a.acc[n]) continue;   // uniform
        const uint32_t r = g.r[n];
        const bool hot = r < a.hot[n];
        if (valid && !hot) {
            double* dst = a.acc[n] + (size_t)r * RP + col;
#pragma unroll
            for (int f = 0; f < F; ++f) unsafeAtomicAdd(dst + f, (double)out[n][f]);
        }
        // the tile is the wave's own; nonzeros of one instruction that share a
        // row are applied one sub-group after the other (LDS operations of a
        // wave complete in order)
        float* t = tile + a.toff[n] + r * RP + col;
#pragma unroll
        for (int s = 0; s < NPI; ++s) {
            if (valid && hot && sub == s) {
#pragma unroll
                for (int f = 0; f < F; ++f) t[f] += out[n][f];
            }
        }
    }
}

template <int RP, int N>
__global__ __launch_bounds__(kLinThreads) void k_lin(LinArgs a) {
    constexpr int LPR = LinShape<RP>::LPR;
    constexpr int F = LinShape<RP>::F;
    constexpr int NPI = LinShape<RP>::NPI;
    extern __shared__ float lds[];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    float* tile = lds + (size_t)wv * a.wave_floats;
    for (uint32_t e = lane; e < a.wave_floats; e += 64) tile[e] = 0.0f;

    // workgroups b, b + 8, ... share an XCD (and its L2): give them adjacent
    // ranges of the Z-order, whose factor rows overlap
    const uint32_t nb = gridDim.x;
    uint32_t blk = blockIdx.x;
    if ((nb & 7u) == 0) blk = (blockIdx.x & 7u) * (nb >> 3) + (blockIdx.x >> 3);
    const uint64_t wave_id = (uint64_t)blk * (kLinThreads / 64) + wv;
    const uint64_t k0 = wave_id * a.per_wave;
    const uint64_t k1 = k0 + a.per_wave < a.nnz ? k0 + a.per_wave : a.nnz;
    const int sub = lane / LPR;
    const int col = (lane % LPR) * F;

    for (uint64_t b = k0; b < k1; b += 64) {
        const uint64_t k = b + lane;
        const bool ok = k < k1;
        const float vv = ok ? a.vals[k] : 0.0f;
        uint32_t ii[N];
#pragma unroll
        for (int n = 0; n < N; ++n) ii[n] = ok ? a.idx[n][k] : 0u;
        const uint32_t cnt = (uint32_t)(k1 - b < 64 ? k1 - b : 64);
        const int J = (int)((cnt + NPI - 1) / NPI);
        // one group of gathers in flight while the previous one is applied
        LinGroup<RP, N> g0, g1;
        lin_load<RP, N>(a, vv, ii, 0, sub, col, g0);
        int j = 0;
        for (; j + 1 < J; j += 2) {
            lin_load<RP, N>(a, vv, ii, j + 1, sub, col, g1);
            lin_apply<RP, N>(a, g0, (uint32_t)(j * NPI + sub) < cnt, sub, col, tile);
            if (j + 2 < J) lin_load<RP, N>(a, vv, ii, j + 2, sub, col, g0);
            lin_apply<RP, N>(a, g1, (uint32_t)((j + 1) * NPI + sub) < cnt, sub, col, tile);
        }
        if (j < J) lin_apply<RP, N>(a, g0, (uint32_t)(j * NPI + sub) < cnt, sub, col, tile);
    }

    // the wave's tiles into the accumulators
#pragma unroll
    for (int n = 0; n < N; ++n) {
        if (!a.acc[n] || a.hot[n] == 0) continue;
        const uint32_t m = a.hot[n] * RP;
        for (uint32_t e = lane; e < m; e += 64) {
            const float x = tile[a.toff[n] + e];
            if (x != 0.0f) unsafeAtomicAdd(a.acc[n] + e, (double)x);
        }
    }
}

template <int RP>
static int launch_lin_n(dyn_gpu_ctx* c, int N, const LinArgs& a, unsigned grid, size_t lds) {
    const void* fn = nullptr;
    switch (N) {
        case 2: fn = (const void*)k_lin<RP, 2>; break;
        case 3: fn = (const void*)k_lin<RP, 3>; break;
        case 4: fn = (const void*)k_lin<RP, 4>; break;
        case 5: fn = (const void*)k_lin<RP, 5>; break;
        case 6: fn = (const void*)k_lin<RP, 6>; break;
        case 7: fn = (const void*)k_lin<RP, 7>; break;
        case 8: fn = (const void*)k_lin<RP, 8>; break;
        default: return set_err(c, DYN_ERR_UNSUPPORTED, "unsupported mode count %d", N);
    }
    static std::atomic<uint64_t> lds_set[7];
    lds_limit_once(c, &lds_set[N - 2], fn, (int)std::min<size_t>(c->max_lds, 160u * 1024u));
    void* args[] = {const_cast<LinArgs*>(&a)};
    if (c->launch_ev[0]) {
        if (hipExtLaunchKernel(fn, dim3(grid), dim3(kLinThreads), args, lds, c->stream,
                               c->launch_ev[0], c->launch_ev[1], 0) != hipSuccess)
            return set_err(c, DYN_ERR_HIP, "linear pass launch failed");
    } else {
        if (hipLaunchKernel(fn, dim3(grid), dim3(kLinThreads), args, lds, c->stream) != hipSuccess)
            return set_err(c, DYN_ERR_HIP, "linear pass launch failed");
    }
    return check_launch(c, "linear all-modes pass");
}

// Rows of each computed mode that get a private LDS tile: the most frequent
// ranks over all modes, as long as a rank receives several updates per wave
// (a tile row costs one flush per wave) and the wave's slice holds them.
static void pick_hot_rows(dyn_gpu_ctx* c, uint32_t mask, uint32_t rows_budget, uint64_t nwaves,
                          uint32_t* hot) {
    const int N = c->N;
    uint32_t pos[kMaxModes] = {0};
    for (int n = 0; n < N; ++n) hot[n] = 0;
    const uint64_t min_count = 4 * nwaves;
    for (uint32_t k = 0; k < rows_budget; ++k) {
        int best = -1;
        uint32_t bc = 0;
        for (int n = 0; n < N; ++n) {
            if (!((mask >> n) & 1u)) continue;
            const std::vector<uint32_t>& rc = c->rank_count_host[n];
            if (pos[n] >= rc.size()) continue;
            if (best < 0 || rc[pos[n]] > bc) {
                best = n;
                bc = rc[pos[n]];
            }
        }
        if (best < 0 || bc < min_count) break;
        ++pos[best];
    }
    for (int n = 0; n < N; ++n) hot[n] = pos[n];
}

void free_linear(dyn_gpu_ctx* c) {
    LinearCopy& L = c->lin;
    dfree(c, L.vals);
    for (int n = 0; n < kMaxModes; ++n) dfree(c, L.idx[n]);
    dfree(c, L.perm);
    dfree(c, L.acc);
    L = LinearCopy();
}

int build_linear(dyn_gpu_ctx* c) {
    if (c->streamed) return reject_streamed(c, "the linearised layout");
    if (c->nnz >= kMaxNnz)
        return set_err(c, DYN_ERR_UNSUPPORTED, "linearised layout: nnz %llu >= 2^32 - 1",
                       (unsigned long long)c->nnz);
    DYN_CHECK(ensure_relabel(c));
    free_linear(c);
    LinearCopy& L = c->lin;
    const int N = c->N;
    const uint32_t n = (uint32_t)c->nnz;
    const size_t nn = std::max<uint32_t>(n, 1);
    const int total = plan_bits(N, c->dims, L.mask, L.bits, L.shift);

    DYN_CHECK(dalloc_t(c, &L.perm, nn));
    DYN_CHECK(dalloc_t(c, &L.vals, nn));
    for (int m = 0; m < N; ++m) DYN_CHECK(dalloc_t(c, &L.idx[m], nn));

    void* sp = nullptr;
    DYN_CHECK(scratch(c, nn * 4 * sizeof(uint32_t), &sp));
    uint32_t* kA = (uint32_t*)sp;
    uint32_t* kB = kA + nn;
    uint32_t* vA = kB + nn;
    uint32_t* khi = vA + nn;
    uint32_t* vB = L.perm;
    if (n) {
        KeyArgs ka;
        std::memset(&ka, 0, sizeof(ka));
        ka.N = N;
        for (int m = 0; m < N; ++m) {
            ka.idx[m] = c->idx[m];
            ka.relab[m] = c->relab[m];
            ka.mask[m] = L.mask[m];
            ka.shift[m] = L.shift[m];
        }
        hipLaunchKernelGGL(k_zkey, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, ka, n, kA, khi);
        DYN_CHECK(check_launch(c, "linear keys"));
        DYN_CHECK(iota_u32(c, vA, n));
        int which = 0;
        DYN_CHECK(sort_pairs(c, kA, vA, kB, vB, n, std::min(total, 32), &which));
        if (total > 32) {
            uint32_t* kcur = which ? kB : kA;
            uint32_t* vcur = which ? vB : vA;
            DYN_CHECK(gather_u32(c, khi, vcur, kcur, n));
            DYN_CHECK(sort_pairs(c, kA, vA, kB, vB, n, total - 32, &which));
        }
        uint32_t* perm = which ? vB : vA;
        if (perm != L.perm)
            DYN_HIP(c, hipMemcpyAsync(L.perm, perm, (size_t)n * 4, hipMemcpyDeviceToDevice,
                                      c->stream));
        DYN_CHECK(gather_f32(c, c->vals, L.perm, L.vals, n));
        for (int m = 0; m < N; ++m)
            DYN_CHECK(gather_relabel(c, c->idx[m], L.perm, c->relab[m], L.idx[m], n));
    }
    // the scratch buffers are reused by the next call: finish before returning
    DYN_HIP(c, hipStreamSynchronize(c->stream));
    L.built = true;
    return DYN_OK;
}

int linear_sweep(dyn_gpu_ctx* c, uint32_t mask) {
    LinearCopy& L = c->lin;
    if (!L.built) DYN_CHECK(build_linear(c));
    DYN_CHECK(ensure_gather(c));
    const int N = c->N;
    const int Rp = c->Rp;
    size_t need = 0, off[kMaxModes] = {0};
    for (int m = 0; m < N; ++m) {
        off[m] = need;
        need += (size_t)c->dims[m] * Rp;
    }
    if (L.acc_count < need) {
        dfree(c, L.acc);
        L.acc = nullptr;
        L.acc_count = 0;
        DYN_CHECK(dalloc_t(c, &L.acc, need));
        L.acc_count = need;
    }
    LinArgs a;
    std::memset(&a, 0, sizeof(a));
    a.vals = L.vals;
    for (int m = 0; m < N; ++m) {
        a.idx[m] = L.idx[m];
        a.Y[m] = c->Yg[m];
        if ((mask >> m) & 1u) {
            a.acc[m] = L.acc + off[m];
            // all-zero bytes are 0.0 in FP64 too
            DYN_CHECK(fill_f32(c, reinterpret_cast<float*>(a.acc[m]), 2 * (size_t)c->dims[m] * Rp,
                               0.0f));
        }
    }
    const unsigned grid = (unsigned)c->num_cus;
    const uint64_t nwaves = (uint64_t)grid * (kLinThreads / 64);
    const size_t lds_total = std::min<size_t>(c->max_lds, 160u * 1024u);
    a.wave_floats = (uint32_t)(lds_total / (kLinThreads / 64) / sizeof(float));
    uint32_t hot[kMaxModes];
    pick_hot_rows(c, mask, a.wave_floats / (uint32_t)Rp, nwaves, hot);
    uint32_t t = 0;
    for (int m = 0; m < N; ++m) {
        a.hot[m] = hot[m];
        a.toff[m] = t;
        t += hot[m] * (uint32_t)Rp;
    }
    a.nnz = c->nnz;
    a.per_wave = ((c->nnz + nwaves - 1) / nwaves + 63) / 64 * 64;
    const size_t lds = (size_t)a.wave_floats * sizeof(float) * (kLinThreads / 64);

    TimedPair tp;
    const bool pooled = !c->seg_pair && timing_acquire(c, &tp);
    if (c->seg_pair) {
        c->launch_ev[0] = c->seg_pair[0];
        c->launch_ev[1] = c->seg_pair[1];
    } else if (pooled) {
        c->launch_ev[0] = tp.a;
        c->launch_ev[1] = tp.b;
    }
    int rc = DYN_OK;
    if (c->nnz == 0) {
        if (c->launch_ev[0]) {
            (void)hipEventRecord(c->launch_ev[0], c->stream);
            (void)hipEventRecord(c->launch_ev[1], c->stream);
        }
    } else {
        switch (Rp) {
            case 16: rc = launch_lin_n<16>(c, N, a, grid, lds); break;
            case 32: rc = launch_lin_n<32>(c, N, a, grid, lds); break;
            case 64: rc = launch_lin_n<64>(c, N, a, grid, lds); break;
            case 128: rc = launch_lin_n<128>(c, N, a, grid, lds); break;
            default: rc = set_err(c, DYN_ERR_UNSUPPORTED, "unsupported padded rank %d", Rp);
        }
    }
    c->launch_ev[0] = c->launch_ev[1] = nullptr;
    if (pooled) {
        if (rc == DYN_OK) c->timed.push_back(tp);
        else c->event_pool.push_back(tp);
    }
    DYN_CHECK(rc);
    for (int m = 0; m < N; ++m) {
        if (!((mask >> m) & 1u)) continue;
        const size_t rows = c->dims[m];
        hipLaunchKernelGGL(k_unrelabel, dim3(grid_for(rows * Rp)), dim3(kBlock), 0, c->stream,
                           a.acc[m], c->relab[m], c->Yhat[m], rows, Rp);
    }
    return check_launch(c, "linear unrelabel");
}

}  // namespace dyn

using namespace dyn;

// ===========================================================================
//  C-ABI: linearised layout
// ===========================================================================
extern "C" int dyn_linear_plan(int N, const uint32_t* dims, uint64_t* mask, uint8_t* bits,
                               uint8_t* shift) {
    if (N < 1 || N > kMaxModes || !dims || !mask || !bits || !shift) return -1;
    for (int n = 0; n < N; ++n)
        if (dims[n] == 0) return -1;
    return plan_bits(N, dims, mask, bits, shift);
}

extern "C" int dyn_gpu_build_linear(dyn_gpu_ctx* c, uint32_t* perm_out) {
    if (!c) return DYN_ERR_INVALID;
    if (c->N == 0) return set_err(c, DYN_ERR_STATE, "no tensor uploaded");
    DYN_HIP(c, hipSetDevice(c->device));
    DYN_CHECK(build_linear(c));
    if (perm_out && c->nnz) {
        DYN_HIP(c, hipMemcpyAsync(perm_out, c->lin.perm, c->nnz * sizeof(uint32_t),
                                  hipMemcpyDeviceToHost, c->stream));
        DYN_HIP(c, hipStreamSynchronize(c->stream));
    }
    return DYN_OK;
}

extern "C" int dyn_gpu_set_layout(dyn_gpu_ctx* c, int layout) {
    if (!c) return DYN_ERR_INVALID;
    if (layout != DYN_LAYOUT_COPIES && layout != DYN_LAYOUT_LINEAR)
        return set_err(c, DYN_ERR_INVALID, "bad layout %d", layout);
    if (layout == DYN_LAYOUT_COPIES) {
        c->use_linear = false;
        return DYN_OK;
    }
    if (c->N == 0) return set_err(c, DYN_ERR_STATE, "no tensor uploaded");
    if (c->streamed) return reject_streamed(c, "the linearised layout");
    DYN_HIP(c, hipSetDevice(c->device));
    if (!c->lin.built) DYN_CHECK(build_linear(c));
    // the per-mode copies are what the linear copy replaces: release them
    for (int m = 0; m < c->N; ++m) free_copy(c, m);
    c->use_linear = true;
    return DYN_OK;
}
