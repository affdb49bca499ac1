// Segmented spMTTKRP on a per-mode sorted copy, plus the kernels that share
// its walk: the adjoint (values replaced by a gradient vector read through
// the copy's permutation) and the fused sparse-completion loss/gradient.
//
// Definition (src/dynasor_reference.cpp:18-50):
//     Yhat_n[i, :] = sum_{k : idx_n[k] = i} v_k * prod_{w != n} Y_w[idx_w[k], :]
// Per nonzero the product is formed in the reference kernel's order
// (src/dynasor_kernel.cpp:675-704): v*Y[w0], times the middle modes, then a
// fused multiply-add with Y[w_last] into the row accumulator.
//
// Work decomposition (gfx950, wave64):
//   * a quad (4 lanes) owns one nonzero at a time; lane q holds Rp/4 columns
//     of the row, loaded 16 bytes at a time (4 FP32 or 8 BF16 columns per
//     load, widened to FP32), so one quad reads a whole factor row per
//     operand and the per-nonzero bookkeeping is paid by 4 lanes only,
//   * each quad walks kSeg consecutive nonzeros of the sorted copy; the value
//     and index columns are loaded 4 at a time (one element per lane,
//     contiguous) and broadcast inside the quad with DPP quad permutes; the
//     gathers of a batch are issued before its arithmetic,
//   * the output row is accumulated in registers while the row pointer says
//     the row continues; a row that starts and ends inside the quad's range
//     is stored directly,
//   * rows cut by a quad boundary leave a partial in LDS; one workgroup-level
//     pass sums partials of the same row in quad order; rows cut by a
//     workgroup boundary leave a partial in a global table that a fix-up
//     kernel sums in workgroup order.  No atomics: results are bitwise
//     reproducible run to run.
#include "dyn_internal.hpp"

#include <algorithm>

namespace dyn {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kQuads = kBlock / 4;            // quads per workgroup
constexpr uint32_t kChunk = (uint32_t)kQuads * kSeg;   // nonzeros per workgroup

struct SegArgs {
    const float*    vals;         // copy values, or the override vector
    const uint32_t* vperm;        // non-null: value of position k is vals[vperm[k]]
    const uint32_t* oidx[kMaxModes - 1];
    const void*     Yo[kMaxModes - 1];
    const uint32_t* rowptr;
    const uint32_t* seg;
    float*          out;
    float*          bpart;        // [nwg][2][RP]
    uint32_t*       brow;         // [nwg][2]
    // completion kernel only
    const float*    An;           // output-mode factor rows (FP32)
    double*         loss_part;    // [nwg] or null
    int             loss;
    uint32_t        nnz;
    uint32_t        nwg;
};

// ---------------------------------------------------------------------------
//  lane <-> column mapping.  A row is read with 16-byte loads; load b of lane
//  q covers columns [b*4*LW + q*LW, +LW) with LW = 4 (FP32) or 8 (BF16, or 4
//  when the row is only 16 wide).  The same mapping is used for the
//  accumulators and the row stores.
// ---------------------------------------------------------------------------
template <int RP, typename FT>
struct Cols {
    static constexpr int LW = (sizeof(FT) == 2 && RP >= 32) ? 8 : 4;
    static constexpr int NB = RP / (4 * LW);
    static constexpr int CPL = RP / 4;         // columns per lane
    static __device__ __forceinline__ int col(int b, int q, int t) { return b * 4 * LW + q * LW + t; }
};

__device__ __forceinline__ void widen(uint32_t u, float& lo, float& hi) {
    lo = __uint_as_float(u << 16);
    hi = __uint_as_float(u & 0xFFFF0000u);
}

template <int RP, typename FT>
__device__ __forceinline__ void load_row(const FT* __restrict__ row, int q, float* v) {
    using C = Cols<RP, FT>;
#pragma unroll
    for (int b = 0; b < C::NB; ++b) {
        if constexpr (sizeof(FT) == 4) {
            const float4 x = *reinterpret_cast<const float4*>(
                reinterpret_cast<const float*>(row) + C::col(b, q, 0));
            v[b * 4 + 0] = x.x;
            v[b * 4 + 1] = x.y;
            v[b * 4 + 2] = x.z;
            v[b * 4 + 3] = x.w;
        } else if constexpr (C::LW == 8) {
            const uint4 x = *reinterpret_cast<const uint4*>(
                reinterpret_cast<const uint16_t*>(row) + C::col(b, q, 0));
            widen(x.x, v[b * 8 + 0], v[b * 8 + 1]);
            widen(x.y, v[b * 8 + 2], v[b * 8 + 3]);
            widen(x.z, v[b * 8 + 4], v[b * 8 + 5]);
            widen(x.w, v[b * 8 + 6], v[b * 8 + 7]);
        } else {
            const uint2 x = *reinterpret_cast<const uint2*>(
                reinterpret_cast<const uint16_t*>(row) + C::col(b, q, 0));
            widen(x.x, v[b * 4 + 0], v[b * 4 + 1]);
            widen(x.y, v[b * 4 + 2], v[b * 4 + 3]);
        }
    }
}

template <int RP, typename FT>
__device__ __forceinline__ void store_row(float* __restrict__ row, int q, const float* v) {
    using C = Cols<RP, FT>;
#pragma unroll
    for (int b = 0; b < C::NB; ++b)
#pragma unroll
        for (int h = 0; h < C::LW / 4; ++h) {
            const int t = b * C::LW + h * 4;
            *reinterpret_cast<float4*>(row + C::col(b, q, h * 4)) =
                make_float4(v[t], v[t + 1], v[t + 2], v[t + 3]);
        }
}

// quad broadcast of lane J (DPP quad_perm [J,J,J,J])
template <int J>
__device__ __forceinline__ uint32_t qb(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, J * 0x55, 0xF, 0xF, false);
}
template <int J>
__device__ __forceinline__ float qbf(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), J * 0x55, 0xF, 0xF, false));
}
__device__ __forceinline__ float quad_sum(float x) {
    x += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xB1, 0xF, 0xF, false));
    x += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x4E, 0xF, 0xF, false));
    return x;
}

// Hardware dispatches workgroup i to XCD i % 8; hand each XCD a contiguous
// range of chunks so neighbouring chunks (which share boundary rows and
// row-pointer lines) stay in one L2.
__device__ __forceinline__ uint32_t xcd_chunk(uint32_t bid, uint32_t nb) {
    const uint32_t nx = 8;
    const uint32_t per = nb / nx, rem = nb % nx;
    const uint32_t x = bid % nx, i = bid / nx;
    return x * per + min(x, rem) + i;
}

__device__ __forceinline__ float dloss(int kind, float m, float y) {
    if (kind == DYN_LOSS_BERNOULLI) return 1.0f / (1.0f + __expf(-m)) - y;
    return m - y;
}
__device__ __forceinline__ double lossval(int kind, float m, float y) {
    if (kind == DYN_LOSS_BERNOULLI) {
        const double md = m;
        const double sp = (md > 0.0 ? md : 0.0) + log1p(exp(-fabs(md)));
        return sp - (double)y * md;
    }
    const double d = (double)m - (double)y;
    return 0.5 * d * d;
}

// nonzeros of a 4-batch whose gathers are issued together
template <int RP, int NO>
constexpr int batch_hoist() {
    constexpr int regs = NO * (RP / 4);
    return regs >= 64 ? 1 : (regs >= 32 ? 2 : 4);
}

template <int J, int NO>
__device__ __forceinline__ void bcast_batch(float v_l, const uint32_t* i_l, float& v, uint32_t* iw) {
    v = qbf<J>(v_l);
#pragma unroll
    for (int w = 0; w < NO; ++w) iw[w] = qb<J>(i_l[w]);
}

// KIND 0: spMTTKRP / adjoint.  KIND 1: fused completion loss + gradient.
template <int RP, int NO, typename FT, int KIND>
__global__ __launch_bounds__(kBlock) void k_seg(SegArgs a) {
    using C = Cols<RP, FT>;
    constexpr int CPL = C::CPL;
    constexpr int HB = batch_hoist<RP, NO>();
    __shared__ float pacc[2 * kQuads][RP];
    __shared__ uint32_t prow[2 * kQuads];
    __shared__ double lred[kBlock / 64];

    const int tid = threadIdx.x;
    const int g = tid >> 2;
    const int q = tid & 3;
    const uint32_t wg = xcd_chunk(blockIdx.x, a.nwg);
    const uint32_t wg_s = wg * kChunk;
    const uint32_t wg_e = min(wg_s + kChunk, a.nnz);

    if (tid < 2 * kQuads) prow[tid] = kNone;
    if (tid < 2) a.brow[2 * wg + tid] = kNone;
    __syncthreads();

    const uint32_t s = wg_s + (uint32_t)g * kSeg;
    const uint32_t e = min(s + (uint32_t)kSeg, a.nnz);
    double lsum = 0.0;

    if (s < e) {
        uint32_t row = a.seg[s / kSeg];
        uint32_t rbeg = a.rowptr[row];
        uint32_t rend = a.rowptr[row + 1];
        float acc[CPL];
#pragma unroll
        for (int c = 0; c < CPL; ++c) acc[c] = 0.f;
        float arow[KIND == 1 ? CPL : 1];
        if constexpr (KIND == 1) load_row<RP, float>(a.An + (size_t)row * RP, q, arow);

        auto flush = [&](uint32_t r, uint32_t rb, uint32_t re) {
            const bool os = rb < s, oe = re > e;
            if (!os && !oe) {
                store_row<RP, FT>(a.out + (size_t)r * RP, q, acc);
            } else {
                const int slot = 2 * g + (os ? 0 : 1);
                store_row<RP, FT>(&pacc[slot][0], q, acc);
                if (q == 0) prow[slot] = r;
            }
        };

        auto step = [&](uint32_t k, float v, const float (*y)[CPL]) {
            if (k >= rend) {
                flush(row, rbeg, rend);
#pragma unroll
                for (int c = 0; c < CPL; ++c) acc[c] = 0.f;
                do {
                    ++row;
                    rbeg = rend;
                    rend = a.rowptr[row + 1];
                } while (k >= rend);
                if constexpr (KIND == 1) load_row<RP, float>(a.An + (size_t)row * RP, q, arow);
            }
            if constexpr (KIND == 0) {
                if constexpr (NO == 1) {
#pragma unroll
                    for (int c = 0; c < CPL; ++c) acc[c] = __fmaf_rn(v, y[0][c], acc[c]);
                } else {
                    float p[CPL];
#pragma unroll
                    for (int c = 0; c < CPL; ++c) p[c] = v * y[0][c];
#pragma unroll
                    for (int w = 1; w < NO - 1; ++w)
#pragma unroll
                        for (int c = 0; c < CPL; ++c) p[c] *= y[w][c];
#pragma unroll
                    for (int c = 0; c < CPL; ++c) acc[c] = __fmaf_rn(p[c], y[NO - 1][c], acc[c]);
                }
            } else {
                float p[CPL];
#pragma unroll
                for (int c = 0; c < CPL; ++c) p[c] = y[0][c];
#pragma unroll
                for (int w = 1; w < NO; ++w)
#pragma unroll
                    for (int c = 0; c < CPL; ++c) p[c] *= y[w][c];
                float part = 0.f;
#pragma unroll
                for (int c = 0; c < CPL; ++c) part = __fmaf_rn(p[c], arow[c], part);
                const float m = quad_sum(part);
                const float gr = dloss(a.loss, m, v);
#pragma unroll
                for (int c = 0; c < CPL; ++c) acc[c] = __fmaf_rn(gr, p[c], acc[c]);
                if (a.loss_part && q == 0) lsum += lossval(a.loss, m, v);
            }
        };

        // value / index columns of the next 4-batch are loaded while the
        // current batch's gathers are in flight
        auto fetch = [&](uint32_t kb, float& v_l, uint32_t* i_l) {
            const uint32_t kk = kb + q;
            if (kk < e) {
                v_l = a.vperm ? a.vals[a.vperm[kk]] : a.vals[kk];
#pragma unroll
                for (int w = 0; w < NO; ++w) i_l[w] = a.oidx[w][kk];
            } else {
                v_l = 0.f;
#pragma unroll
                for (int w = 0; w < NO; ++w) i_l[w] = 0u;
            }
        };
        float v_n;
        uint32_t i_n[NO];
        fetch(s, v_n, i_n);
        for (uint32_t k0 = s; k0 < e; k0 += 4) {
            const float v_l = v_n;
            uint32_t i_l[NO];
#pragma unroll
            for (int w = 0; w < NO; ++w) i_l[w] = i_n[w];
            if (k0 + 4 < e) fetch(k0 + 4, v_n, i_n);
            const uint32_t cnt = min(4u, e - k0);
            float vb[4];
            uint32_t ib[4][NO];
            bcast_batch<0, NO>(v_l, i_l, vb[0], ib[0]);
            bcast_batch<1, NO>(v_l, i_l, vb[1], ib[1]);
            bcast_batch<2, NO>(v_l, i_l, vb[2], ib[2]);
            bcast_batch<3, NO>(v_l, i_l, vb[3], ib[3]);
#pragma unroll
            for (int j0 = 0; j0 < 4; j0 += HB) {
                float y[HB][NO][CPL];
#pragma unroll
                for (int h = 0; h < HB; ++h)
#pragma unroll
                    for (int w = 0; w < NO; ++w)
                        load_row<RP, FT>(reinterpret_cast<const FT*>(a.Yo[w]) +
                                             (size_t)ib[j0 + h][w] * RP,
                                         q, y[h][w]);
#pragma unroll
                for (int h = 0; h < HB; ++h)
                    if ((uint32_t)(j0 + h) < cnt) step(k0 + j0 + h, vb[j0 + h], y[h]);
            }
        }
        flush(row, rbeg, rend);
    }
    __syncthreads();

    // workgroup-level combine of the partial rows, in quad order
    for (int half = 0; half < 2; ++half) {
        const int slot = 2 * g + half;
        const uint32_t r = prow[slot];
        if (r == kNone) continue;
        int p = slot - 1;
        while (p >= 0 && prow[p] == kNone) --p;
        if (p >= 0 && prow[p] == r) continue;   // not the first partial of its row
        float sum[CPL];
#pragma unroll
        for (int b = 0; b < C::NB; ++b)
#pragma unroll
            for (int t = 0; t < C::LW; ++t) sum[b * C::LW + t] = pacc[slot][C::col(b, q, t)];
        for (int t2 = slot + 1; t2 < 2 * kQuads; ++t2) {
            const uint32_t rt = prow[t2];
            if (rt == kNone) continue;
            if (rt != r) break;
#pragma unroll
            for (int b = 0; b < C::NB; ++b)
#pragma unroll
                for (int t = 0; t < C::LW; ++t) sum[b * C::LW + t] += pacc[t2][C::col(b, q, t)];
        }
        const bool os = a.rowptr[r] < wg_s;
        const bool oe = a.rowptr[r + 1] > wg_e;
        if (!os && !oe) {
            store_row<RP, FT>(a.out + (size_t)r * RP, q, sum);
        } else {
            const int bs = os ? 0 : 1;
            store_row<RP, FT>(a.bpart + ((size_t)wg * 2 + bs) * RP, q, sum);
            if (q == 0) a.brow[2 * wg + bs] = r;
        }
    }

    if (KIND == 1 && a.loss_part) {
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) lsum += __shfl_xor(lsum, o, 64);
        if ((tid & 63) == 0) lred[tid >> 6] = lsum;
        __syncthreads();
        if (tid == 0) {
            double t = 0.0;
            for (int w = 0; w < kBlock / 64; ++w) t += lred[w];
            a.loss_part[wg] = t;
        }
    }
}

// Rows cut by workgroup boundaries: a partial in slot 1 of workgroup b opens
// the row; slot 0 of workgroups b+1, b+2, ... continue it while the row id
// matches.  Each quad of a fix-up block owns one workgroup slot; chains of up
// to kShortChain partials are summed by their quad in order, longer ones
// (heavy rows) by the whole block: quads stride the chain, then an ordered
// LDS reduction.  Deterministic.
constexpr int kShortChain = 8;

template <int RP>
__global__ __launch_bounds__(kBlock) void k_fixup(const float* __restrict__ bpart,
                                                  const uint32_t* __restrict__ brow,
                                                  uint32_t nwg, float* __restrict__ out) {
    using C = Cols<RP, float>;
    constexpr int CPL = C::CPL;
    __shared__ float red[kQuads][RP];
    __shared__ uint32_t longb[kQuads];
    __shared__ uint32_t nlong;
    const int tid = threadIdx.x;
    const int g = tid >> 2;
    const int q = tid & 3;
    if (tid == 0) nlong = 0;
    __syncthreads();
    const uint32_t b = blockIdx.x * kQuads + g;
    if (b < nwg) {
        const uint32_t r = brow[2 * b + 1];
        if (r != kNone) {
            float sum[CPL];
            load_row<RP, float>(bpart + ((size_t)b * 2 + 1) * RP, q, sum);
            uint32_t c = b + 1;
            int len = 0;
            for (; c < nwg && len < kShortChain; ++c, ++len) {
                if (brow[2 * c] != r) break;
                float t[CPL];
                load_row<RP, float>(bpart + ((size_t)c * 2) * RP, q, t);
#pragma unroll
                for (int i = 0; i < CPL; ++i) sum[i] += t[i];
            }
            const bool more = (c < nwg) && (len == kShortChain) && (brow[2 * c] == r);
            if (!more) {
                store_row<RP, float>(out + (size_t)r * RP, q, sum);
            } else if (q == 0) {
                const uint32_t slot = atomicAdd(&nlong, 1u);
                longb[slot] = b;
            }
        }
    }
    __syncthreads();
    const uint32_t nl = nlong;
    // process long chains in ascending workgroup order (slot order from the
    // atomic is arbitrary, the arithmetic below does not depend on it)
    for (uint32_t i = 0; i < nl; ++i) {
        const uint32_t lb = longb[i];
        const uint32_t r = brow[2 * lb + 1];
        float sum[CPL];
#pragma unroll
        for (int j = 0; j < CPL; ++j) sum[j] = 0.f;
        if (g == 0) load_row<RP, float>(bpart + ((size_t)lb * 2 + 1) * RP, q, sum);
        for (uint32_t c = lb + 1 + g; c < nwg; c += kQuads) {
            if (brow[2 * c] != r) break;
            float t[CPL];
            load_row<RP, float>(bpart + ((size_t)c * 2) * RP, q, t);
#pragma unroll
            for (int j = 0; j < CPL; ++j) sum[j] += t[j];
        }
        store_row<RP, float>(&red[g][0], q, sum);
        __syncthreads();
        if (g == 0) {
            float t[CPL];
            load_row<RP, float>(&red[0][0], q, t);
            for (int j = 1; j < kQuads; ++j) {
                float u[CPL];
                load_row<RP, float>(&red[j][0], q, u);
#pragma unroll
                for (int k = 0; k < CPL; ++k) t[k] += u[k];
            }
            store_row<RP, float>(out + (size_t)r * RP, q, t);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
//  host-side dispatch
// ---------------------------------------------------------------------------
template <int RP, int NO, typename FT, int KIND>
static void launch_seg(dyn_gpu_ctx* c, const SegArgs& a) {
    hipLaunchKernelGGL((k_seg<RP, NO, FT, KIND>), dim3(a.nwg), dim3(kBlock), 0, c->stream, a);
}

template <int RP, typename FT, int KIND>
static int dispatch_no(dyn_gpu_ctx* c, int NO, const SegArgs& a) {
    switch (NO) {
        case 1: launch_seg<RP, 1, FT, KIND>(c, a); break;
        case 2: launch_seg<RP, 2, FT, KIND>(c, a); break;
        case 3: launch_seg<RP, 3, FT, KIND>(c, a); break;
        case 4: launch_seg<RP, 4, FT, KIND>(c, a); break;
        case 5: launch_seg<RP, 5, FT, KIND>(c, a); break;
        case 6: launch_seg<RP, 6, FT, KIND>(c, a); break;
        case 7: launch_seg<RP, 7, FT, KIND>(c, a); break;
        default: return set_err(c, DYN_ERR_UNSUPPORTED, "unsupported mode count %d", NO + 1);
    }
    return DYN_OK;
}

template <typename FT, int KIND>
static int dispatch_rp(dyn_gpu_ctx* c, int Rp, int NO, const SegArgs& a) {
    switch (Rp) {
        case 16: return dispatch_no<16, FT, KIND>(c, NO, a);
        case 32: return dispatch_no<32, FT, KIND>(c, NO, a);
        case 64: return dispatch_no<64, FT, KIND>(c, NO, a);
        case 128: return dispatch_no<128, FT, KIND>(c, NO, a);
        default: return set_err(c, DYN_ERR_UNSUPPORTED, "unsupported padded rank %d", Rp);
    }
}

static int launch_fixup(dyn_gpu_ctx* c, int Rp, const float* bpart, const uint32_t* brow,
                        uint32_t nwg, float* out) {
    const unsigned nb = (nwg + kQuads - 1) / kQuads;
    switch (Rp) {
        case 16: hipLaunchKernelGGL(k_fixup<16>, dim3(nb), dim3(kBlock), 0, c->stream, bpart, brow, nwg, out); break;
        case 32: hipLaunchKernelGGL(k_fixup<32>, dim3(nb), dim3(kBlock), 0, c->stream, bpart, brow, nwg, out); break;
        case 64: hipLaunchKernelGGL(k_fixup<64>, dim3(nb), dim3(kBlock), 0, c->stream, bpart, brow, nwg, out); break;
        case 128: hipLaunchKernelGGL(k_fixup<128>, dim3(nb), dim3(kBlock), 0, c->stream, bpart, brow, nwg, out); break;
        default: return set_err(c, DYN_ERR_UNSUPPORTED, "unsupported padded rank %d", Rp);
    }
    return check_launch(c, "fixup");
}

static int prepare_seg(dyn_gpu_ctx* c, int mode, int Rp, const void* const* Yin, SegArgs* a,
                       int* NO) {
    const ModeCopy& cp = c->copy[mode];
    if (!cp.built) return set_err(c, DYN_ERR_STATE, "mode %d copy not built", mode);
    *a = SegArgs();
    int j = 0;
    for (int w = 0; w < c->N; ++w) {
        if (w == mode) continue;
        a->oidx[j] = cp.idx[w];
        a->Yo[j] = Yin[w];
        ++j;
    }
    *NO = j;
    a->vals = cp.vals;
    a->rowptr = cp.rowptr;
    a->seg = cp.seg;
    a->nnz = (uint32_t)c->nnz;
    a->nwg = (uint32_t)((c->nnz + kChunk - 1) / kChunk);
    void* sp = nullptr;
    const size_t nwg = std::max<uint32_t>(a->nwg, 1);
    DYN_CHECK(scratch2(c, nwg * 2 * (size_t)Rp * sizeof(float) + nwg * 2 * sizeof(uint32_t) +
                              nwg * sizeof(double) + 1024,
                       &sp));
    a->bpart = (float*)sp;
    a->brow = (uint32_t*)((char*)sp + nwg * 2 * (size_t)Rp * sizeof(float));
    a->loss_part = nullptr;
    return DYN_OK;
}

int mttkrp_mode(dyn_gpu_ctx* c, int mode, int prec, const void* const* Yin,
                const float* vals_override, float* out, int Rp) {
    SegArgs a;
    int NO = 0;
    DYN_CHECK(prepare_seg(c, mode, Rp, Yin, &a, &NO));
    if (vals_override) {
        a.vals = vals_override;
        a.vperm = c->copy[mode].perm;
    }
    a.out = out;
    DYN_CHECK(fill_f32(c, out, (size_t)c->dims[mode] * Rp, 0.0f));
    if (a.nwg == 0) return DYN_OK;
    TimedPair tp;
    timing_begin(c, &tp);
    if (prec == DYN_PREC_BF16)
        DYN_CHECK((dispatch_rp<bf16, 0>(c, Rp, NO, a)));
    else
        DYN_CHECK((dispatch_rp<float, 0>(c, Rp, NO, a)));
    DYN_CHECK(check_launch(c, "mttkrp_seg"));
    timing_end(c, &tp);
    return launch_fixup(c, Rp, a.bpart, a.brow, a.nwg, out);
}

// ---------------------------------------------------------------------------
//  forward model evaluation in input order (+ optional FP64 <vals, m>)
// ---------------------------------------------------------------------------
struct FwdArgs {
    const uint32_t* idx[kMaxModes];
    const float*    A[kMaxModes];
    const double*   lambda;      // null: ones
    float*          m;           // null: not written
    const float*    vals;        // null: no dot product
    double*         dot_part;    // [nblocks]
    uint32_t        nnz;
};

template <int J, int N>
__device__ __forceinline__ void bcast_idx(const uint32_t* i_l, uint32_t* iw) {
#pragma unroll
    for (int n = 0; n < N; ++n) iw[n] = qb<J>(i_l[n]);
}

template <int RP, int N>
__global__ __launch_bounds__(kBlock) void k_forward(FwdArgs a) {
    using C = Cols<RP, float>;
    constexpr int CPL = C::CPL;
    __shared__ double red[kBlock / 64];
    const int tid = threadIdx.x;
    const int g = tid >> 2;
    const int q = tid & 3;
    float lam[CPL];
#pragma unroll
    for (int b = 0; b < C::NB; ++b)
#pragma unroll
        for (int t = 0; t < C::LW; ++t)
            lam[b * C::LW + t] = a.lambda ? (float)a.lambda[C::col(b, q, t)] : 1.f;
    double dot = 0.0;
    const uint64_t per_block = (uint64_t)kQuads * 4;
    for (uint64_t k0 = (uint64_t)blockIdx.x * per_block + (uint64_t)g * 4; k0 < a.nnz;
         k0 += (uint64_t)gridDim.x * per_block) {
        const uint64_t kk = k0 + q;
        uint32_t i_l[N];
#pragma unroll
        for (int n = 0; n < N; ++n) i_l[n] = (kk < a.nnz) ? a.idx[n][kk] : 0u;
        const float v_l = (a.vals && kk < a.nnz) ? a.vals[kk] : 0.f;
        uint32_t ib[4][N];
        bcast_idx<0, N>(i_l, ib[0]);
        bcast_idx<1, N>(i_l, ib[1]);
        bcast_idx<2, N>(i_l, ib[2]);
        bcast_idx<3, N>(i_l, ib[3]);
        float mj[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float p[CPL];
#pragma unroll
            for (int c = 0; c < CPL; ++c) p[c] = lam[c];
#pragma unroll
            for (int n = 0; n < N; ++n) {
                float y[CPL];
                load_row<RP, float>(a.A[n] + (size_t)ib[j][n] * RP, q, y);
#pragma unroll
                for (int c = 0; c < CPL; ++c) p[c] *= y[c];
            }
            float part = 0.f;
#pragma unroll
            for (int c = 0; c < CPL; ++c) part += p[c];
            mj[j] = quad_sum(part);
        }
        const float mine = (q == 0) ? mj[0] : (q == 1) ? mj[1] : (q == 2) ? mj[2] : mj[3];
        if (kk < a.nnz) {
            if (a.m) a.m[kk] = mine;
            if (a.vals) dot += (double)v_l * (double)mine;
        }
    }
    if (a.dot_part) {
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) dot += __shfl_xor(dot, o, 64);
        if ((tid & 63) == 0) red[tid >> 6] = dot;
        __syncthreads();
        if (tid == 0) {
            double t = 0.0;
            for (int w = 0; w < kBlock / 64; ++w) t += red[w];
            a.dot_part[blockIdx.x] = t;
        }
    }
}

__global__ void k_sum_parts(const double* __restrict__ part, uint32_t n, double* __restrict__ out) {
    __shared__ double red[kBlock];
    double s = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += kBlock) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < kBlock; ++i) t += red[i];
        *out = t;
    }
}

template <int RP>
static int launch_fwd_n(dyn_gpu_ctx* c, int N, unsigned nb, const FwdArgs& a) {
    switch (N) {
        case 2: hipLaunchKernelGGL((k_forward<RP, 2>), dim3(nb), dim3(kBlock), 0, c->stream, a); break;
        case 3: hipLaunchKernelGGL((k_forward<RP, 3>), dim3(nb), dim3(kBlock), 0, c->stream, a); break;
        case 4: hipLaunchKernelGGL((k_forward<RP, 4>), dim3(nb), dim3(kBlock), 0, c->stream, a); break;
        case 5: hipLaunchKernelGGL((k_forward<RP, 5>), dim3(nb), dim3(kBlock), 0, c->stream, a); break;
        case 6: hipLaunchKernelGGL((k_forward<RP, 6>), dim3(nb), dim3(kBlock), 0, c->stream, a); break;
        case 7: hipLaunchKernelGGL((k_forward<RP, 7>), dim3(nb), dim3(kBlock), 0, c->stream, a); break;
        case 8: hipLaunchKernelGGL((k_forward<RP, 8>), dim3(nb), dim3(kBlock), 0, c->stream, a); break;
        default: return set_err(c, DYN_ERR_UNSUPPORTED, "unsupported mode count %d", N);
    }
    return check_launch(c, "forward");
}

int forward_eval(dyn_gpu_ctx* c, const float* const* A, int Rp, const double* lambda,
                 float* m_out, const float* vals_dot, double* dot_out_dev) {
    FwdArgs a = FwdArgs();
    for (int n = 0; n < c->N; ++n) {
        a.idx[n] = c->idx[n];
        a.A[n] = A[n];
    }
    a.lambda = lambda;
    a.m = m_out;
    a.vals = vals_dot;
    a.nnz = (uint32_t)c->nnz;
    const unsigned nb = 4096;
    double* part = nullptr;
    if (vals_dot) {
        void* sp = nullptr;
        DYN_CHECK(scratch2(c, nb * sizeof(double), &sp));
        part = (double*)sp;
    }
    a.dot_part = part;
    int rc = DYN_OK;
    switch (Rp) {
        case 16: rc = launch_fwd_n<16>(c, c->N, nb, a); break;
        case 32: rc = launch_fwd_n<32>(c, c->N, nb, a); break;
        case 64: rc = launch_fwd_n<64>(c, c->N, nb, a); break;
        case 128: rc = launch_fwd_n<128>(c, c->N, nb, a); break;
        default: return set_err(c, DYN_ERR_UNSUPPORTED, "unsupported padded rank %d", Rp);
    }
    DYN_CHECK(rc);
    if (vals_dot) {
        hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kBlock), 0, c->stream, part, nb,
                           dot_out_dev);
        DYN_CHECK(check_launch(c, "sum_parts"));
    }
    return DYN_OK;
}

// ---------------------------------------------------------------------------
//  fused completion loss + gradient over every mode copy
// ---------------------------------------------------------------------------
int completion_grad(dyn_gpu_ctx* c, int loss, const float* const* A, int Rp,
                    float* const* grad, double* loss_out) {
    const void* Yin[kMaxModes];
    for (int n = 0; n < c->N; ++n) Yin[n] = A[n];
    double* dloss_dev = nullptr;
    for (int mode = 0; mode < c->N; ++mode) {
        SegArgs a;
        int NO = 0;
        DYN_CHECK(prepare_seg(c, mode, Rp, Yin, &a, &NO));
        a.out = grad[mode];
        a.An = A[mode];
        a.loss = loss;
        if (mode == 0 && loss_out) {
            const size_t nwg = std::max<uint32_t>(a.nwg, 1);
            uintptr_t lp = (uintptr_t)((char*)a.brow + nwg * 2 * sizeof(uint32_t));
            a.loss_part = (double*)((lp + 7) & ~(uintptr_t)7);
            dloss_dev = a.loss_part;
        }
        DYN_CHECK(fill_f32(c, grad[mode], (size_t)c->dims[mode] * Rp, 0.0f));
        if (a.nwg == 0) continue;
        DYN_CHECK((dispatch_rp<float, 1>(c, Rp, NO, a)));
        DYN_CHECK(check_launch(c, "completion_seg"));
        DYN_CHECK(launch_fixup(c, Rp, a.bpart, a.brow, a.nwg, grad[mode]));
        if (mode == 0 && loss_out) {
            void* sp = nullptr;
            DYN_CHECK(scratch(c, 64, &sp));
            hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kBlock), 0, c->stream, dloss_dev,
                               a.nwg, (double*)sp);
            DYN_CHECK(check_launch(c, "sum_parts"));
            DYN_HIP(c, hipMemcpyAsync(loss_out, sp, sizeof(double), hipMemcpyDeviceToHost,
                                      c->stream));
            DYN_HIP(c, hipStreamSynchronize(c->stream));
        }
    }
    if (loss_out && c->nnz == 0) *loss_out = 0.0;
    return DYN_OK;
}

}  // namespace dyn
