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
//   * a lane group of G = Rp/4 lanes owns one nonzero at a time; lane q
//     holds columns [4q, 4q+4) of the row as a float4 (BF16 operands: 4
//     bf16 = 8 bytes per lane, widened to FP32),
//   * each group walks kSeg consecutive nonzeros of the sorted copy; the
//     value and index columns are loaded G at a time (one element per lane,
//     contiguous) and broadcast inside the group with lane shuffles,
//   * the output row is accumulated in registers while the row pointer says
//     the row continues; a row that starts and ends inside the group's range
//     is stored directly,
//   * rows cut by a group boundary leave a partial in LDS; one workgroup-level
//     pass sums partials of the same row in group order; rows cut by a
//     workgroup boundary leave a partial in a global table that a fix-up
//     kernel sums in workgroup order.  No atomics: results are bitwise
//     reproducible run to run.
#include "dyn_internal.hpp"

#include <algorithm>

namespace dyn {

constexpr uint32_t kNone = 0xFFFFFFFFu;

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

__device__ __forceinline__ float4 ld4(const float* p) {
    return *reinterpret_cast<const float4*>(p);
}
__device__ __forceinline__ float4 ld4(const bf16* p) {
    const uint2 u = *reinterpret_cast<const uint2*>(p);
    float4 r;
    r.x = __uint_as_float(u.x << 16);
    r.y = __uint_as_float(u.x & 0xFFFF0000u);
    r.z = __uint_as_float(u.y << 16);
    r.w = __uint_as_float(u.y & 0xFFFF0000u);
    return r;
}
__device__ __forceinline__ void st4(float* p, float4 v) { *reinterpret_cast<float4*>(p) = v; }
__device__ __forceinline__ float4 f4(float a) { return make_float4(a, a, a, a); }
__device__ __forceinline__ float4 add4(float4 a, float4 b) {
    return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}
__device__ __forceinline__ float4 mul4(float4 a, float4 b) {
    return make_float4(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w);
}
__device__ __forceinline__ float4 fma4(float4 a, float4 b, float4 c) {
    return make_float4(__fmaf_rn(a.x, b.x, c.x), __fmaf_rn(a.y, b.y, c.y),
                       __fmaf_rn(a.z, b.z, c.z), __fmaf_rn(a.w, b.w, c.w));
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

// KIND 0: spMTTKRP / adjoint.  KIND 1: fused completion loss + gradient.
template <int RP, int NO, typename FT, int KIND>
__global__ __launch_bounds__(kBlock) void k_seg(SegArgs a) {
    constexpr int G = RP / 4;
    constexpr int NG = kBlock / G;
    constexpr uint32_t CH = (uint32_t)NG * kSeg;
    __shared__ float4 pacc[2 * NG][G];
    __shared__ uint32_t prow[2 * NG];
    __shared__ double lred[kBlock / 64];

    const int tid = threadIdx.x;
    const int g = tid / G;
    const int q = tid % G;
    const uint32_t wg = xcd_chunk(blockIdx.x, a.nwg);
    const uint32_t wg_s = wg * CH;
    const uint32_t wg_e = min(wg_s + CH, a.nnz);

    if (tid < 2 * NG) prow[tid] = kNone;
    if (tid < 2) a.brow[2 * wg + tid] = kNone;
    __syncthreads();

    const uint32_t s = wg_s + (uint32_t)g * kSeg;
    const uint32_t e = min(s + (uint32_t)kSeg, a.nnz);
    double lsum = 0.0;

    if (s < e) {
        uint32_t row = a.seg[s / kSeg];
        uint32_t rbeg = a.rowptr[row];
        uint32_t rend = a.rowptr[row + 1];
        float4 acc = f4(0.f);
        float4 arow = f4(0.f);
        if (KIND == 1) arow = ld4(a.An + (size_t)row * RP + q * 4);

        auto flush = [&](uint32_t r, uint32_t rb, uint32_t re) {
            const bool os = rb < s, oe = re > e;
            if (!os && !oe) {
                st4(a.out + (size_t)r * RP + q * 4, acc);
            } else {
                const int slot = 2 * g + (os ? 0 : 1);
                pacc[slot][q] = acc;
                if (q == 0) prow[slot] = r;
            }
        };

        for (uint32_t k0 = s; k0 < e; k0 += G) {
            const uint32_t kk = k0 + q;
            float v_l = 0.f;
            uint32_t i_l[NO];
            if (kk < e) {
                v_l = a.vperm ? a.vals[a.vperm[kk]] : a.vals[kk];
#pragma unroll
                for (int w = 0; w < NO; ++w) i_l[w] = a.oidx[w][kk];
            } else {
#pragma unroll
                for (int w = 0; w < NO; ++w) i_l[w] = 0;
            }
            const uint32_t cnt = min((uint32_t)G, e - k0);
#pragma unroll 4
            for (uint32_t j = 0; j < cnt; ++j) {
                const uint32_t k = k0 + j;
                if (k >= rend) {
                    flush(row, rbeg, rend);
                    acc = f4(0.f);
                    do {
                        ++row;
                        rbeg = rend;
                        rend = a.rowptr[row + 1];
                    } while (k >= rend);
                    if (KIND == 1) arow = ld4(a.An + (size_t)row * RP + q * 4);
                }
                const float v = __shfl(v_l, (int)j, G);
                uint32_t iw[NO];
#pragma unroll
                for (int w = 0; w < NO; ++w) iw[w] = __shfl(i_l[w], (int)j, G);
                const FT* y0p = reinterpret_cast<const FT*>(a.Yo[0]) + (size_t)iw[0] * RP + q * 4;
                if (KIND == 0) {
                    if (NO == 1) {
                        acc = fma4(f4(v), ld4(y0p), acc);
                    } else {
                        float4 p = mul4(f4(v), ld4(y0p));
#pragma unroll
                        for (int w = 1; w < NO - 1; ++w)
                            p = mul4(p, ld4(reinterpret_cast<const FT*>(a.Yo[w]) +
                                            (size_t)iw[w] * RP + q * 4));
                        acc = fma4(p, ld4(reinterpret_cast<const FT*>(a.Yo[NO - 1]) +
                                          (size_t)iw[NO - 1] * RP + q * 4),
                                   acc);
                    }
                } else {
                    float4 p = ld4(y0p);
#pragma unroll
                    for (int w = 1; w < NO; ++w)
                        p = mul4(p, ld4(reinterpret_cast<const FT*>(a.Yo[w]) +
                                        (size_t)iw[w] * RP + q * 4));
                    const float4 t = mul4(p, arow);
                    float m = (t.x + t.y) + (t.z + t.w);
#pragma unroll
                    for (int o = 1; o < G; o <<= 1) m += __shfl_xor(m, o, G);
                    const float gr = dloss(a.loss, m, v);
                    acc = fma4(f4(gr), p, acc);
                    if (a.loss_part && q == 0) lsum += lossval(a.loss, m, v);
                }
            }
        }
        flush(row, rbeg, rend);
    }
    __syncthreads();

    // workgroup-level combine of the partial rows, in group order
    for (int half = 0; half < 2; ++half) {
        const int slot = 2 * g + half;
        const uint32_t r = prow[slot];
        if (r == kNone) continue;
        int p = slot - 1;
        while (p >= 0 && prow[p] == kNone) --p;
        if (p >= 0 && prow[p] == r) continue;   // not the first partial of its row
        float4 sum = pacc[slot][q];
        for (int t = slot + 1; t < 2 * NG; ++t) {
            const uint32_t rt = prow[t];
            if (rt == kNone) continue;
            if (rt != r) break;
            sum = add4(sum, pacc[t][q]);
        }
        const bool os = a.rowptr[r] < wg_s;
        const bool oe = a.rowptr[r + 1] > wg_e;
        if (!os && !oe) {
            st4(a.out + (size_t)r * RP + q * 4, sum);
        } else {
            const int bs = os ? 0 : 1;
            st4(a.bpart + ((size_t)wg * 2 + bs) * RP + q * 4, sum);
            if (q == 0) a.brow[2 * wg + bs] = r;
        }
    }

    if (KIND == 1 && a.loss_part) {
        // deterministic block sum of the per-group loss (lane q == 0 holds it)
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
// matches.  One workgroup per opening partial sums the chain in a fixed
// order (strided by lane group, then an ordered LDS reduction).
template <int RP>
__global__ __launch_bounds__(kBlock) void k_fixup(const float* __restrict__ bpart,
                                                  const uint32_t* __restrict__ brow,
                                                  uint32_t nwg, float* __restrict__ out) {
    constexpr int G = RP / 4;
    constexpr int NG = kBlock / G;
    __shared__ float4 red[NG][G];
    const uint32_t b = blockIdx.x;
    const uint32_t r = brow[2 * b + 1];
    if (r == kNone) return;
    const int tid = threadIdx.x;
    const int g = tid / G;
    const int q = tid % G;
    float4 sum = f4(0.f);
    if (g == 0) sum = ld4(bpart + ((size_t)b * 2 + 1) * RP + q * 4);
    for (uint32_t c = b + 1 + g; c < nwg; c += NG) {
        if (brow[2 * c] != r) break;
        sum = add4(sum, ld4(bpart + ((size_t)c * 2) * RP + q * 4));
    }
    red[g][q] = sum;
    __syncthreads();
    if (g == 0) {
        float4 t = red[0][q];
        for (int j = 1; j < NG; ++j) t = add4(t, red[j][q]);
        st4(out + (size_t)r * RP + q * 4, t);
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
    switch (Rp) {
        case 16: hipLaunchKernelGGL(k_fixup<16>, dim3(nwg), dim3(kBlock), 0, c->stream, bpart, brow, nwg, out); break;
        case 32: hipLaunchKernelGGL(k_fixup<32>, dim3(nwg), dim3(kBlock), 0, c->stream, bpart, brow, nwg, out); break;
        case 64: hipLaunchKernelGGL(k_fixup<64>, dim3(nwg), dim3(kBlock), 0, c->stream, bpart, brow, nwg, out); break;
        case 128: hipLaunchKernelGGL(k_fixup<128>, dim3(nwg), dim3(kBlock), 0, c->stream, bpart, brow, nwg, out); break;
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
    const uint32_t G = (uint32_t)Rp / 4;
    const uint32_t CH = (kBlock / G) * kSeg;
    a->vals = cp.vals;
    a->rowptr = cp.rowptr;
    a->seg = cp.seg;
    a->nnz = (uint32_t)c->nnz;
    a->nwg = (uint32_t)((c->nnz + CH - 1) / CH);
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

template <int RP, int N>
__global__ __launch_bounds__(kBlock) void k_forward(FwdArgs a) {
    constexpr int G = RP / 4;
    constexpr int NG = kBlock / G;
    __shared__ double red[kBlock / 64];
    const int tid = threadIdx.x;
    const int g = tid / G;
    const int q = tid % G;
    float4 lam = f4(1.f);
    if (a.lambda) {
        lam = make_float4((float)a.lambda[q * 4 + 0], (float)a.lambda[q * 4 + 1],
                          (float)a.lambda[q * 4 + 2], (float)a.lambda[q * 4 + 3]);
    }
    double dot = 0.0;
    const uint64_t per_block = (uint64_t)NG * G;   // nonzeros per block iteration
    for (uint64_t k0 = (uint64_t)blockIdx.x * per_block + (uint64_t)g * G; k0 < a.nnz;
         k0 += (uint64_t)gridDim.x * per_block) {
        const uint64_t kk = k0 + q;
        uint32_t i_l[N];
#pragma unroll
        for (int n = 0; n < N; ++n) i_l[n] = (kk < a.nnz) ? a.idx[n][kk] : 0u;
        float v_l = (a.vals && kk < a.nnz) ? a.vals[kk] : 0.f;
        const uint32_t cnt = (uint32_t)min((uint64_t)G, a.nnz - k0);
        float mine = 0.f;
        for (uint32_t j = 0; j < cnt; ++j) {
            float4 p = lam;
#pragma unroll
            for (int n = 0; n < N; ++n) {
                const uint32_t ii = __shfl(i_l[n], (int)j, G);
                p = mul4(p, ld4(a.A[n] + (size_t)ii * RP + q * 4));
            }
            float m = (p.x + p.y) + (p.z + p.w);
#pragma unroll
            for (int o = 1; o < G; o <<= 1) m += __shfl_xor(m, o, G);
            if ((uint32_t)q == j) mine = m;
        }
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
    const unsigned nb = 2048;
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
    double* dloss = nullptr;
    for (int mode = 0; mode < c->N; ++mode) {
        SegArgs a;
        int NO = 0;
        DYN_CHECK(prepare_seg(c, mode, Rp, Yin, &a, &NO));
        a.out = grad[mode];
        a.An = A[mode];
        a.loss = loss;
        if (mode == 0 && loss_out) {
            const size_t nwg = std::max<uint32_t>(a.nwg, 1);
            a.loss_part = (double*)((char*)a.brow + nwg * 2 * sizeof(uint32_t));
            a.loss_part = (double*)(((uintptr_t)a.loss_part + 7) & ~(uintptr_t)7);
            dloss = a.loss_part;
        }
        DYN_CHECK(fill_f32(c, grad[mode], (size_t)c->dims[mode] * Rp, 0.0f));
        if (a.nwg == 0) continue;
        DYN_CHECK((dispatch_rp<float, 1>(c, Rp, NO, a)));
        DYN_CHECK(check_launch(c, "completion_seg"));
        DYN_CHECK(launch_fixup(c, Rp, a.bpart, a.brow, a.nwg, grad[mode]));
        if (mode == 0 && loss_out) {
            void* sp = nullptr;
            DYN_CHECK(scratch(c, 64, &sp));
            hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kBlock), 0, c->stream, dloss, a.nwg,
                               (double*)sp);
            DYN_CHECK(check_launch(c, "sum_parts"));
            DYN_HIP(c, hipMemcpyAsync(loss_out, sp, sizeof(double), hipMemcpyDeviceToHost,
                                      c->stream));
            DYN_HIP(c, hipStreamSynchronize(c->stream));
        }
    }
    return DYN_OK;
}

}  // namespace dyn
