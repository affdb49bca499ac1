// Host side of the segmented spMTTKRP (the kernel is in dyn_seg.hpp): argument
// set-up and dispatch, the heavy-row combine, and the kernels that share the
// walk -- the adjoint (values replaced by a gradient vector read through the
// copy's permutation), the forward model evaluation and the fused
// sparse-completion loss/gradient.
#include "dyn_seg.hpp"

namespace dyn {

// Heavy rows split over several waves' piece ranges: sum the partials the
// waves stored (the last piece of each range inside the row, and the row's
// last piece) in slot order -- quads stride the ranges, then a butterfly
// over quads.  One wave per heavy row; rows held whole by one wave were
// stored by the segment kernel and are skipped.
template <int RP>
__global__ __launch_bounds__(kBlock) void k_heavy(const float* __restrict__ part,
                                                  const uint32_t* __restrict__ hrow,
                                                  const uint32_t* __restrict__ hfirst,
                                                  const uint32_t* __restrict__ hcount,
                                                  uint32_t nheavy, uint32_t npieces,
                                                  uint32_t nwaves, float* __restrict__ out) {
    using C = Cols<RP, float>;
    constexpr int CPL = C::CPL;
    const int lane = threadIdx.x & 63;
    const int g = lane >> 2;
    const int q = lane & 3;
    const uint32_t h = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    if (h >= nheavy) return;
    const uint32_t first = hfirst[h], last = first + hcount[h] - 1;
    const uint32_t pq = npieces / nwaves, prem = npieces % nwaves;
    const uint32_t w0 = wave_of(first, pq, prem), w1 = wave_of(last, pq, prem);
    if (w0 == w1) return;
    float sum[CPL];
#pragma unroll
    for (int j = 0; j < CPL; ++j) sum[j] = 0.f;
    for (uint32_t w = w0 + (uint32_t)g; w <= w1; w += 16) {
        const uint32_t slot = min(range_begin(w + 1, pq, prem) - 1, last);
        float t[CPL];
        load_row<RP, float>(part + (size_t)slot * RP, q, t);
#pragma unroll
        for (int j = 0; j < CPL; ++j) sum[j] += t[j];
    }
#pragma unroll
    for (int o = 4; o < 64; o <<= 1)
#pragma unroll
        for (int j = 0; j < CPL; ++j) sum[j] += __shfl_xor(sum[j], o, 64);
    if (g == 0) store_row<RP, float>(out + (size_t)hrow[h] * RP, q, sum);
}

// ---------------------------------------------------------------------------
//  host-side dispatch
// ---------------------------------------------------------------------------
// The hot split used by spMTTKRP launches on 3-mode tensors: operand 1 (the
// mode after the secondary sort mode) is split at hot_split_rows.  Tensors
// with more modes (more operands than the split serves) measured faster on
// the plain order with rank-ordered LDS rows of every operand read through
// generic loads; their BF16 launches add the all-hot class split (rows of
// kClassMin+ nonzeros lead with the nonzeros whose every gathered row has an
// LDS rank: those pieces read LDS only, by ds_read).  It measured 1.2 % faster
// at 4-mode rank 64 and 0.3 % at 5-mode rank 32: the remaining memory
// gathers still bound the launch.
static int prepare_split(dyn_gpu_ctx* c, int mode, int Rp, int prec) {
    if (c->N < 3) return DYN_OK;
    if (c->N > 3) {
        // plain (row, secondary) order: undo a 3-mode split; BF16 spMTTKRP on
        // up to 7 modes and 192 gathered columns then splits off the nonzeros
        // whose every row sits in LDS (an FP32 launch runs on such a copy as
        // on the plain order)
        const ModeCopy& cp = c->copy[mode];
        if (cp.hot_w >= 0) DYN_CHECK(remap_mode(c, mode, cp.secondary, cp.rank_key));
        // (not on streamed chunks: a chunk's copy is rebuilt every sweep, and
        // the regroup would cost more than its launch saves)
        if (prec != DYN_PREC_BF16 || c->streamed || c->N > 7 || Rp * (c->N - 1) > 192 ||
            (uint64_t)c->nnz * 4u > 0xFFFFFFF0ull)
            return DYN_OK;
        uint32_t dims_o[kMaxModes], H[kMaxModes] = {0};
        int NO = 0;
        for (int w = 0; w < c->N; ++w)
            if (w != mode) dims_o[NO++] = c->dims[w];
        plan_lds_rows(c, (size_t)Rp * sizeof(bf16), NO, dims_o, H);
        bool same = true;
        for (int j = 0; j < NO; ++j) {
            if (H[j] == 0) return DYN_OK;
            same = same && c->copy[mode].cls_h[j] == H[j];
        }
        if (same) return DYN_OK;
        // split from the plain order, so each class keeps the secondary order
        if (c->copy[mode].cls_h[0] != 0)
            DYN_CHECK(remap_mode(c, mode, cp.secondary, cp.rank_key));
        return ensure_class_split(c, mode, H);
    }
    int others[kMaxModes], j = 0;
    for (int w = 0; w < c->N; ++w)
        if (w != mode) others[j++] = w;
    return ensure_hot_split(c, mode, others[1],
                            hot_split_rows(c, Rp, c->dims[others[1]]));
}

// compiled in dyn_mttkrp_big.hip
extern template int dispatch_rp<float, 0, true>(dyn_gpu_ctx*, int, int, SegArgs&, const uint32_t*);
extern template int dispatch_rp<bf16, 0, true>(dyn_gpu_ctx*, int, int, SegArgs&, const uint32_t*);
extern template int dispatch_rp<float, 1, true>(dyn_gpu_ctx*, int, int, SegArgs&, const uint32_t*);

// Columns of 4 GiB or more take the BIG instantiations.
template <typename FT, int KIND>
static int seg_dispatch(dyn_gpu_ctx* c, int Rp, int NO, SegArgs& a, const uint32_t* d) {
    return (uint64_t)a.nnz * 4u > 0xFFFFFFF0ull ? dispatch_rp<FT, KIND, true>(c, Rp, NO, a, d)
                                             : dispatch_rp<FT, KIND, false>(c, Rp, NO, a, d);
}

static int launch_heavy(dyn_gpu_ctx* c, int Rp, const ModeCopy& cp, const float* part,
                        uint32_t nwaves, float* out) {
    if (cp.nheavy == 0) return DYN_OK;
#define DYN_HEAVY(RPV)                                                                        \
    hipLaunchKernelGGL(k_heavy<RPV>, dim3((cp.nheavy + kBlock / 64 - 1) / (kBlock / 64)),  \
                       dim3(kBlock), 0, c->stream, part, cp.heavy_row, cp.heavy_first,       \
                       cp.heavy_count, cp.nheavy, cp.npieces, nwaves, out)
    switch (Rp) {
        case 16: DYN_HEAVY(16); break;
        case 32: DYN_HEAVY(32); break;
        case 64: DYN_HEAVY(64); break;
        case 128: DYN_HEAVY(128); break;
        default: return set_err(c, DYN_ERR_UNSUPPORTED, "unsupported padded rank %d", Rp);
    }
#undef DYN_HEAVY
    return check_launch(c, "heavy_rows");
}

size_t seg_part_bytes(const dyn_gpu_ctx* c, int mode, int Rp) {
    const size_t pieces = std::max<uint32_t>(c->copy[mode].npieces, 1);
    return ((pieces * Rp * sizeof(float)) + 255) & ~size_t(255);
}

bool split_ready(const dyn_gpu_ctx* c, int mode, int Rp) {
    const ModeCopy& cp = c->copy[mode];
    if (!cp.built) return false;
    if (c->N < 3) return true;
    if (c->N > 3) return cp.hot_w < 0;
    int others[kMaxModes], j = 0;
    for (int w = 0; w < c->N; ++w)
        if (w != mode) others[j++] = w;
    return cp.hot_w == others[1] &&
           cp.hot_h == hot_split_rows(const_cast<dyn_gpu_ctx*>(c), Rp, c->dims[others[1]]);
}

// Segment-kernel arguments of copy `mode`.  The piece partials go to
// `part_buf` when given (at least seg_part_bytes(c, mode, Rp) bytes, no loss
// slots), else to the context's second scratch buffer.
static int prepare_seg(dyn_gpu_ctx* c, int mode, int Rp, const void* const* Yin, SegArgs* a,
                       int* NO, uint32_t* dims_o, bool loss_slots, float* part_buf = nullptr) {
    const ModeCopy& cp = c->copy[mode];
    if (!cp.built) return set_err(c, DYN_ERR_STATE, "mode %d copy not built", mode);
    *a = SegArgs();
    int j = 0;
    for (int w = 0; w < c->N; ++w) {
        if (w == mode) continue;
        a->oidx[j] = cp.idx[w];
        a->Yo[j] = Yin[w];
        dims_o[j] = c->dims[w];
        ++j;
    }
    *NO = j;
    a->vals = cp.vals;
    a->rowptr = cp.rowptr;
    a->seg_k = cp.seg_k;
    a->seg_e = cp.seg_e;
    a->seg_row = cp.seg_row;
    a->nseg = cp.nseg;
    a->pc_k = cp.pc_k;
    a->pc_e = cp.pc_e;
    a->pc_row = cp.pc_row;
    a->pc_hot = cp.pc_hot;
    for (int w = 0; w + 1 < kMaxModes; ++w) a->cls_h[w] = cp.cls_h[w];
    a->npieces = cp.npieces;
    a->empty_rows = cp.empty_rows;
    a->nempty = cp.nempty;
    a->stream_bytes = (uint32_t)std::min<uint64_t>(c->nnz * 4ull, 0xFFFFFFF0ull);
    a->nnz = (uint32_t)c->nnz;
    if (part_buf) {
        a->part = part_buf;
        a->loss_part = nullptr;
        a->loss_dev = nullptr;
        return DYN_OK;
    }
    void* sp = nullptr;
    const size_t pbytes = seg_part_bytes(c, mode, Rp);
    const size_t nloss = (size_t)cp.npieces + cp.nseg;
    const size_t lbytes = loss_slots ? ((std::max<size_t>(nloss, 1) * sizeof(double) + 255) & ~size_t(255)) : 0;
    DYN_CHECK(scratch2(c, pbytes + lbytes + (1 + kMaxChunks) * sizeof(double) + 256, &sp));
    a->part = (float*)sp;
    a->loss_part = loss_slots ? (double*)((char*)sp + pbytes) : nullptr;
    a->loss_dev = (double*)((char*)sp + pbytes + lbytes);
    return DYN_OK;
}

int mttkrp_mode(dyn_gpu_ctx* c, int mode, int prec, const void* const* Yin,
                const float* vals_override, float* out, int Rp, float* part_buf) {
    SegArgs a;
    int NO = 0;
    uint32_t dims_o[kMaxModes];
    if (!c->copy[mode].built) return set_err(c, DYN_ERR_STATE, "mode %d copy not built", mode);
    if (part_buf) {
        // caller-provided partials: the copy must already be in launch order
        // (re-sorting it would allocate)
        if (!split_ready(c, mode, Rp))
            return set_err(c, DYN_ERR_STATE, "mode %d copy is not prepared for this launch", mode);
    } else {
        DYN_CHECK(prepare_split(c, mode, Rp, prec));
    }
    DYN_CHECK(prepare_seg(c, mode, Rp, Yin, &a, &NO, dims_o, false, part_buf));
    if (vals_override) {
        a.vals = vals_override;
        a.vperm = c->copy[mode].perm;
    }
    a.out = out;
    // the kernel writes every row; with no nonzeros at all it is not launched
    if (a.nseg == 0 && a.npieces == 0)
        DYN_CHECK(fill_f32(c, out, (size_t)c->dims[mode] * Rp, 0.0f));
    // timing events (the CP-ALS loop's pair, or a pool pair) are recorded by
    // the launch itself
    TimedPair tp;
    const bool pooled = !c->seg_pair && !part_buf && timing_acquire(c, &tp);
    if (c->seg_pair) {
        c->launch_ev[0] = c->seg_pair[0];
        c->launch_ev[1] = c->seg_pair[1];
    } else if (pooled) {
        c->launch_ev[0] = tp.a;
        c->launch_ev[1] = tp.b;
    }
    int rc = DYN_OK;
    if (a.nseg != 0 || a.npieces != 0) {
        if (prec == DYN_PREC_BF16) rc = seg_dispatch<bf16, 0>(c, Rp, NO, a, dims_o);
        else rc = seg_dispatch<float, 0>(c, Rp, NO, a, dims_o);
        if (rc == DYN_OK) rc = check_launch(c, "mttkrp_seg");
    } else if (c->launch_ev[0]) {
        (void)hipEventRecord(c->launch_ev[0], c->stream);
        (void)hipEventRecord(c->launch_ev[1], c->stream);
    }
    c->launch_ev[0] = c->launch_ev[1] = nullptr;
    if (pooled) {
        if (rc == DYN_OK) c->timed.push_back(tp);
        else c->event_pool.push_back(tp);
    }
    DYN_CHECK(rc);
    if (a.nseg == 0 && a.npieces == 0) return DYN_OK;
    return launch_heavy(c, Rp, c->copy[mode], a.part, a.nwaves, out);
}

// Relabelled copies (rows in frequency-rank order) of caller factors A[n]
// (input order, Rp-strided), in the first scratch buffer; mode `skip` is not
// gathered and gets none.
static int relabeled_operands(dyn_gpu_ctx* c, const float* const* A, int Rp, int skip,
                              const void** Yin) {
    DYN_CHECK(ensure_relabel(c));
    size_t total = 0;
    for (int n = 0; n < c->N; ++n) total += (size_t)c->dims[n] * Rp;
    void* sp = nullptr;
    DYN_CHECK(scratch(c, total * sizeof(float) + 256 * kMaxModes, &sp));
    char* p = (char*)sp;
    for (int n = 0; n < c->N; ++n) {
        if (n == skip) {
            Yin[n] = nullptr;
            continue;
        }
        float* dst = (float*)p;
        DYN_CHECK(relabel_rows(c, A[n], c->relab[n], dst, nullptr, c->dims[n], Rp));
        Yin[n] = dst;
        p += (((size_t)c->dims[n] * Rp * sizeof(float)) + 255) & ~size_t(255);
    }
    return DYN_OK;
}

int adjoint_mode(dyn_gpu_ctx* c, int mode, const float* const* A, int Rp, const float* g,
                 float* out) {
    const void* Yin[kMaxModes];
    DYN_CHECK(relabeled_operands(c, A, Rp, mode, Yin));
    return mttkrp_mode(c, mode, DYN_PREC_FP32, Yin, g, out, Rp);
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
    // lambda stays FP64: a diverging CP-ALS drives it past the FP32 range,
    // and the model value is only formed after the FP32 factor products
    double lam[CPL];
#pragma unroll
    for (int b = 0; b < C::NB; ++b)
#pragma unroll
        for (int t = 0; t < C::LW; ++t)
            lam[b * C::LW + t] = a.lambda ? a.lambda[C::col(b, q, t)] : 1.0;
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
        double mj[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float p[CPL];
            load_row<RP, float>(a.A[0] + (size_t)ib[j][0] * RP, q, p);
#pragma unroll
            for (int n = 1; n < N; ++n) {
                float y[CPL];
                load_row<RP, float>(a.A[n] + (size_t)ib[j][n] * RP, q, y);
#pragma unroll
                for (int c = 0; c < CPL; ++c) p[c] *= y[c];
            }
            double part = 0.0;
#pragma unroll
            for (int c = 0; c < CPL; ++c) part = fma(lam[c], (double)p[c], part);
            part += __shfl_xor(part, 1, 64);
            part += __shfl_xor(part, 2, 64);
            mj[j] = part;
        }
        const double mine = (q == 0) ? mj[0] : (q == 1) ? mj[1] : (q == 2) ? mj[2] : mj[3];
        if (kk < a.nnz) {
            if (a.m) a.m[kk] = (float)mine;
            if (a.vals) dot += (double)v_l * mine;
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

// Block b sums part[b * chunk, min((b + 1) * chunk, n)) in a fixed order
// (thread-strided, then the threads in order) into out[b].
__global__ __launch_bounds__(kBlock) void k_sum_chunks(const double* __restrict__ part, uint32_t n,
                                                       uint32_t chunk, double* __restrict__ out) {
    __shared__ double red[kBlock];
    const uint64_t b0 = (uint64_t)blockIdx.x * chunk;
    const uint64_t e = std::min<uint64_t>(b0 + chunk, n);
    double s = 0.0;
    for (uint64_t i = b0 + threadIdx.x; i < e; i += kBlock) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < kBlock; ++i) t += red[i];
        out[blockIdx.x] = t;
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

// Deterministic FP64 sum of n device values into *out: per-chunk partials
// (one workgroup each, up to kMaxChunks of them) then one workgroup over the
// partials.  `tmp` holds at least kMaxChunks doubles and must not alias part.
static int sum_doubles(dyn_gpu_ctx* c, const double* part, uint32_t n, double* tmp, double* out) {
    const uint32_t chunk = std::max<uint32_t>(kBlock * 8, (n + kMaxChunks - 1) / kMaxChunks);
    const uint32_t nb = std::max<uint32_t>(1, (n + chunk - 1) / chunk);
    hipLaunchKernelGGL(k_sum_chunks, dim3(nb), dim3(kBlock), 0, c->stream, part, n, chunk, tmp);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kBlock), 0, c->stream, tmp, nb, out);
    return check_launch(c, "sum_doubles");
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
    DYN_CHECK(relabeled_operands(c, A, Rp, -1, Yin));
    if (loss_out) *loss_out = 0.0;
    for (int mode = 0; mode < c->N; ++mode) {
        SegArgs a;
        int NO = 0;
        uint32_t dims_o[kMaxModes];
        const bool want_loss = (mode == 0 && loss_out);
        if (!c->copy[mode].built) return set_err(c, DYN_ERR_STATE, "mode %d copy not built", mode);
        DYN_CHECK(prepare_split(c, mode, Rp, DYN_PREC_FP32));
        DYN_CHECK(prepare_seg(c, mode, Rp, Yin, &a, &NO, dims_o, want_loss));
        a.out = grad[mode];
        a.An = A[mode];
        a.loss = loss;
        if (a.nseg == 0 && a.npieces == 0) {
            DYN_CHECK(fill_f32(c, grad[mode], (size_t)c->dims[mode] * Rp, 0.0f));
            continue;
        }
        DYN_CHECK((seg_dispatch<float, 1>(c, Rp, NO, a, dims_o)));
        DYN_CHECK(check_launch(c, "completion_seg"));
        DYN_CHECK(launch_heavy(c, Rp, c->copy[mode], a.part, a.nwaves, grad[mode]));
        if (want_loss) {
            double* d = a.loss_dev;
            DYN_CHECK(sum_doubles(c, a.loss_part, a.npieces + a.nseg, d + 1, d));
            DYN_HIP(c, hipMemcpyAsync(loss_out, d, sizeof(double), hipMemcpyDeviceToHost,
                                      c->stream));
            DYN_HIP(c, hipStreamSynchronize(c->stream));
        }
    }
    return DYN_OK;
}

}  // namespace dyn
