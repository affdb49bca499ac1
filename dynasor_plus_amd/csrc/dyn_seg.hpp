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
//   * persistent workgroups (one per CU) first process the heavy-row pieces:
//     every wave owns a contiguous range of pieces; per piece it loads the
//     piece's values and index columns up front (kPieceLane per lane), quad g
//     takes nonzeros [16 g, 16 g + 16) in rounds of 4 (DPP quad broadcasts),
//     and the wave keeps a running row sum that it reduces across its 16
//     quads and stores once per row it touches.  A second kernel sums each
//     heavy row's stored partials in slot order,
//   * then the light segments, one quad each, walked in batches of 4 (next
//     batch prefetched), rows stored as they complete,
//   * factor rows are gathered by frequency rank: the hottest rows of the
//     gathered operands are staged in LDS once per workgroup, the rest come
//     from L2 / HBM; the secondary sort mode's row is reused from registers
//     while consecutive nonzeros of a quad share it.
//   No atomics on data: results are bitwise reproducible run to run.
//
// This header holds the kernel and its launch dispatch, templated on the
// column size class so that two translation units compile the two classes.
#pragma once
#include "dyn_internal.hpp"

#include <hip/hip_ext.h>

#include <algorithm>
#include <atomic>
#include <type_traits>

namespace dyn {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kBufDword3 = 0x00020000;   // raw buffer resource word 3 on gfx9-family parts
// cache policy of the stream loads: non-temporal (each value and index is
// read once), so the stream does not push gathered factor rows out of L2
constexpr int kStreamPolicy = 2;
constexpr int kQuads = kBlock / 4;            // quads per 256-thread block
constexpr uint32_t kMaxChunks = 2048;         // partials of the two-stage FP64 sums

struct SegArgs {
    const float*    vals;         // copy values, or the override vector
    const uint32_t* vperm;        // non-null: value of position k is vals[vperm[k]]
    const uint32_t* oidx[kMaxModes - 1];   // frequency-ranked indices of the other modes
    const void*     Yo[kMaxModes - 1];     // gathered operands, rows in rank order
    uint32_t        hot[kMaxModes - 1];    // leading rows of each operand staged in LDS
    const uint32_t* rowptr;
    const uint32_t* seg_k;        // light segments
    const uint32_t* seg_e;
    const uint32_t* seg_row;
    uint32_t        nseg;
    const uint32_t* pc_k;         // heavy-row pieces
    const uint32_t* pc_e;
    const uint32_t* pc_row;
    uint32_t        npieces;
    float*          out;
    float*          part;         // [npieces][RP]
    double*         loss_dev;     // completion: summed loss (device scalar)
    uint32_t        nwaves;       // waves of the launch (set by the launcher)
    uint32_t        stream_bytes; // bytes of one value / index column (buffer range)
    const uint32_t* pc_hot;       // hot split of operand 1 (null: none)
    const uint32_t* empty_rows;   // rows without nonzeros, zeroed by the kernel
    uint32_t        nempty;
    // completion kernel only
    const float*    An;           // output-mode factor rows (FP32, input order)
    double*         loss_part;    // [npieces + nseg] or null
    int             loss;
    uint32_t        nnz;          // nonzeros of the copy (value / index column length)
    uint32_t        allhot;       // bit w: every row of operand w is staged in LDS
    uint32_t        cls_h[kMaxModes - 1];   // the copy's all-hot class thresholds (0: none)
    uint32_t        cls_on;       // pc_hot marks pieces whose every gather is an LDS row
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
    static constexpr int RD = RP * (int)sizeof(FT) / 16;   // raw 32-bit words per lane
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

// A lane's part of a row as raw 32-bit words (no conversion), and its
// conversion to FP32 columns in load_row's order.  Gathers load raw words so
// that a conditional gather needs no wait inside its branch.
template <int RP, typename FT>
__device__ __forceinline__ void load_raw(const FT* __restrict__ lane_row, uint32_t* r) {
    using C = Cols<RP, FT>;
#pragma unroll
    for (int b = 0; b < C::NB; ++b) {
        if constexpr (C::LW * sizeof(FT) == 16) {
            const uint4 x = *reinterpret_cast<const uint4*>(lane_row + b * 4 * C::LW);
            r[b * 4 + 0] = x.x;
            r[b * 4 + 1] = x.y;
            r[b * 4 + 2] = x.z;
            r[b * 4 + 3] = x.w;
        } else {
            const uint2 x = *reinterpret_cast<const uint2*>(lane_row + b * 4 * C::LW);
            r[b * 2 + 0] = x.x;
            r[b * 2 + 1] = x.y;
        }
    }
}

template <int RP, typename FT>
__device__ __forceinline__ void widen_raw(const uint32_t* r, float* v) {
    using C = Cols<RP, FT>;
#pragma unroll
    for (int i = 0; i < C::RD; ++i) {
        if constexpr (sizeof(FT) == 4) v[i] = __uint_as_float(r[i]);
        else widen(r[i], v[2 * i], v[2 * i + 1]);
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

// The loss derivative and value of one nonzero, in FP32 (m itself is an FP32
// sum); the values are summed in FP64.  The sigmoid takes the hardware
// reciprocal (1 ulp) and the softplus its stable form max(m, 0) +
// log(1 + e^-|m|) with the hardware exp and log (a few ulp; where 1 + e^-|m|
// rounds to 1 the term is below 6e-8 and is dropped): an FP64 softplus per
// nonzero made the launch that sums the loss 1.8x slower than the other two,
// and the library log1pf still cost it three times the VALU of the others.
__device__ __forceinline__ float dloss(int kind, float m, float y) {
    if (kind == DYN_LOSS_BERNOULLI) return __builtin_amdgcn_rcpf(1.0f + __expf(-m)) - y;
    return m - y;
}
__device__ __forceinline__ float lossval(int kind, float m, float y) {
    if (kind == DYN_LOSS_BERNOULLI) return fmaxf(m, 0.0f) + __logf(1.0f + __expf(-fabsf(m))) - y * m;
    const float d = m - y;
    return 0.5f * d * d;
}

// Register plan of the segment kernel.  A quad holds per gathered row CPL =
// RP/4 floats per lane; HB nonzeros have their gathers in flight together; a
// piece's stream takes kPieceLane * (NO + 1) registers, its broadcast round
// 4 * (NO + 1); the completion kernel adds the output row and the FP64 loss.
// The workgroup is as wide as the register file allows (16 waves need <= 128
// VGPRs per lane, 8 waves <= 256; 4 waves only add AGPR spill room).
template <int RP, int NO, int KIND>
constexpr int seg_regs(int hb) {
    return hb * NO * (RP / 4) + (2 + KIND) * (RP / 4) + (int)(kPieceLane + 4) * (NO + 1) + 24 +
           64 * KIND;
}
template <int RP, int NO, int KIND>
constexpr int seg_threads() {
    return seg_regs<RP, NO, KIND>(1) <= 116 ? 1024 : (seg_regs<RP, NO, KIND>(1) <= 240 ? 512 : 256);
}
// Nonzeros whose gathers are in flight together, counted with the raw row
// words of the operand precision (BF16 rows take half the registers of FP32
// ones: 5-mode rank-32 and 4-mode rank-64 BF16 launches go from 1 and 2 to 2
// and 4 nonzeros in flight, 3 % and 5 % faster).
template <int RP, int NO, typename FT, int KIND>
constexpr int seg_hoist() {
    constexpr int budget = seg_threads<RP, NO, KIND>() == 1024 ? 128 : 240;
    constexpr int rows = NO * Cols<RP, FT>::RD;
    constexpr int rest = seg_regs<RP, NO, KIND>(0);
    return rest + 4 * rows <= budget ? 4 : (rest + 2 * rows <= budget ? 2 : 1);
}

// Heavy pieces: rounds of 4 nonzeros per quad whose gathers are in flight
// together -- two when the raw words of 8 nonzeros fit beside the rest.
template <int RP, int NO, typename FT, int KIND>
constexpr int seg_rounds() {
    return (KIND == 0 && seg_hoist<RP, NO, float, KIND>() == 4 &&
            seg_threads<RP, NO, KIND>() == 1024 &&
            8 * NO * Cols<RP, FT>::RD + seg_regs<RP, NO, KIND>(4) - 4 * NO * (RP / 4) <= 128)
               ? 2
               : 1;
}

template <int J, int NO>
__device__ __forceinline__ void bcast_batch(float v_l, const uint32_t* i_l, float& v, uint32_t* iw) {
    v = qbf<J>(v_l);
#pragma unroll
    for (int w = 0; w < NO; ++w) iw[w] = qb<J>(i_l[w]);
}

// Round T of a heavy piece: nonzero j of the round is component j of lane T.
template <int T, int NO>
__device__ __forceinline__ void bcast_round(const float* sv4, const uint32_t (*si4)[4], float* vb,
                                            uint32_t (*ib)[NO]) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        vb[j] = qbf<T>(sv4[j]);
#pragma unroll
        for (int w = 0; w < NO; ++w) ib[j][w] = qb<T>(si4[w][j]);
    }
}

// The same load from a known address space: LDS (ds_read) or global memory.
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));

template <int RP, typename FT, int AS>
__device__ __forceinline__ void load_raw_as(const FT* lane_row, uint32_t* r) {
    using C = Cols<RP, FT>;
    typedef __attribute__((address_space(AS))) const u32x4_t v4p;
    typedef __attribute__((address_space(AS))) const u32x2_t v2p;
#pragma unroll
    for (int b = 0; b < C::NB; ++b) {
        if constexpr (C::LW * sizeof(FT) == 16) {
            const u32x4_t x = *(v4p*)(lane_row + b * 4 * C::LW);
            r[b * 4 + 0] = x.x;
            r[b * 4 + 1] = x.y;
            r[b * 4 + 2] = x.z;
            r[b * 4 + 3] = x.w;
        } else {
            const u32x2_t x = *(v2p*)(lane_row + b * 4 * C::LW);
            r[b * 2 + 0] = x.x;
            r[b * 2 + 1] = x.y;
        }
    }
}

// A lane's part of a row from a 32-bit LDS byte address (ds_read).
template <int RP, typename FT>
__device__ __forceinline__ void load_raw_lds(uint32_t addr, uint32_t* r) {
    using C = Cols<RP, FT>;
    typedef __attribute__((address_space(3))) const u32x4_t v4p;
    typedef __attribute__((address_space(3))) const u32x2_t v2p;
    constexpr uint32_t LB = C::LW * sizeof(FT);   // bytes of one lane load
#pragma unroll
    for (int b = 0; b < C::NB; ++b) {
        const uint32_t ab = addr + (uint32_t)b * 4u * LB;
        if constexpr (LB == 16) {
            const u32x4_t x = *(v4p*)(size_t)ab;
            r[b * 4 + 0] = x.x;
            r[b * 4 + 1] = x.y;
            r[b * 4 + 2] = x.z;
            r[b * 4 + 3] = x.w;
        } else {
            const u32x2_t x = *(v2p*)(size_t)ab;
            r[b * 2 + 0] = x.x;
            r[b * 2 + 1] = x.y;
        }
    }
}

// per-operand gather bases of one lane
template <int NO, typename FT>
struct GatherBase {
    const FT* g[NO];        // rows in memory (this lane's first column)
    const FT* l[NO];        // rows in LDS (generic addresses; the low word is the LDS address)
};

// A factor row by frequency rank (raw words of this lane): from the LDS copy
// when hot, else from memory.  With AH, an operand whose rows all sit in LDS
// (`all`, uniform over the launch) is read by ds_read at 32-bit addresses:
// its rows then skip the vector-memory path and take no 64-bit address
// arithmetic.  AH is on for BF16 rows up to 6 operands (FP32 rows and wider
// launches spill with the second path).
template <int RP, typename FT, bool AH, typename GB>
__device__ __forceinline__ void gather_raw(const GB& gb, int w, uint32_t hot, bool all, uint32_t i,
                                           uint32_t* r) {
    if (AH && all)
        load_raw_lds<RP, FT>((uint32_t)reinterpret_cast<uint64_t>(gb.l[w]) + i * (uint32_t)(RP * sizeof(FT)),
                             r);
    else load_raw<RP, FT>(((i < hot) ? gb.l[w] : gb.g[w]) + (size_t)i * RP, r);
}

// Whether heavy pieces take the hot-split path (3-mode tensors, both
// operand precisions).
template <int NO, typename FT>
constexpr bool kSplitPath() { return NO == 2; }

// Whether a launch may run all-hot class pieces (every gather from LDS by
// ds_read): BF16 spMTTKRP on 4 to 7 modes, ordinary column sizes, and at most
// 192 gathered columns per nonzero (wider launches spill registers with both
// piece variants in one kernel).
template <int RP, int NO, typename FT, int KIND, bool BIG>
constexpr bool kClassPath() {
    return sizeof(FT) == 2 && KIND == 0 && NO >= 3 && NO <= 6 && RP * NO <= 192 && !BIG;
}

// Gathers of operand w.  SPLIT 0: no hot split (LDS or memory by rank).
// SPLIT 1: a hot-split piece -- operand 1 comes from LDS when the piece is
// hot (`hot1`, uniform over the wave) and from memory otherwise; operand 0
// (reused across a fiber) from LDS or memory by rank; the others from memory.
// SPLIT 2: an all-hot class piece -- every row from LDS at 32-bit addresses.
template <int RP, typename FT, int SPLIT, bool AH, typename GB>
__device__ __forceinline__ void gather_op(const SegArgs& a, bool hot1, int w, const GB& gb,
                                          uint32_t i, uint32_t* r) {
    if constexpr (SPLIT == 2) {
        load_raw_lds<RP, FT>((uint32_t)reinterpret_cast<uint64_t>(gb.l[w]) + i * (uint32_t)(RP * sizeof(FT)),
                             r);
    } else if constexpr (SPLIT == 0) {
        gather_raw<RP, FT, AH>(gb, w, a.hot[w], (a.allhot >> w) & 1u, i, r);
    } else {
        if (w == 0) gather_raw<RP, FT, AH>(gb, 0, a.hot[0], a.allhot & 1u, i, r);
        else if (w == 1 && hot1) load_raw_as<RP, FT, 3>(gb.l[1] + (size_t)i * RP, r);
        else load_raw_as<RP, FT, 1>(gb.g[w] + (size_t)i * RP, r);
    }
}

// One nonzero's contribution to the quad's accumulator (y[w] = gathered rows).
// KIND 0 forms v*y0, times the middle operands, then a fused multiply-add
// with the last operand; KIND 1 forms the model value m = <prod_w y_w, a_row>,
// the loss derivative g(m, v), and accumulates g * prod_w y_w.
template <int CPL, int NO, int KIND>
__device__ __forceinline__ void accumulate(float* acc, float v, const float (*y)[CPL],
                                           const float* arow, int loss, double* lsum, int q,
                                           bool want_loss) {
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
        const float gr = dloss(loss, m, v);
#pragma unroll
        for (int c = 0; c < CPL; ++c) acc[c] = __fmaf_rn(gr, p[c], acc[c]);
        if (want_loss && q == 0) *lsum += (double)lossval(loss, m, v);
    }
}

// Whether a quad reuses operand 0's row (the secondary sort mode's) from
// registers while consecutive nonzeros share it.  FP32 rows always; BF16 rows
// only when long (Rp >= 64) and with 3+ operands: at 4-mode rank 64 the reuse
// is 3 % faster, at 5-mode rank 32 5 % slower and at 3-mode rank 32 2 %
// slower than gathering every row (the selects cost more than the half-length
// rows they save).
template <int RP, int NO, typename FT>
constexpr bool kReuse0() { return sizeof(FT) == 4 || (NO >= 3 && RP >= 64); }

// Gathers of HB broadcast nonzeros (operand 0 reused from raw0prev when the
// index repeats and kReuse0 holds), then their accumulation in order.  `live(j)` says whether
// nonzero j of the batch exists; `before_step(j)` runs ahead of its
// accumulation (row changes of light segments).
template <int RP, int NO, typename FT, int KIND, int J, int HB, int CPL, int SPLIT = 0,
          typename GB, typename Live, typename Before>
__device__ __forceinline__ void batch4(const SegArgs& a, const GB& gb,
                                       const float* vb, const uint32_t (*ib)[NO], int q,
                                       float* acc, const float* arow, double* lsum,
                                       uint32_t& prev0, uint32_t* raw0prev, Live live,
                                       Before before_step, bool hot1 = false) {
    constexpr int RD = Cols<RP, FT>::RD;
    constexpr bool AH = sizeof(FT) == 2 && NO <= 6;
#pragma unroll
    for (int j0 = 0; j0 < J; j0 += HB) {
        uint32_t raw[HB][NO][RD];
        bool reuse[HB];
#pragma unroll
        for (int h = 0; h < HB; ++h) {
            const int j = j0 + h;
            const uint32_t before = (j == 0) ? prev0 : ib[j - 1][0];
            reuse[h] = kReuse0<RP, NO, FT>() && ib[j][0] == before;
            if (!reuse[h])
                gather_op<RP, FT, SPLIT, AH>(a, hot1, 0, gb, ib[j][0], raw[h][0]);
#pragma unroll
            for (int w = 1; w < NO; ++w)
                gather_op<RP, FT, SPLIT, AH>(a, hot1, w, gb, ib[j][w], raw[h][w]);
        }
#pragma unroll
        for (int h = 0; h < HB; ++h) {
            const int j = j0 + h;
            if (live(j)) {
#pragma unroll
                for (int d = 0; d < RD; ++d) raw0prev[d] = reuse[h] ? raw0prev[d] : raw[h][0][d];
                float y[NO][CPL];
                widen_raw<RP, FT>(raw0prev, y[0]);
#pragma unroll
                for (int w = 1; w < NO; ++w) widen_raw<RP, FT>(raw[h][w], y[w]);
                before_step(j);
                accumulate<CPL, NO, KIND>(acc, vb[j], y, arow, a.loss, lsum, q,
                                          a.loss_part != nullptr);
                prev0 = ib[j][0];
            }
        }
    }
}

// Quad broadcast from lane t, t known at compile time once the caller's loop
// is unrolled (the switch folds to one DPP move).
__device__ __forceinline__ uint32_t qb_at(uint32_t v, int t) {
    switch (t) {
        case 0: return qb<0>(v);
        case 1: return qb<1>(v);
        case 2: return qb<2>(v);
        default: return qb<3>(v);
    }
}
__device__ __forceinline__ float qbf_at(float v, int t) {
    return __uint_as_float(qb_at(__float_as_uint(v), t));
}

// Nonzeros [S, S + J) of the quad's 16 in a hot-split piece whose operand-1
// rows all sit in LDS (nonzero s is component s % 4 of lane s / 4).  The
// operand-0 gathers of all J are issued first; then, per nonzero, its value
// and operand-1 index are broadcast and its operand-1 row read from LDS just
// before the accumulation.  Operand-1 rows thus hold no registers while the
// gathers are in flight, and J = 6 memory gathers fit where the general
// batch keeps 4 (plus 4 LDS rows).
template <int RP, typename FT, int KIND, int S, int J, int CPL, typename GB, typename Live>
__device__ __forceinline__ void batch_hot(const SegArgs& a, const GB& gb,
                                          const float* sv4, const uint32_t (*si4)[4], int q,
                                          float* acc, const float* arow, double* lsum,
                                          uint32_t& prev0, uint32_t* raw0prev, Live live) {
    constexpr int RD = Cols<RP, FT>::RD;
    uint32_t raw0[J][RD];
    uint32_t i0[J];
    bool reuse[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
        i0[j] = qb_at(si4[0][(S + j) % 4], (S + j) / 4);
        const uint32_t before = (j == 0) ? prev0 : i0[j - 1];
        reuse[j] = sizeof(FT) == 4 && i0[j] == before;
        if (!reuse[j]) gather_raw<RP, FT, false>(gb, 0, a.hot[0], false, i0[j], raw0[j]);
    }
#pragma unroll
    for (int j = 0; j < J; ++j) {
        if (live(j)) {
            const uint32_t i1 = qb_at(si4[1][(S + j) % 4], (S + j) / 4);
            const float v = qbf_at(sv4[(S + j) % 4], (S + j) / 4);
            uint32_t r1[RD];
            load_raw_as<RP, FT, 3>(gb.l[1] + (size_t)i1 * RP, r1);
#pragma unroll
            for (int d = 0; d < RD; ++d) raw0prev[d] = reuse[j] ? raw0prev[d] : raw0[j][d];
            float y[2][CPL];
            widen_raw<RP, FT>(raw0prev, y[0]);
            widen_raw<RP, FT>(r1, y[1]);
            accumulate<CPL, 2, KIND>(acc, v, y, arow, a.loss, lsum, q, a.loss_part != nullptr);
            prev0 = i0[j];
        }
    }
}

// Contiguous split of P pieces over W waves: wave w owns
// [range_begin(w), range_begin(w + 1)) with P = W q + rem, the first rem
// waves taking q + 1 pieces.
__device__ __forceinline__ uint32_t range_begin(uint32_t w, uint32_t q, uint32_t rem) {
    return w * q + min(w, rem);
}
// the wave whose range holds piece p
__device__ __forceinline__ uint32_t wave_of(uint32_t p, uint32_t q, uint32_t rem) {
    const uint32_t big = rem * (q + 1);
    return (p < big) ? p / (q + 1) : rem + (p - big) / q;
}

// KIND 0: spMTTKRP / adjoint.  KIND 1: fused completion loss + gradient.
// BIG: value / index columns of 4 GiB or more (2^30 nonzeros or more), read
// through buffer resources based at each heavy piece instead of one resource
// per column (a buffer offset is 32 bits).  Its own instantiations, so the
// code of ordinary tensors is not touched by it.
template <int RP, int NO, typename FT, int KIND, bool BIG>
__global__ __launch_bounds__((seg_threads<RP, NO, KIND>())) void k_seg(SegArgs a) {
    using C = Cols<RP, FT>;
    constexpr int CPL = C::CPL;
    constexpr int HB = seg_hoist<RP, NO, FT, KIND>();
    constexpr int TB = seg_threads<RP, NO, KIND>();
    constexpr int PL = (int)kPieceLane;
    constexpr int NR = seg_rounds<RP, NO, FT, KIND>();
    extern __shared__ __align__(16) unsigned char lds_raw[];
    FT* lds = reinterpret_cast<FT*>(lds_raw);

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int qw = lane >> 2;         // quad within the wave
    const int q = lane & 3;

    // stage the hot rows of every operand (16-byte copies)
    uint32_t off[NO];
    {
        uint32_t o = 0;
#pragma unroll
        for (int w = 0; w < NO; ++w) {
            off[w] = o;
            const uint32_t nvec = a.hot[w] * (uint32_t)(RP * sizeof(FT) / 16);
            const uint4* src = reinterpret_cast<const uint4*>(a.Yo[w]);
            uint4* dst = reinterpret_cast<uint4*>(lds + (size_t)o * RP);
            for (uint32_t e = tid; e < nvec; e += TB) dst[e] = src[e];
            o += a.hot[w];
        }
    }
    __syncthreads();
    GatherBase<NO, FT> gb;
#pragma unroll
    for (int w = 0; w < NO; ++w) {
        gb.g[w] = reinterpret_cast<const FT*>(a.Yo[w]) + C::col(0, q, 0);
        gb.l[w] = lds + (size_t)off[w] * RP + C::col(0, q, 0);
    }

    // Static schedule (pieces have equal length but the last of each row, so
    // round-robin balances them; a shared work counter would serialise on one
    // L2 atomic per claim).  Wave numbers are strided over the workgroups, so
    // the ranges one CU runs come from across the whole piece list and every
    // CU gets about the same mix of rows.
    const uint32_t nwaves = gridDim.x * (uint32_t)(TB / 64);
    const uint32_t gw = (uint32_t)__builtin_amdgcn_readfirstlane(
        (int)((uint32_t)(tid >> 6) * gridDim.x + blockIdx.x));

    // rows without nonzeros belong to no piece or segment: one wave zeroes
    // each (RP / 4 lanes, 16 bytes each)
    for (uint32_t e = gw; e < a.nempty; e += nwaves)
        if (lane < RP / 4)
            reinterpret_cast<float4*>(a.out + (size_t)a.empty_rows[e] * RP)[lane] =
                make_float4(0.f, 0.f, 0.f, 0.f);

    // ---- phase 1: heavy-row pieces, one wave at a time ----
    // Wave gw owns pieces [range_begin(gw), range_begin(gw + 1)); its running row
    // sum is flushed into the slot of the last piece of each row it touches,
    // which is where k_heavy looks for it.
    {
        // buffer resources over the copy's value / index columns (dword ranges;
        // unused by BIG launches)
        __amdgpu_buffer_rsrc_t rs_idx[NO];
#pragma unroll
        for (int w = 0; w < NO; ++w)
            rs_idx[w] = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t*>(a.oidx[w]), (short)0,
                                                          (int)a.stream_bytes, kBufDword3);
        const __amdgpu_buffer_rsrc_t rs_val = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<float*>(a.vals), (short)0, (int)a.stream_bytes, kBufDword3);
        const uint32_t pq = a.npieces / nwaves, prem = a.npieces % nwaves;
        const uint32_t p0 = range_begin(gw, pq, prem);
        const uint32_t p1 = range_begin(gw + 1, pq, prem);
        float acc[CPL];
#pragma unroll
        for (int c = 0; c < CPL; ++c) acc[c] = 0.f;
        float arow[KIND == 1 ? CPL : 1];
        double lsum = 0.0;
        uint32_t prev0 = kNone;
        uint32_t raw0prev[C::RD];
#pragma unroll
        for (int d = 0; d < C::RD; ++d) raw0prev[d] = 0u;
        bool whole = false;   // the current row starts inside this range
        for (uint32_t p = p0; p < p1; ++p) {
            const uint32_t kb = a.pc_k[p];
            const uint32_t n = a.pc_e[p] - kb;
            const uint32_t row = a.pc_row[p];
            if (p == p0 || a.pc_row[p - 1] != row) {   // the row's first piece in this range
                whole = (p > p0) || (p == 0) || (a.pc_row[p - 1] != row);
                if constexpr (KIND == 1) load_row<RP, float>(a.An + (size_t)row * RP, q, arow);
            }
            // the piece's stream: quad g owns nonzeros [16 g, 16 g + 16), lane q
            // of it holds the 4 consecutive ones from 16 g + 4 q (one 16-byte
            // buffer load per column; dword-aligned offsets)
            static_assert(PL == 4, "a lane holds 4 consecutive nonzeros of its quad's 16");
            const uint32_t e0 = (uint32_t)(16 * qw + 4 * q);
            float sv4[4];
            uint32_t si4[NO][4];
            {
                __amdgpu_buffer_rsrc_t ri[NO], rv;
                uint32_t boff;
                if constexpr (BIG) {
                    // resources based at the piece's first nonzero (wave-uniform:
                    // scalar arithmetic), byte offsets below 1 KiB
                    const uint32_t rem = a.nnz - kb;
                    const int nrec = (int)(rem >= kPiece ? kPiece * 4u : rem * 4u);
#pragma unroll
                    for (int w = 0; w < NO; ++w)
                        ri[w] = __builtin_amdgcn_make_buffer_rsrc(
                            const_cast<uint32_t*>(a.oidx[w] + kb), (short)0, nrec, kBufDword3);
                    rv = __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(a.vals + kb), (short)0,
                                                           nrec, kBufDword3);
                    boff = e0 * 4u;
                } else {
#pragma unroll
                    for (int w = 0; w < NO; ++w) ri[w] = rs_idx[w];
                    rv = rs_val;
                    boff = (kb + e0) * 4u;
                }
#pragma unroll
                for (int w = 0; w < NO; ++w) {
                    const auto x = __builtin_amdgcn_raw_buffer_load_b128(ri[w], boff, 0, kStreamPolicy);
#pragma unroll
                    for (int c = 0; c < 4; ++c) si4[w][c] = (e0 + c < n) ? x[c] : 0u;
                }
                if (a.vperm) {
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        sv4[c] = (e0 + c < n) ? a.vals[a.vperm[kb + e0 + c]] : 0.f;
                } else {
                    const auto x = __builtin_amdgcn_raw_buffer_load_b128(rv, boff, 0, kStreamPolicy);
#pragma unroll
                    for (int c = 0; c < 4; ++c) sv4[c] = (e0 + c < n) ? __uint_as_float(x[c]) : 0.f;
                }
            }
            // round t: nonzeros 16 g + 4 t + j, held by lane t of the quad
            // a hot-split piece reads operand 1 from LDS (hot) or memory (cold)
            const bool hot1 = kSplitPath<NO, FT>() && a.pc_hot[p] != 0u;
            auto do_rounds = [&](auto t0c, auto splitc) {
                constexpr int T0 = decltype(t0c)::value;
                constexpr int SP = decltype(splitc)::value;
                const uint32_t k0 = (uint32_t)(16 * qw + 4 * T0);
                if (k0 < n) {
                    float vb[4 * NR];
                    uint32_t ib[4 * NR][NO];
                    bcast_round<T0, NO>(sv4, si4, vb, ib);
                    if constexpr (NR > 1) bcast_round<T0 + 1, NO>(sv4, si4, vb + 4, ib + 4);
                    batch4<RP, NO, FT, KIND, 4 * NR, (NR > 1 ? 4 * NR : HB), CPL, SP>(
                        a, gb, vb, ib, q, acc, arow, &lsum, prev0, raw0prev,
                        [&](int j) { return k0 + (uint32_t)j < n; }, [&](int) {}, hot1);
                }
            };
            // FP32 hot pieces at Rp <= 32: nonzeros in parts of 6, 6 and 4
            // (batch_hot; wider rows would spill its registers)
            auto hot_part = [&](auto sc, auto jc) {
                constexpr int S = decltype(sc)::value;
                constexpr int JN = decltype(jc)::value;
                const uint32_t k0 = (uint32_t)(16 * qw + S);
                if (k0 < n) {
                    if constexpr (NO == 2)
                        batch_hot<RP, FT, KIND, S, JN, CPL>(
                            a, gb, sv4, si4, q, acc, arow, &lsum, prev0, raw0prev,
                            [&](int j) { return k0 + (uint32_t)j < n; });
                }
            };
            auto do_piece = [&](auto splitc) {
                if constexpr (decltype(splitc)::value == 1 && sizeof(FT) == 4 && KIND == 0 &&
                              RP <= 32) {
                    if (hot1) {
                        hot_part(std::integral_constant<int, 0>{}, std::integral_constant<int, 6>{});
                        hot_part(std::integral_constant<int, 6>{}, std::integral_constant<int, 6>{});
                        hot_part(std::integral_constant<int, 12>{}, std::integral_constant<int, 4>{});
                        return;
                    }
                }
                do_rounds(std::integral_constant<int, 0>{}, splitc);
                if constexpr (NR == 1) do_rounds(std::integral_constant<int, 1>{}, splitc);
                do_rounds(std::integral_constant<int, 2>{}, splitc);
                if constexpr (NR == 1) do_rounds(std::integral_constant<int, 3>{}, splitc);
            };
            // launches on three modes run on a hot-split copy; an all-hot
            // class piece (wave-uniform flag) reads every row from LDS
            if constexpr (kClassPath<RP, NO, FT, KIND, BIG>()) {
                if (a.cls_on && a.pc_hot[p] != 0u) do_piece(std::integral_constant<int, 2>{});
                else do_piece(std::integral_constant<int, 0>{});
            } else {
                do_piece(std::integral_constant<int, kSplitPath<NO, FT>() ? 1 : 0>{});
            }
            const bool row_ends = (p + 1 == a.npieces) || (a.pc_row[p + 1] != row);
            const bool flush = row_ends || (p + 1 == p1);
            if (flush) {
                // sum the 16 quad partials (butterfly over quads, fixed order);
                // a row held whole by this wave is final, else a partial
#pragma unroll
                for (int o = 4; o < 64; o <<= 1)
#pragma unroll
                    for (int c = 0; c < CPL; ++c) acc[c] += __shfl_xor(acc[c], o, 64);
                float* dst = (whole && row_ends) ? a.out + (size_t)row * RP : a.part + (size_t)p * RP;
                if (qw == 0) store_row<RP, FT>(dst, q, acc);
#pragma unroll
                for (int c = 0; c < CPL; ++c) acc[c] = 0.f;
            }
            if constexpr (KIND == 1) {
                if (a.loss_part) {
                    double l = 0.0;
                    if (flush) {
                        l = lsum;
#pragma unroll
                        for (int o = 4; o < 64; o <<= 1) l += __shfl_xor(l, o, 64);
                        lsum = 0.0;
                    }
                    if (lane == 0) a.loss_part[p] = l;
                }
            }
        }
    }

    // ---- phase 2: light segments, one quad each (16 per wave step) ----
    for (uint32_t base = gw * 16u; base < a.nseg; base += nwaves * 16u) {
        const uint32_t i = base + (uint32_t)qw;
        if (i >= a.nseg) continue;

        const uint32_t kb = a.seg_k[i];
        const uint32_t ke = a.seg_e[i];
        uint32_t row = a.seg_row[i];
        uint32_t rend = a.rowptr[row + 1];
        float acc[CPL];
#pragma unroll
        for (int c = 0; c < CPL; ++c) acc[c] = 0.f;
        float arow[KIND == 1 ? CPL : 1];
        if constexpr (KIND == 1) load_row<RP, float>(a.An + (size_t)row * RP, q, arow);
        double lsum = 0.0;
        uint32_t prev0 = kNone;
        uint32_t raw0prev[C::RD];
#pragma unroll
        for (int d = 0; d < C::RD; ++d) raw0prev[d] = 0u;

        auto fetch = [&](uint32_t kx, float& v_l, uint32_t* i_l) {
            const uint32_t kk = kx + q;
            if (kk < ke) {
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
        fetch(kb, v_n, i_n);
        for (uint32_t k0 = kb; k0 < ke; k0 += 4) {
            const float v_l = v_n;
            uint32_t i_l[NO];
#pragma unroll
            for (int w = 0; w < NO; ++w) i_l[w] = i_n[w];
            if (k0 + 4 < ke) fetch(k0 + 4, v_n, i_n);
            float vb[4];
            uint32_t ib[4][NO];
            bcast_batch<0, NO>(v_l, i_l, vb[0], ib[0]);
            bcast_batch<1, NO>(v_l, i_l, vb[1], ib[1]);
            bcast_batch<2, NO>(v_l, i_l, vb[2], ib[2]);
            bcast_batch<3, NO>(v_l, i_l, vb[3], ib[3]);
            batch4<RP, NO, FT, KIND, 4, HB, CPL>(
                a, gb, vb, ib, q, acc, arow, &lsum, prev0, raw0prev,
                [&](int j) { return k0 + (uint32_t)j < ke; },
                [&](int j) {
                    const uint32_t k = k0 + (uint32_t)j;
                    if (k >= rend) {   // the row is complete
                        store_row<RP, FT>(a.out + (size_t)row * RP, q, acc);
#pragma unroll
                        for (int c = 0; c < CPL; ++c) acc[c] = 0.f;
                        do {
                            ++row;
                            rend = a.rowptr[row + 1];
                        } while (k >= rend);
                        if constexpr (KIND == 1)
                            load_row<RP, float>(a.An + (size_t)row * RP, q, arow);
                    }
                });
        }
        store_row<RP, FT>(a.out + (size_t)row * RP, q, acc);
        if (KIND == 1 && a.loss_part && q == 0) a.loss_part[a.npieces + i] = lsum;
    }
}

// ---------------------------------------------------------------------------
//  launch dispatch
// ---------------------------------------------------------------------------
// Rows of operand 1 held in LDS under a hot split: as many FP32 rows as the
// budget holds, for either precision, so FP32 and BF16 launches share one
// split copy.  BF16 rows take half the budget; the rest holds operand 0's
// leading rows.
inline uint32_t hot_split_rows(dyn_gpu_ctx* c, int Rp, uint32_t rows) {
    const size_t budget = std::min<size_t>(c->max_lds, 160u * 1024u) - 1024u;
    return (uint32_t)std::min<size_t>(rows, budget / ((size_t)Rp * 4));
}

// LDS rows of each operand on the plain order.  Operands whose rows all fit
// take them all, smallest first, up to half the budget (their gathers then
// never leave the LDS); the rest of the budget is shared out smallest first,
// each taking at most an equal share of what is left.
inline void plan_lds_rows(dyn_gpu_ctx* c, size_t rowb, int NO, const uint32_t* dims_o,
                          uint32_t* hot) {
    const size_t budget = std::min<size_t>(c->max_lds, 160u * 1024u) - 1024u;
    int ord[kMaxModes - 1];
    for (int w = 0; w < NO; ++w) ord[w] = w;
    std::sort(ord, ord + NO, [&](int x, int y) { return dims_o[x] < dims_o[y]; });
    size_t whole = 0;
    int k = 0;
    for (; k < NO && whole + (size_t)dims_o[ord[k]] * rowb <= budget / 2; ++k) {
        hot[ord[k]] = dims_o[ord[k]];
        whole += (size_t)dims_o[ord[k]] * rowb;
    }
    size_t left = budget - whole;
    for (int m = k; m < NO; ++m) {
        const size_t share = left / ((size_t)(NO - m) * rowb);
        hot[ord[m]] = (uint32_t)std::min<size_t>(share, dims_o[ord[m]]);
        left -= (size_t)hot[ord[m]] * rowb;
    }
}

template <int RP, int NO, typename FT, int KIND, bool BIG>
int launch_seg(dyn_gpu_ctx* c, SegArgs& a, const uint32_t* dims_o) {
    const size_t budget = std::min<size_t>(c->max_lds, 160u * 1024u) - 1024u;
    const size_t rowb = (size_t)RP * sizeof(FT);
    size_t bytes = 0;
    if (kSplitPath<NO, FT>() && a.pc_hot) {
        // hot split: operand 1's leading rows, exactly the split, then as many
        // of operand 0's as the rest of the budget holds
        for (int w = 0; w < NO; ++w) a.hot[w] = 0;
        a.hot[1] = hot_split_rows(c, RP, dims_o[1]);
        a.hot[0] = std::min<uint32_t>(
            dims_o[0], (uint32_t)((budget - std::min(budget, (size_t)a.hot[1] * rowb)) / rowb));
        bytes = ((size_t)a.hot[0] + a.hot[1]) * rowb;
    } else {
        plan_lds_rows(c, rowb, NO, dims_o, a.hot);
        for (int w = 0; w < NO; ++w) bytes += (size_t)a.hot[w] * rowb;
    }
    // all-hot class pieces need the copy split at exactly these LDS rows
    a.cls_on = 0;
    if (kClassPath<RP, NO, FT, KIND, BIG>() && a.pc_hot) {
        a.cls_on = 1;
        for (int w = 0; w < NO; ++w)
            if (a.cls_h[w] == 0 || a.cls_h[w] != a.hot[w]) a.cls_on = 0;
    }
    a.allhot = 0;
    for (int w = 0; w < NO; ++w)
        if (a.hot[w] >= dims_o[w]) a.allhot |= 1u << w;
    // a 3-mode launch reads the copy's hot-piece flags (the hot split)
    if (kSplitPath<NO, FT>() && !a.pc_hot)
        return set_err(c, DYN_ERR_STATE, "3-mode spMTTKRP launch without a hot split");
    auto fn = k_seg<RP, NO, FT, KIND, BIG>;
    static std::atomic<uint64_t> lds_set{0};
    lds_limit_once(c, &lds_set, (const void*)fn, (int)std::min<size_t>(c->max_lds, 160u * 1024u));
    a.nwaves = (uint32_t)c->num_cus * (uint32_t)(seg_threads<RP, NO, KIND>() / 64);
    if (c->launch_ev[0]) {
        hipExtLaunchKernelGGL(fn, dim3(c->num_cus), dim3(seg_threads<RP, NO, KIND>()),
                              (uint32_t)bytes, c->stream, c->launch_ev[0], c->launch_ev[1], 0u, a);
    } else {
        hipLaunchKernelGGL(fn, dim3(c->num_cus), dim3(seg_threads<RP, NO, KIND>()), bytes,
                           c->stream, a);
    }
    return DYN_OK;
}

template <int RP, typename FT, int KIND, bool BIG>
int dispatch_no(dyn_gpu_ctx* c, int NO, SegArgs& a, const uint32_t* d) {
    switch (NO) {
        case 1: return launch_seg<RP, 1, FT, KIND, BIG>(c, a, d);
        case 2: return launch_seg<RP, 2, FT, KIND, BIG>(c, a, d);
        case 3: return launch_seg<RP, 3, FT, KIND, BIG>(c, a, d);
        case 4: return launch_seg<RP, 4, FT, KIND, BIG>(c, a, d);
        case 5: return launch_seg<RP, 5, FT, KIND, BIG>(c, a, d);
        case 6: return launch_seg<RP, 6, FT, KIND, BIG>(c, a, d);
        case 7: return launch_seg<RP, 7, FT, KIND, BIG>(c, a, d);
        default: return set_err(c, DYN_ERR_UNSUPPORTED, "unsupported mode count %d", NO + 1);
    }
}

// Launches the segment kernel for padded rank Rp and NO gathered operands.
// The BIG = true instantiations are compiled in their own file.
template <typename FT, int KIND, bool BIG>
int dispatch_rp(dyn_gpu_ctx* c, int Rp, int NO, SegArgs& a, const uint32_t* d) {
    switch (Rp) {
        case 16: return dispatch_no<16, FT, KIND, BIG>(c, NO, a, d);
        case 32: return dispatch_no<32, FT, KIND, BIG>(c, NO, a, d);
        case 64: return dispatch_no<64, FT, KIND, BIG>(c, NO, a, d);
        case 128: return dispatch_no<128, FT, KIND, BIG>(c, NO, a, d);
        default: return set_err(c, DYN_ERR_UNSUPPORTED, "unsupported padded rank %d", Rp);
    }
}

}  // namespace dyn
