// CP-ALS algebra on the device and the CP-ALS driver.
//
//   Gram        G_n = Y_n^T Y_n from FP32 factors on MFMA (v_mfma_f32_32x32x2f32
//               / 16x16x4f32), FP32 per-workgroup partials summed in FP64 in a
//               fixed order (la::gram, src/dynasor_linalg.cpp:47-99).
//   Cholesky    V_n = ones .* G_m (m != n) in FP64, ridge * max(diag) added,
//               right-looking factorisation with the reference's defensive
//               zero pivots (src/dynasor_linalg.cpp:104-200,
//               src/dynasor_cpals.cpp:292-305).
//   Solve       X = Yhat V^{-1} by per-row forward/back substitution in FP64
//               (src/dynasor_linalg.cpp:210-246); column sums of x^2 in FP64.
//   Normalise   lambda_r = ||X[:, r]||_2, Y = X / lambda (zero when lambda
//               <= 1e-20) (src/dynasor_linalg.cpp:251-316); the same pass
//               refreshes the BF16 shadow and the Gram partials, and for the
//               last mode the <Y, Yhat> column partials of the fit.
//   Fit         ||T||^2 - 2 <T, M> + ||M||^2 in FP64 with the Yhat identity
//               (src/dynasor_cpals.cpp:79-113, 332-373): from the last mode's
//               Yhat under Gauss-Seidel; under Jacobi from the next sweep's
//               Yhat of the smallest mode, which reads exactly the model being
//               measured (the reference's Jacobi print uses a stale Yhat).
#include "dyn_internal.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace dyn {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int kMaxGramWG = 1024;
constexpr int kSolveBlock = 256;
constexpr int kMaxSolveWG = 1024;
// relative rounding allowance of ||T||^2 - 2 <T, M> + ||M||^2 (FP32 spMTTKRP
// inputs), measured against the largest of the three terms
constexpr double kFitRoundRel = 1e-4;

// ---------------------------------------------------------------------------
//  FP64 reduction of ||vals||^2
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_sumsq(const float* __restrict__ v, uint64_t n,
                                                  double* __restrict__ part) {
    __shared__ double red[kBlock / 64];
    double s = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * kBlock) {
        const double x = v[i];
        s += x * x;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) t += red[w];
        part[blockIdx.x] = t;
    }
}

__global__ void k_sum_d(const double* __restrict__ part, uint32_t n, double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double t = 0.0;
    for (uint32_t i = 0; i < n; ++i) t += part[i];
    *out = t;
}

int tensor_norm2(dyn_gpu_ctx* c, double* out) {
    if (c->tnorm2 >= 0.0) {
        *out = c->tnorm2;
        return DYN_OK;
    }
    const unsigned nb = 1024;
    void* sp = nullptr;
    DYN_CHECK(scratch2(c, (nb + 8) * sizeof(double), &sp));
    double* part = (double*)sp;
    hipLaunchKernelGGL(k_sumsq, dim3(nb), dim3(kBlock), 0, c->stream, c->vals, c->nnz, part);
    hipLaunchKernelGGL(k_sum_d, dim3(1), dim3(64), 0, c->stream, part, nb, part + nb);
    DYN_CHECK(check_launch(c, "sumsq"));
    double h = 0.0;
    DYN_HIP(c, hipMemcpyAsync(&h, part + nb, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    DYN_HIP(c, hipStreamSynchronize(c->stream));
    c->tnorm2 = h;
    *out = h;
    return DYN_OK;
}

// sum over k of p[(first + k step) ld + col] for first + k step < n, added in
// k order; the loads are issued eight at a time so that their latencies overlap
template <typename T>
__device__ __forceinline__ double strided_sum(const T* __restrict__ p, uint32_t first,
                                              uint32_t step, uint32_t n, size_t ld, int col) {
    double t = 0.0;
    uint32_t b = first;
    for (; b + 7u * step < n; b += 8u * step) {
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = (double)p[(size_t)(b + k * step) * ld + col];
#pragma unroll
        for (int k = 0; k < 8; ++k) t += v[k];
    }
    for (; b < n; b += step) t += (double)p[(size_t)b * ld + col];
    return t;
}

// ---------------------------------------------------------------------------
//  Scale + Gram pass.  Rows are consumed in MFMA operand order: with the
//  32x32x2 instruction lane l holds row (l / 32) of a row pair at column
//  (l % 32) of a 32-column block (A = Y^T and B = Y take the same value);
//  with 16x16x4, row (l / 16) of a row quad at column (l % 16).
// ---------------------------------------------------------------------------
struct ScaleArgs {
    float*        Y;          // rows x RP, scaled in place when sq_part != null
    float*        Yg;         // gathered FP32 operand, rows in rank order (nullable)
    bf16*         Ygb;        // gathered BF16 operand (with Yg)
    const uint32_t* relab;    // row -> rank (with Yg)
    const double* sq_part;    // [sq_nwg][RP] column sums of x^2 from the solve
                              // (nullable: Gram only, no scaling)
    uint32_t      sq_nwg;
    int           R;
    double*       lambda_out; // written by workgroup 0 when scaling
    double*       inv_out;
    const float*  Yhat;       // non-null: accumulate <Y, Yhat> per column
    float*        gpart;      // [nwg][RP][RP]
    double*       ipart;      // [nwg][RP]
    uint32_t      rows;
    uint32_t      nwg;
};

__device__ __forceinline__ bf16 to_bf16(float x) { return __float2bfloat16(x); }

template <int RP>
__global__ __launch_bounds__(kBlock) void k_scale_gram(ScaleArgs a) {
    __shared__ float tile[RP * RP];
    __shared__ double ired[kBlock / 64][RP];
    __shared__ double s_inv[RP];
    __shared__ double s_red[kBlock];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    constexpr int W = kBlock / 64;

    // lambda_r = ||X[:, r]||_2 from the solve's per-workgroup partials: every
    // workgroup repeats this small reduction (fixed order: strided subsets,
    // then the subsets in order) instead of paying a separate launch
    const bool scale = a.sq_part != nullptr;
    if (scale) {
        constexpr int NS = kBlock / RP;
        const int col = tid % RP, sub = tid / RP;
        s_red[tid] = strided_sum(a.sq_part, (uint32_t)sub, (uint32_t)NS, a.sq_nwg, RP, col);
        __syncthreads();
        if (tid < RP) {
            double v = 0.0;
            for (int j = 0; j < NS; ++j) v += s_red[j * RP + tid];
            const double l = (tid < a.R) ? sqrt(v) : 0.0;
            const double iv = (l > 1e-20) ? 1.0 / l : 0.0;
            s_inv[tid] = iv;
            if (blockIdx.x == 0) {
                a.lambda_out[tid] = l;
                a.inv_out[tid] = iv;
            }
        }
        __syncthreads();
    }

    if constexpr (RP == 16) {
        // 16x16x4: one tile; each wave step consumes 4 rows
        v4f acc = {0.f, 0.f, 0.f, 0.f};
        const int col = lane & 15;
        const int rsub = lane >> 4;
        const double inv = scale ? s_inv[col] : 1.0;
        double ip = 0.0;
        const uint32_t stride = a.nwg * W * 4u;
        // U row quads per trip: all loads first, then the updates in row order
        constexpr int U = 4;
        for (uint32_t r0 = (blockIdx.x * W + wave) * 4u; r0 < a.rows; r0 += stride * U) {
            float x[U], yh[U];
            uint32_t rl[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t r = r0 + u * stride + rsub;
                const uint32_t rc = r < a.rows ? r : a.rows - 1;
                x[u] = a.Y[(size_t)rc * RP + col];
                yh[u] = a.Yhat ? a.Yhat[(size_t)rc * RP + col] : 0.f;
                rl[u] = a.Yg ? a.relab[rc] : 0u;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t r = r0 + u * stride + rsub;
                const bool ok = r < a.rows;
                float y = x[u];
                if (scale) y = (float)((double)y * inv);
                if (ok) {
                    if (scale) a.Y[(size_t)r * RP + col] = y;
                    if (a.Yg) {
                        const size_t g = (size_t)rl[u] * RP + col;
                        a.Yg[g] = y;
                        a.Ygb[g] = to_bf16(y);
                    }
                    if (a.Yhat) ip += (double)y * (double)yh[u];
                } else {
                    y = 0.f;
                }
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(y, y, acc, 0, 0, 0);
            }
        }
        // deterministic cross-wave sum through LDS (wave order)
        for (int w = 0; w < W; ++w) {
            if (wave == w) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = 4 * (lane >> 4) + v;
                    const int j = lane & 15;
                    if (w == 0) tile[i * RP + j] = acc[v];
                    else tile[i * RP + j] += acc[v];
                }
            }
            __syncthreads();
        }
        if (a.Yhat) {
            // lanes with the same column: l, l+16, l+32, l+48
            ip += __shfl_xor(ip, 16, 64);
            ip += __shfl_xor(ip, 32, 64);
            if (lane < 16) ired[wave][lane] = ip;
        }
    } else {
        constexpr int TB = RP / 32;                       // 32-column blocks
        constexpr bool kSplitRows = (TB <= 2);            // else: wave w owns block-row w
        constexpr int NT = kSplitRows ? TB * TB : TB;     // tiles per wave
        v16f acc[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[t][v] = 0.f;
        const int col = lane & 31;
        const int rsub = lane >> 5;
        double inv[TB];
#pragma unroll
        for (int b = 0; b < TB; ++b) inv[b] = scale ? s_inv[b * 32 + col] : 1.0;
        double ip[TB];
#pragma unroll
        for (int b = 0; b < TB; ++b) ip[b] = 0.0;
        const uint32_t first = kSplitRows ? (blockIdx.x * W + wave) * 2u : blockIdx.x * 2u;
        const uint32_t stride = kSplitRows ? a.nwg * W * 2u : a.nwg * 2u;
        // U row pairs per trip: all loads first, then the updates in row order
        constexpr int U = 4;
        for (uint32_t r00 = first; r00 < a.rows; r00 += stride * U) {
          float xl[U][TB], yh[U][TB];
          uint32_t rl[U];
#pragma unroll
          for (int u = 0; u < U; ++u) {
              const uint32_t r = r00 + u * stride + rsub;
              const uint32_t rc = r < a.rows ? r : a.rows - 1;
#pragma unroll
              for (int b = 0; b < TB; ++b) {
                  const size_t off = (size_t)rc * RP + b * 32 + col;
                  xl[u][b] = a.Y[off];
                  yh[u][b] = a.Yhat ? a.Yhat[off] : 0.f;
              }
              rl[u] = a.Yg ? a.relab[rc] : 0u;
          }
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const uint32_t r = r00 + u * stride + rsub;
            float y[TB];
#pragma unroll
            for (int b = 0; b < TB; ++b) {
                y[b] = 0.f;
                if (r < a.rows) {
                    const size_t off = (size_t)r * RP + b * 32 + col;
                    float x = xl[u][b];
                    if (scale) x = (float)((double)x * inv[b]);
                    y[b] = x;
                    // in the block-row split every wave reads the whole row;
                    // wave b alone writes column block b
                    const bool writer = kSplitRows || wave == b;
                    if (writer) {
                        if (scale) a.Y[off] = x;
                        if (a.Yg) {
                            const size_t g = (size_t)rl[u] * RP + b * 32 + col;
                            a.Yg[g] = x;
                            a.Ygb[g] = to_bf16(x);
                        }
                        if (a.Yhat) ip[b] += (double)x * (double)yh[u][b];
                    }
                }
            }
            if constexpr (kSplitRows) {
#pragma unroll
                for (int bm = 0; bm < TB; ++bm)
#pragma unroll
                    for (int bn = 0; bn < TB; ++bn)
                        acc[bm * TB + bn] = __builtin_amdgcn_mfma_f32_32x32x2f32(
                            y[bm], y[bn], acc[bm * TB + bn], 0, 0, 0);
            } else {
                float ym = y[0];
#pragma unroll
                for (int b = 0; b < TB; ++b)
                    if (b == wave) ym = y[b];
#pragma unroll
                for (int bn = 0; bn < TB; ++bn)
                    acc[bn] = __builtin_amdgcn_mfma_f32_32x32x2f32(ym, y[bn], acc[bn], 0, 0, 0);
            }
          }
        }
        // accumulator layout (32x32): register v of lane l holds
        // C[8*(v/4) + 4*(l/32) + v%4][l%32]
        if constexpr (kSplitRows) {
            for (int w = 0; w < W; ++w) {
                if (wave == w) {
#pragma unroll
                    for (int t = 0; t < NT; ++t) {
                        const int bm = t / TB, bn = t % TB;
#pragma unroll
                        for (int v = 0; v < 16; ++v) {
                            const int i = bm * 32 + 8 * (v / 4) + 4 * (lane >> 5) + (v % 4);
                            const int j = bn * 32 + (lane & 31);
                            if (w == 0) tile[i * RP + j] = acc[t][v];
                            else tile[i * RP + j] += acc[t][v];
                        }
                    }
                }
                __syncthreads();
            }
        } else {
#pragma unroll
            for (int bn = 0; bn < NT; ++bn)
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int i = wave * 32 + 8 * (v / 4) + 4 * (lane >> 5) + (v % 4);
                    const int j = bn * 32 + (lane & 31);
                    tile[i * RP + j] = acc[bn][v];
                }
            __syncthreads();
        }
        if (a.Yhat) {
#pragma unroll
            for (int b = 0; b < TB; ++b) {
                double t = ip[b];
                t += __shfl_xor(t, 32, 64);
                if (lane < 32) {
                    if (kSplitRows) ired[wave][b * 32 + lane] = t;
                    else if (wave == b) ired[0][b * 32 + lane] = t;
                }
            }
        }
    }
    __syncthreads();
    float* gp = a.gpart + (size_t)blockIdx.x * RP * RP;
    for (int i = tid; i < RP * RP; i += kBlock) gp[i] = tile[i];
    if (a.Yhat && tid < RP) {
        double t = 0.0;
        const bool split = (RP == 16) || (RP / 32 <= 2);
        if (split) {
            for (int w = 0; w < W; ++w) t += ired[w][tid];
        } else {
            t = ired[0][tid];
        }
        a.ipart[(size_t)blockIdx.x * RP + tid] = t;
    }
}

// G = sum of the per-workgroup FP32 Gram partials in FP64.  Block of 256:
// 64 entries x 4 interleaved partial subsets, combined in subset order.
template <int RP>
__global__ __launch_bounds__(kBlock) void k_gram_reduce(const float* __restrict__ gpart,
                                                        uint32_t nwg, double* __restrict__ G) {
    // 16 elements per block, 16 strided subsets of the partials per element,
    // subsets combined in order
    __shared__ double red[16][16];
    const int le = threadIdx.x & 15, sub = threadIdx.x >> 4;
    const int e = blockIdx.x * 16 + le;
    red[sub][le] = (e < RP * RP) ? strided_sum(gpart, (uint32_t)sub, 16u, nwg, (size_t)RP * RP, e)
                                 : 0.0;
    __syncthreads();
    if (sub == 0 && e < RP * RP) {
        double t = 0.0;
        for (int j = 0; j < 16; ++j) t += red[j][le];
        G[e] = t;
    }
}

// ---------------------------------------------------------------------------
//  V = ones .* G_m (m != n), ridge, Cholesky, then V^{-1} (one workgroup,
//  FP64 in LDS).
//
//  The reference solves every row with a forward and a backward substitution
//  whose zero pivots yield zero components (src/dynasor_linalg.cpp:210-246).
//  Both substitutions are linear maps: with S the set of non-zero pivots the
//  forward one is M = embed(L_SS^{-1}) and the backward one is M^T, so the
//  row update is x = yhat . (M^T M).  The workgroup forms M column by column
//  (forward substitution on unit vectors under the same zero-pivot rule) and
//  then W = M^T M; the row pass is a GEMV against W.
// ---------------------------------------------------------------------------
struct CholArgs {
    const double* G[kMaxModes];
    int           N;
    int           n;
    int           R;
    double        ridge;
    double*       W;          // RP x RP out: V^{-1} (zeros outside R x R)
};

template <int RP>
__global__ __launch_bounds__(kBlock) void k_chol(CholArgs a) {
    __shared__ double V[RP * RP];     // lower: L ; strict upper: M^T (V[j][i] = M[i][j])
    __shared__ double Md[RP];         // diag of M
    __shared__ double s_inv;
    __shared__ double s_dmax;
    const int tid = threadIdx.x;
    const int R = a.R;
    for (int e = tid; e < RP * RP; e += kBlock) {
        const int r = e / RP, c = e % RP;
        double v = 0.0;
        if (r < R && c < R) {
            v = 1.0;
            for (int m = 0; m < a.N; ++m) {
                if (m == a.n) continue;
                v *= a.G[m][e];
            }
        }
        V[e] = v;
    }
    __syncthreads();
    if (tid == 0) {
        double dmax = 0.0;
        for (int r = 0; r < R; ++r) dmax = fmax(dmax, V[r * RP + r]);
        s_dmax = dmax;
    }
    __syncthreads();
    const double eps = fmax(a.ridge * s_dmax, 0.0);
    if (tid < R) V[tid * RP + tid] += eps;
    __syncthreads();
    // right-looking factorisation, one row per thread: in step j thread i
    // forms V[i][j] - sum_{k<j} V[i][k] V[j][k] (thread j: the pivot), then
    // the below-diagonal entries are scaled by 1 / pivot
    for (int j = 0; j < R; ++j) {
        const int i = tid;
        double t = 0.0;
        if (i >= j && i < R) {
            t = V[i * RP + j];
            for (int k = 0; k < j; ++k) t -= V[i * RP + k] * V[j * RP + k];
        }
        if (i == j) {
            const double sp = (t < 0.0) ? 0.0 : t;
            const double d = sqrt(sp);
            V[j * RP + j] = d;
            s_inv = (d > 0.0) ? 1.0 / d : 0.0;
        }
        __syncthreads();
        if (i > j && i < R) V[i * RP + j] = t * s_inv;
        __syncthreads();
    }
    // column j of M: L m = e_j with zero pivots -> zero components
    for (int j = tid; j < R; j += kBlock) {
        const double dj = V[j * RP + j];
        const double mj = (dj > 0.0) ? 1.0 / dj : 0.0;
        Md[j] = mj;
        for (int i = j + 1; i < R; ++i) {
            double s = -V[i * RP + j] * mj;
            for (int k = j + 1; k < i; ++k) s -= V[i * RP + k] * V[j * RP + k];
            const double di = V[i * RP + i];
            V[j * RP + i] = (di > 0.0) ? s / di : 0.0;
        }
    }
    __syncthreads();
    // W[r][c] = sum_{i >= max(r, c)} M[i][r] M[i][c]
    for (int e = tid; e < RP * RP; e += kBlock) {
        const int r = e / RP, c = e % RP;
        double w = 0.0;
        if (r < R && c < R) {
            for (int i = (r > c ? r : c); i < R; ++i) {
                const double mr = (i == r) ? Md[r] : V[r * RP + i];
                const double mc = (i == c) ? Md[c] : V[c * RP + i];
                w += mr * mc;
            }
        }
        a.W[e] = w;
    }
}

// ---------------------------------------------------------------------------
//  X = Yhat W per row (FP64 GEMV, W = V^{-1} in LDS), column sums of x^2 per
//  workgroup in FP64.
// ---------------------------------------------------------------------------
template <int RP>
__global__ __launch_bounds__(kSolveBlock) void k_solve(const float* __restrict__ Yhat,
                                                       float* __restrict__ X,
                                                       const double* __restrict__ Wg,
                                                       uint32_t rows,
                                                       double* __restrict__ sq_part) {
    // four lanes per row, lane q of a row owning columns [q CL, (q + 1) CL)
    constexpr int CL = RP / 4;
    __shared__ double W[RP * RP];
    __shared__ double red[kSolveBlock / 64][RP];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int q = lane & 3;
    const int g = lane >> 2;          // row of the wave's 16
    for (int e = tid; e < RP * RP; e += kSolveBlock) W[e] = Wg[e];
    __syncthreads();
    double sq[CL];
#pragma unroll
    for (int c = 0; c < CL; ++c) sq[c] = 0.0;

    // NRW rows per quad share every W element read from LDS (the LDS reads,
    // RP doubles per row and lane, bound the kernel at rank 64); each row's
    // products are formed in the same order as with one row.  Wave-uniform
    // trip count, so the cross-lane reductions see all lanes.
    constexpr int NRW = (RP == 64) ? 2 : 1;
    const uint32_t stride = gridDim.x * (kSolveBlock / 4) * NRW;
    for (uint32_t b = (blockIdx.x * (kSolveBlock / 4) + (uint32_t)(wave * 16)) * NRW; b < rows;
         b += stride) {
        uint32_t i[NRW];
        bool ok[NRW];
        const float* yrow[NRW];
        double x[NRW][CL];
#pragma unroll
        for (int r = 0; r < NRW; ++r) {
            i[r] = b + (uint32_t)(r * 16) + (uint32_t)g;
            ok[r] = i[r] < rows;
            yrow[r] = Yhat + (size_t)(ok[r] ? i[r] : 0) * RP;
#pragma unroll
            for (int c = 0; c < CL; ++c) x[r][c] = 0.0;
        }
        for (int r4 = 0; r4 < RP / 4; ++r4) {
            double yv[NRW][4];
#pragma unroll
            for (int r = 0; r < NRW; ++r) {
                float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
                if (ok[r]) v = *reinterpret_cast<const float4*>(yrow[r] + r4 * 4);
                yv[r][0] = (double)v.x;
                yv[r][1] = (double)v.y;
                yv[r][2] = (double)v.z;
                yv[r][3] = (double)v.w;
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const double* wr = W + (r4 * 4 + k) * RP + q * CL;
#pragma unroll
                for (int c = 0; c < CL; ++c) {
                    const double wc = wr[c];
#pragma unroll
                    for (int r = 0; r < NRW; ++r) x[r][c] = fma(yv[r][k], wc, x[r][c]);
                }
            }
        }
#pragma unroll
        for (int r = 0; r < NRW; ++r) {
            if (ok[r]) {
#pragma unroll
                for (int c4 = 0; c4 < CL / 4; ++c4) {
                    float4 v = make_float4((float)x[r][c4 * 4 + 0], (float)x[r][c4 * 4 + 1],
                                           (float)x[r][c4 * 4 + 2], (float)x[r][c4 * 4 + 3]);
                    *reinterpret_cast<float4*>(X + (size_t)i[r] * RP + q * CL + c4 * 4) = v;
                }
#pragma unroll
                for (int c = 0; c < CL; ++c) sq[c] = fma(x[r][c], x[r][c], sq[c]);
            }
        }
    }
    // column sums of x^2: over the 16 rows of the wave (lanes q, q+4, ...)
#pragma unroll
    for (int c = 0; c < CL; ++c) {
        double t = sq[c];
#pragma unroll
        for (int o = 4; o < 64; o <<= 1) t += __shfl_xor(t, o, 64);
        sq[c] = t;
    }
    if (g == 0) {
#pragma unroll
        for (int c = 0; c < CL; ++c) red[wave][q * CL + c] = sq[c];
    }
    __syncthreads();
    if (tid < RP) {
        double t = 0.0;
        for (int w = 0; w < kSolveBlock / 64; ++w) t += red[w][tid];
        sq_part[(size_t)blockIdx.x * RP + tid] = t;
    }
}

// fit: out[0] = fit, [1] = residual, [2] = <T,M>, [3] = ||M||^2
struct FitArgs {
    const double* G[kMaxModes];
    int           N;
    int           R;
    int           RP;
    const double* lambda;
    const double* ipart;      // [nwg][RP] <Y, Yhat> column partials, or null
    uint32_t      nwg;
    const double* inner_dev;  // used when ipart is null (forward-model dot)
    double        tnorm2;
    double*       out;
};

// sum of one value per thread in a fixed order: butterfly within each wave,
// then the waves in order; every thread gets the result
__device__ __forceinline__ double block_sum(double v, double* red) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) t += red[w];
    return t;
}

__global__ __launch_bounds__(kBlock) void k_fit(FitArgs a) {
    __shared__ double part[kBlock];
    __shared__ double red[kBlock / 64];
    const int tid = threadIdx.x;
    // <T, M> = sum_r lambda_r * colsum_r; column sums over the partial blocks
    // in strided subsets, subsets combined in order
    double inner;
    if (a.ipart) {
        const int NS = kBlock / a.RP;
        const int col = tid % a.RP, sub = tid / a.RP;
        part[tid] = (sub < NS) ? strided_sum(a.ipart, (uint32_t)sub, (uint32_t)NS, a.nwg,
                                             (size_t)a.RP, col)
                               : 0.0;
        __syncthreads();
        double v = 0.0;
        if (tid < a.R) {
            double cs = 0.0;
            for (int j = 0; j < NS; ++j) cs += part[j * a.RP + tid];
            v = a.lambda[tid] * cs;
        }
        inner = block_sum(v, red);
    } else {
        inner = *a.inner_dev;
    }
    // ||M||^2 = sum_{r,c} lambda_r lambda_c prod_n G_n[r, c]
    double rs = 0.0;
    for (int e = tid; e < a.R * a.R; e += kBlock) {
        const int r = e / a.R, c = e % a.R;
        double h = 1.0;
        for (int n = 0; n < a.N; ++n) h *= a.G[n][(size_t)r * a.RP + c];
        rs += a.lambda[r] * (a.lambda[c] * h);
    }
    const double mn = block_sum(rs, red);
    if (tid == 0) {
        // ||T - M||^2 >= 0 holds exactly; a negative value within the rounding
        // of its three terms is clamped to zero (the reference's clamp,
        // src/dynasor_cpals.cpp:345-346).  A value further below zero, or a
        // non-finite term, means the model has left the range where the three
        // terms can be formed (a diverged update): the fit and residual are
        // then reported as NaN rather than as a perfect fit.
        double r2 = a.tnorm2 - 2.0 * inner + mn;
        const double scale = fmax(a.tnorm2, fmax(2.0 * fabs(inner), fabs(mn)));
        const bool finite = isfinite(inner) && isfinite(mn) && isfinite(r2);
        double fit, resid;
        if (finite && r2 >= -kFitRoundRel * scale) {
            resid = sqrt(fmax(r2, 0.0));
            fit = 1.0 - resid / fmax(sqrt(a.tnorm2), 1e-30);
        } else {
            resid = __longlong_as_double(0x7FF8000000000000ll);
            fit = resid;
        }
        a.out[0] = fit;
        a.out[1] = resid;
        a.out[2] = inner;
        a.out[3] = mn;
    }
}

// ---------------------------------------------------------------------------
//  host wrappers
// ---------------------------------------------------------------------------
static uint32_t gram_wgs(uint32_t rows, int RP) {
    // ~64 rows per workgroup: several hundred workgroups for the largest
    // factors (the pass is latency-bound), and the partial Grams stay small
    const uint32_t per = 64u;
    uint32_t want = (rows + per - 1) / per;
    return std::max<uint32_t>(1, std::min<uint32_t>(want, kMaxGramWG));
}

struct SolveParts {
    const double* part = nullptr;
    uint32_t      nwg = 0;
};

template <int RP>
static int scale_gram_rp(dyn_gpu_ctx* c, int mode, const SolveParts* sq, bool inner,
                         double** ipart_out, uint32_t* nwg_out) {
    ScaleArgs a;
    a.Y = c->Y[mode];
    // scaling rewrites Y: the gathered FP32 / BF16 operands follow in the same pass
    a.Yg = sq ? c->Yg[mode] : nullptr;
    a.Ygb = sq ? c->Ygb[mode] : nullptr;
    a.relab = sq ? c->relab[mode] : nullptr;
    a.sq_part = sq ? sq->part : nullptr;
    a.sq_nwg = sq ? sq->nwg : 0;
    a.R = c->R;
    a.lambda_out = c->lambda;
    a.inv_out = c->inv_lambda;
    a.Yhat = inner ? c->Yhat[mode] : nullptr;
    a.rows = c->dims[mode];
    a.nwg = gram_wgs(a.rows, RP);
    void* sp = nullptr;
    DYN_CHECK(scratch2(c, (size_t)a.nwg * RP * RP * sizeof(float) +
                              (size_t)a.nwg * RP * sizeof(double) + 256,
                       &sp));
    a.ipart = (double*)sp;
    a.gpart = (float*)((char*)sp + (size_t)a.nwg * RP * sizeof(double));
    hipLaunchKernelGGL(k_scale_gram<RP>, dim3(a.nwg), dim3(kBlock), 0, c->stream, a);
    DYN_CHECK(check_launch(c, "scale_gram"));
    hipLaunchKernelGGL(k_gram_reduce<RP>, dim3((RP * RP + 15) / 16), dim3(kBlock), 0,
                       c->stream, a.gpart, a.nwg, c->G[mode]);
    DYN_CHECK(check_launch(c, "gram_reduce"));
    if (ipart_out) *ipart_out = a.ipart;
    if (nwg_out) *nwg_out = a.nwg;
    return DYN_OK;
}

static int scale_gram(dyn_gpu_ctx* c, int mode, const SolveParts* sq, bool inner,
                      double** ipart_out, uint32_t* nwg_out) {
    switch (c->Rp) {
        case 16: return scale_gram_rp<16>(c, mode, sq, inner, ipart_out, nwg_out);
        case 32: return scale_gram_rp<32>(c, mode, sq, inner, ipart_out, nwg_out);
        case 64: return scale_gram_rp<64>(c, mode, sq, inner, ipart_out, nwg_out);
        case 128: return scale_gram_rp<128>(c, mode, sq, inner, ipart_out, nwg_out);
        default: return set_err(c, DYN_ERR_UNSUPPORTED, "unsupported padded rank %d", c->Rp);
    }
}

int gram_mode(dyn_gpu_ctx* c, int mode) {
    return scale_gram(c, mode, nullptr, false, nullptr, nullptr);
}

// V_n^{-1} into c->L on `stream`
static int chol(dyn_gpu_ctx* c, int n, double ridge, hipStream_t stream) {
    CholArgs a;
    for (int m = 0; m < c->N; ++m) a.G[m] = c->G[m];
    a.N = c->N;
    a.n = n;
    a.R = c->R;
    a.ridge = ridge;
    a.W = c->L;
    switch (c->Rp) {
        case 16: hipLaunchKernelGGL(k_chol<16>, dim3(1), dim3(kBlock), 0, stream, a); break;
        case 32: hipLaunchKernelGGL(k_chol<32>, dim3(1), dim3(kBlock), 0, stream, a); break;
        case 64: hipLaunchKernelGGL(k_chol<64>, dim3(1), dim3(kBlock), 0, stream, a); break;
        case 128: hipLaunchKernelGGL(k_chol<128>, dim3(1), dim3(kBlock), 0, stream, a); break;
        default: return set_err(c, DYN_ERR_UNSUPPORTED, "unsupported padded rank %d", c->Rp);
    }
    return check_launch(c, "chol");
}

static int solve(dyn_gpu_ctx* c, int n, SolveParts* out) {
    const uint32_t rows = c->dims[n];
    const uint32_t nwg = std::max<uint32_t>(
        1, std::min<uint32_t>((rows + kSolveBlock / 4 - 1) / (kSolveBlock / 4), kMaxSolveWG));
    void* sp = nullptr;
    DYN_CHECK(scratch(c, (size_t)nwg * c->Rp * sizeof(double), &sp));
    double* part = (double*)sp;
    switch (c->Rp) {
        case 16: hipLaunchKernelGGL(k_solve<16>, dim3(nwg), dim3(kSolveBlock), 0, c->stream, c->Yhat[n], c->Y[n], c->L, rows, part); break;
        case 32: hipLaunchKernelGGL(k_solve<32>, dim3(nwg), dim3(kSolveBlock), 0, c->stream, c->Yhat[n], c->Y[n], c->L, rows, part); break;
        case 64: hipLaunchKernelGGL(k_solve<64>, dim3(nwg), dim3(kSolveBlock), 0, c->stream, c->Yhat[n], c->Y[n], c->L, rows, part); break;
        case 128: hipLaunchKernelGGL(k_solve<128>, dim3(nwg), dim3(kSolveBlock), 0, c->stream, c->Yhat[n], c->Y[n], c->L, rows, part); break;
        default: return set_err(c, DYN_ERR_UNSUPPORTED, "unsupported padded rank %d", c->Rp);
    }
    out->part = part;
    out->nwg = nwg;
    return check_launch(c, "solve");
}

static int fit_eval(dyn_gpu_ctx* c, const double* ipart, uint32_t nwg, const double* inner_dev,
                    double tnorm2, double* out) {
    FitArgs a;
    for (int m = 0; m < c->N; ++m) a.G[m] = c->G[m];
    a.N = c->N;
    a.R = c->R;
    a.RP = c->Rp;
    a.lambda = c->lambda;
    a.ipart = ipart;
    a.nwg = nwg;
    a.inner_dev = inner_dev;
    a.tnorm2 = tnorm2;
    a.out = out;
    hipLaunchKernelGGL(k_fit, dim3(1), dim3(kBlock), 0, c->stream, a);
    return check_launch(c, "fit");
}

}  // namespace dyn

using namespace dyn;

// ===========================================================================
//  C-ABI: CP-ALS
// ===========================================================================
extern "C" int dyn_gpu_cpals(dyn_gpu_ctx* c, const dyn_cp_opts* o, dyn_cp_result* res) {
    if (!c || !o || !res) return DYN_ERR_INVALID;
    if (c->N == 0) return set_err(c, DYN_ERR_STATE, "no tensor uploaded");
    if (!c->Y[0]) return set_err(c, DYN_ERR_STATE, "factors not set");
    if (o->kernel != DYN_CP_GAUSS_SEIDEL && o->kernel != DYN_CP_JACOBI)
        return set_err(c, DYN_ERR_INVALID, "bad kernel %d", o->kernel);
    if (o->precision != DYN_PREC_FP32 && o->precision != DYN_PREC_BF16)
        return set_err(c, DYN_ERR_INVALID, "bad precision %d", o->precision);
    if (o->max_iters < 0) return set_err(c, DYN_ERR_INVALID, "max_iters < 0");
    DYN_HIP(c, hipSetDevice(c->device));
    using clk = std::chrono::steady_clock;
    auto secs = [](clk::time_point a, clk::time_point b) {
        return std::chrono::duration<double>(b - a).count();
    };
    const int N = c->N;
    bool jacobi = o->kernel == DYN_CP_JACOBI;
    if (c->streamed && !jacobi) {
        // src/dynasor_cpals.cpp:223-231: the streaming driver is all-modes only
        std::fprintf(stderr,
                     "[cpals] WARN: the tensor is streamed from host memory; Gauss-Seidel has "
                     "no streaming driver: running Jacobi (all-modes) instead\n");
        jacobi = true;
    }
    if (c->use_linear && !jacobi) {
        // src/dynasor_cpals.cpp:232-240: the linearised layout is all-modes only
        std::fprintf(stderr,
                     "[cpals] WARN: the tensor is linearised (Z-order); Gauss-Seidel needs the "
                     "per-mode copies: running Jacobi (all-modes) instead\n");
        jacobi = true;
    }
    // the linearised pass gathers FP32 rows (it is bound by its scattered
    // additions, so BF16 operands would not shorten it)
    const bool bf = o->precision == DYN_PREC_BF16 && !c->use_linear;

    double* hist = res->fit_history;
    double* itim = res->iter_time_s;
    std::memset(res, 0, sizeof(*res));
    res->fit_history = hist;
    res->iter_time_s = itim;

    auto t_r0 = clk::now();
    if (c->use_linear) {
        if (!c->lin.built) DYN_CHECK(build_linear(c));
    } else {
        for (int m = 0; m < N && !c->streamed; ++m)
            if (!c->copy[m].built) DYN_CHECK(remap_mode(c, m, m == 0 ? 1 : 0));
    }
    DYN_HIP(c, hipStreamSynchronize(c->stream));
    res->remap_time_s = secs(t_r0, clk::now());

    double tn2 = 0.0;
    DYN_CHECK(tensor_norm2(c, &tn2));
    res->tensor_norm = std::sqrt(tn2);

    // initial Grams, lambda = 1, gathered operands
    for (int m = 0; m < N; ++m) DYN_CHECK(gram_mode(c, m));
    {
        std::vector<double> ones(c->Rp, 0.0);
        for (int r = 0; r < c->R; ++r) ones[r] = 1.0;
        DYN_HIP(c, hipMemcpyAsync(c->lambda, ones.data(), c->Rp * sizeof(double),
                                  hipMemcpyHostToDevice, c->stream));
    }
    DYN_CHECK(ensure_gather(c));
    DYN_HIP(c, hipStreamSynchronize(c->stream));

    if (o->verbose) {
        std::printf("[cpals] modes=%d  rank=%d  padded=%d  nnz=%llu  ||T|| = %.6e  "
                    "max_iters=%d  tol=%.1e  kernel=%s  precision=%s\n",
                    N, c->R, c->Rp, (unsigned long long)c->nnz, res->tensor_norm, o->max_iters,
                    o->tol, jacobi ? "jacobi" : "gauss-seidel", bf ? "bf16-gather" : "fp32");
    }

    // With tol <= 0 the run never stops early, so the fits stay on the device
    // and are read once at the end (no host round trip per iteration); the
    // per-iteration times then come from events.
    const int iters = std::max(o->max_iters, 0);
    const bool deferred = !(o->tol > 0.0) && iters > 0;
    const int per = 2 * N + 1;   // iteration start, then a begin/end pair per spMTTKRP
    std::vector<hipEvent_t> evs((size_t)std::max(iters, 1) * per + 1, nullptr);
    double* hist_dev = nullptr;
    // the Hadamard / Cholesky / inverse of mode n depends only on the Grams,
    // so it runs on the auxiliary stream while mode n's spMTTKRP runs
    hipEvent_t ev_g = nullptr, ev_l = nullptr;
    auto destroy_events = [&]() {
        for (hipEvent_t e : evs)
            if (e) (void)hipEventDestroy(e);
        if (ev_g) (void)hipEventDestroy(ev_g);
        if (ev_l) (void)hipEventDestroy(ev_l);
        if (hist_dev) dfree(c, hist_dev);
    };
    {
        bool ok = hipEventCreateWithFlags(&ev_g, hipEventDisableTiming) == hipSuccess &&
                  hipEventCreateWithFlags(&ev_l, hipEventDisableTiming) == hipSuccess;
        for (size_t i = 0; ok && i < evs.size(); ++i)
            ok = hipEventCreateWithFlags(&evs[i], 0) == hipSuccess;
        int arc = ok ? DYN_OK : set_err(c, DYN_ERR_HIP, "event creation failed");
        if (arc == DYN_OK && deferred) arc = dalloc_t(c, &hist_dev, (size_t)iters * 4);
        if (arc != DYN_OK) {
            destroy_events();
            return arc;
        }
    }

    int rc = DYN_OK;
    double fit_prev = 0.0;
    auto t_start = clk::now();
    double* inner_dev = c->fitbuf + 8;
    auto mttkrp_secs = [&](const hipEvent_t* E, int nev) {
        double mt = 0.0;
        for (int i = 0; i < nev; ++i) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, E[1 + 2 * i], E[2 + 2 * i]) == hipSuccess) {
                mt += ms * 1e-3;
                // Gauss-Seidel pairs bracket exactly one segment kernel
                if (c->timing && !jacobi) {
                    c->timed_extra_ms += ms;
                    c->timed_extra_n += 1;
                }
            }
        }
        return mt;
    };
    // a NaN fit marks a model whose fit terms are no longer finite or
    // consistent (k_fit); it never counts as converged, and is reported once
    bool warned = false;
    auto warn_fit = [&](int iter, double fit) {
        if (std::isfinite(fit) || warned) return;
        warned = true;
        std::fprintf(stderr,
                     "[cpals] WARN: iteration %d: the model diverged (||T - M||^2 is not "
                     "finite or is negative beyond rounding); fit reported as NaN\n",
                     iter);
    };
    // Jacobi: the fit of iteration t is formed from the spMTTKRPs of iteration
    // t + 1, which read exactly the model of iteration t: for any mode s,
    // <T, M_t> = sum_r lambda_r <Y_s[:, r], Yhat_s[:, r]> with Yhat_s taken from
    // the factors of M_t (the Gauss-Seidel identity with a consistent Yhat).
    // So no forward pass is needed per iteration: the fit of iteration t is
    // evaluated after the next sweep (before any factor of M_t is replaced),
    // and the last iteration's fit costs one extra spMTTKRP of mode s.
    int s_mode = 0;
    for (int m = 1; m < N; ++m)
        if (c->dims[m] < c->dims[s_mode]) s_mode = m;
    const void* Yin[kMaxModes];
    const void* Yin32[kMaxModes];
    for (int m = 0; m < N; ++m) {
        Yin[m] = bf ? (const void*)c->Ygb[m] : (const void*)c->Yg[m];
        Yin32[m] = c->Yg[m];
    }
    // With BF16 gathers the mode whose Yhat carries the fit (Gauss-Seidel: the
    // last mode; Jacobi: s_mode) still gathers FP32 rows, so the reported fit
    // is the fit of the FP32 model that the run returns.
    const int fit_mode = jacobi ? s_mode : N - 1;
    auto mttkrp_n = [&](int n) {
        const bool f32 = bf && n == fit_mode;
        const int prec = (f32 || !bf) ? DYN_PREC_FP32 : DYN_PREC_BF16;
        return mttkrp_mode(c, n, prec, f32 ? Yin32 : Yin, nullptr, c->Yhat[n], c->Rp);
    };
    int precs[kMaxModes];
    for (int m = 0; m < N; ++m) precs[m] = (bf && m != fit_mode) ? DYN_PREC_BF16 : DYN_PREC_FP32;
    // the spMTTKRPs of the modes in `mask` (Jacobi), in core or streamed
    auto sweep = [&](uint32_t mask) -> int {
        if (c->streamed) return stream_sweep(c, mask, precs);
        if (c->use_linear) return linear_sweep(c, mask);
        for (int m = 0; m < N; ++m)
            if ((mask >> m) & 1u) DYN_CHECK(mttkrp_n(m));
        return DYN_OK;
    };
    // fit of the current device model from Yhat[s_mode] (Jacobi) into `out`
    auto jacobi_fit = [&](double* out) -> int {
        double* ipart = nullptr;
        uint32_t inwg = 0;
        DYN_CHECK(scale_gram(c, s_mode, nullptr, true, &ipart, &inwg));
        return fit_eval(c, ipart, inwg, inner_dev, tn2, out);
    };
    // per-mode updates of one iteration (Gauss-Seidel runs each spMTTKRP
    // here, Jacobi before); Gauss-Seidel's last mode also forms the fit
    auto update_modes = [&](hipEvent_t* E, int& nev, double* fit_out) -> int {
        double* ipart = nullptr;
        uint32_t inwg = 0;
        for (int n = 0; n < N; ++n) {
            if (jacobi) {
                // nothing to overlap with (mode n's V needs mode n - 1's new
                // Gram): V_n^{-1} in order on the engine stream, without the
                // cross-stream hand-offs
                DYN_CHECK(chol(c, n, o->ridge, c->stream));
            } else {
                // Grams current -> V_n^{-1} on the auxiliary stream while
                // mode n's spMTTKRP runs
                if (hipEventRecord(ev_g, c->stream) != hipSuccess ||
                    hipStreamWaitEvent(c->aux_stream, ev_g, 0) != hipSuccess)
                    return set_err(c, DYN_ERR_HIP, "event record/wait failed");
                DYN_CHECK(chol(c, n, o->ridge, c->aux_stream));
                if (hipEventRecord(ev_l, c->aux_stream) != hipSuccess)
                    return set_err(c, DYN_ERR_HIP, "event record failed");
                c->seg_pair = &E[1 + 2 * n];
                const int mrc = mttkrp_n(n);
                c->seg_pair = nullptr;
                nev = n + 1;
                DYN_CHECK(mrc);
                if (hipStreamWaitEvent(c->stream, ev_l, 0) != hipSuccess)
                    return set_err(c, DYN_ERR_HIP, "stream wait failed");
            }
            SolveParts sq;
            DYN_CHECK(solve(c, n, &sq));
            const bool last = (n == N - 1) && !jacobi;
            DYN_CHECK(scale_gram(c, n, &sq, last, last ? &ipart : nullptr, last ? &inwg : nullptr));
        }
        if (!jacobi) return fit_eval(c, ipart, inwg, inner_dev, tn2, fit_out);
        return DYN_OK;
    };
    auto read_fit = [&](int iter) -> int {
        hipError_t e = hipMemcpyAsync(c->fitbuf_host, c->fitbuf, 4 * sizeof(double),
                                      hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess)
            return set_err(c, DYN_ERR_HIP, "cpals iteration %d failed: %s", iter,
                           hipGetErrorString(e));
        return DYN_OK;
    };
    // one iteration's results, in order; returns true when it converged
    auto record = [&](int iter, double fit, double resid, double it, const hipEvent_t* E,
                      int nev) -> bool {
        warn_fit(iter, fit);
        if (res->fit_history) res->fit_history[iter] = fit;
        if (res->iter_time_s) res->iter_time_s[iter] = it;
        res->mttkrp_time_s += mttkrp_secs(E, nev);
        res->final_fit = fit;
        res->final_residual = resid;
        res->iters_done = iter + 1;
        if (o->verbose)
            std::printf("[cpals] iter %3d  fit = %.6f  dfit = %.2e  resid = %.3e  iter = %.3f s\n",
                        iter, fit, iter == 0 ? 0.0 : std::fabs(fit - fit_prev), resid, it);
        const bool conv = !deferred && iter > 0 && std::fabs(fit - fit_prev) < o->tol;
        if (conv && o->verbose)
            std::printf("[cpals] converged after %d iterations (|dfit|=%.2e < %.1e)\n", iter + 1,
                        std::fabs(fit - fit_prev), o->tol);
        fit_prev = fit;
        return conv;
    };
    auto ev_ms = [](hipEvent_t a, hipEvent_t b) {
        float ms = 0.f;
        return hipEventElapsedTime(&ms, a, b) == hipSuccess ? ms * 1e-3 : 0.0;
    };
    hipEvent_t E_end = evs.back();
    int nev = jacobi ? 1 : 0;
    int done = 0;           // iterations whose factor updates ran
    bool stopped = false;   // converged (Jacobi: before the updates of `done`)
    for (int iter = 0; iter < iters && rc == DYN_OK && !stopped; ++iter) {
        auto t0 = clk::now();
        hipEvent_t* E = &evs[(size_t)iter * per];
        (void)hipEventRecord(E[0], c->stream);
        if (jacobi) {
            (void)hipEventRecord(E[1], c->stream);
            rc = sweep((1u << N) - 1u);
            (void)hipEventRecord(E[2], c->stream);
            if (rc != DYN_OK) break;
            if (iter > 0) {
                // the fit of the previous iteration's model
                double* out = deferred ? hist_dev + (size_t)(iter - 1) * 4 : c->fitbuf;
                if ((rc = jacobi_fit(out)) != DYN_OK) break;
                if (!deferred) {
                    if ((rc = read_fit(iter)) != DYN_OK) break;
                    const hipEvent_t* Ep = &evs[(size_t)(iter - 1) * per];
                    if (record(iter - 1, c->fitbuf_host[0], c->fitbuf_host[1], ev_ms(Ep[0], E[0]),
                               Ep, nev)) {
                        stopped = true;
                        break;
                    }
                }
            }
            if ((rc = update_modes(E, nev, nullptr)) != DYN_OK) break;
            done = iter + 1;
            continue;
        }
        double* fit_out = deferred ? hist_dev + (size_t)iter * 4 : c->fitbuf;
        if ((rc = update_modes(E, nev, fit_out)) != DYN_OK) break;
        done = iter + 1;
        if (deferred) continue;
        if ((rc = read_fit(iter)) != DYN_OK) break;
        stopped = record(iter, c->fitbuf_host[0], c->fitbuf_host[1], secs(t0, clk::now()), E, nev);
    }
    if (rc == DYN_OK && done > 0) (void)hipEventRecord(E_end, c->stream);
    // Jacobi: the last updated model's fit, from one more spMTTKRP of s_mode
    if (rc == DYN_OK && jacobi && !stopped && done > 0) {
        double* out = deferred ? hist_dev + (size_t)(done - 1) * 4 : c->fitbuf;
        rc = sweep(1u << s_mode);
        if (rc == DYN_OK) rc = jacobi_fit(out);
        if (rc == DYN_OK && !deferred) {
            rc = read_fit(done - 1);
            if (rc == DYN_OK) {
                const hipEvent_t* Ep = &evs[(size_t)(done - 1) * per];
                (void)record(done - 1, c->fitbuf_host[0], c->fitbuf_host[1], ev_ms(Ep[0], E_end),
                             Ep, nev);
            }
        }
    }
    if (deferred && rc == DYN_OK && done > 0) {
        std::vector<double> h((size_t)done * 4);
        hipError_t e = hipMemcpyAsync(h.data(), hist_dev, h.size() * sizeof(double),
                                      hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            rc = set_err(c, DYN_ERR_HIP, "cpals failed: %s", hipGetErrorString(e));
        } else {
            for (int iter = 0; iter < done; ++iter) {
                const hipEvent_t* E = &evs[(size_t)iter * per];
                const hipEvent_t next = (iter + 1 < done) ? evs[(size_t)(iter + 1) * per] : E_end;
                (void)record(iter, h[(size_t)iter * 4], h[(size_t)iter * 4 + 1], ev_ms(E[0], next),
                             E, nev);
            }
        }
    }
    destroy_events();
    res->total_time_s = secs(t_start, clk::now());
    res->linalg_time_s = std::max(0.0, res->total_time_s - res->mttkrp_time_s);
    c->gather_valid = (rc == DYN_OK);
    if (rc == DYN_OK && o->verbose) {
        std::printf("[cpals] done  iters=%d  final_fit=%.6f  resid=%.3e\n", res->iters_done,
                    res->final_fit, res->final_residual);
    }
    return rc;
}

// ===========================================================================
//  C-ABI: forward / adjoint / completion
// ===========================================================================
static bool rp_ok(int ld) { return ld == 16 || ld == 32 || ld == 64 || ld == 128; }

extern "C" int dyn_gpu_forward(dyn_gpu_ctx* c, int use_lambda, float* m_out) {
    if (!c || !m_out) return DYN_ERR_INVALID;
    if (!c->Y[0]) return set_err(c, DYN_ERR_STATE, "factors not set");
    DYN_CHECK(reject_streamed(c, "dyn_gpu_forward"));
    DYN_HIP(c, hipSetDevice(c->device));
    if (c->nnz == 0) return DYN_OK;
    void* sp = nullptr;
    DYN_CHECK(scratch(c, c->nnz * sizeof(float), &sp));
    float* Ys[kMaxModes];
    for (int m = 0; m < c->N; ++m) Ys[m] = c->Y[m];
    DYN_CHECK(forward_eval(c, Ys, c->Rp, use_lambda ? c->lambda : nullptr, (float*)sp, nullptr,
                           nullptr));
    DYN_HIP(c, hipMemcpyAsync(m_out, sp, c->nnz * sizeof(float), hipMemcpyDeviceToHost,
                              c->stream));
    DYN_HIP(c, hipStreamSynchronize(c->stream));
    return DYN_OK;
}

extern "C" int dyn_gpu_adjoint(dyn_gpu_ctx* c, const float* g, int mode, float* grad_out,
                               int ld) {
    if (!c || !g || !grad_out) return DYN_ERR_INVALID;
    if (!c->Y[0]) return set_err(c, DYN_ERR_STATE, "factors not set");
    DYN_CHECK(reject_streamed(c, "dyn_gpu_adjoint"));
    if (mode < 0 || mode >= c->N) return set_err(c, DYN_ERR_INVALID, "bad mode %d", mode);
    if (ld < c->R) return set_err(c, DYN_ERR_INVALID, "ld %d < rank %d", ld, c->R);
    DYN_HIP(c, hipSetDevice(c->device));
    if (!c->copy[mode].built) DYN_CHECK(remap_mode(c, mode, mode == 0 ? 1 : 0));
    const size_t rows = c->dims[mode];
    void* sp = nullptr;
    DYN_CHECK(scratch(c, std::max<uint64_t>(c->nnz, 1) * sizeof(float) +
                             rows * c->Rp * sizeof(float) + 256,
                      &sp));
    float* gd = (float*)sp;
    float* od = (float*)((char*)sp + ((std::max<uint64_t>(c->nnz, 1) * sizeof(float) + 255) & ~size_t(255)));
    if (c->nnz)
        DYN_HIP(c, hipMemcpyAsync(gd, g, c->nnz * sizeof(float), hipMemcpyHostToDevice,
                                  c->stream));
    DYN_CHECK(ensure_gather(c));
    const void* Yin[kMaxModes];
    for (int m = 0; m < c->N; ++m) Yin[m] = c->Yg[m];
    DYN_CHECK(mttkrp_mode(c, mode, DYN_PREC_FP32, Yin, gd, od, c->Rp));
    std::vector<float> pad(rows * c->Rp);
    DYN_HIP(c, hipMemcpyAsync(pad.data(), od, pad.size() * sizeof(float), hipMemcpyDeviceToHost,
                              c->stream));
    DYN_HIP(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < rows; ++i)
        std::memcpy(grad_out + i * (size_t)ld, &pad[i * c->Rp], sizeof(float) * c->R);
    return DYN_OK;
}

extern "C" int dyn_gpu_forward_dev(dyn_gpu_ctx* c, const float* const* A, int ld, float* m) {
    if (!c || !A || !m) return DYN_ERR_INVALID;
    if (c->N == 0) return set_err(c, DYN_ERR_STATE, "no tensor uploaded");
    DYN_CHECK(reject_streamed(c, "dyn_gpu_forward_dev"));
    if (!rp_ok(ld)) return set_err(c, DYN_ERR_INVALID, "ld must be 16, 32, 64 or 128 (got %d)", ld);
    DYN_HIP(c, hipSetDevice(c->device));
    if (c->nnz == 0) return DYN_OK;
    return forward_eval(c, A, ld, nullptr, m, nullptr, nullptr);
}

extern "C" int dyn_gpu_adjoint_dev(dyn_gpu_ctx* c, const float* g, int mode,
                                   const float* const* A, int ld, float* grad) {
    if (!c || !g || !A || !grad) return DYN_ERR_INVALID;
    if (c->N == 0) return set_err(c, DYN_ERR_STATE, "no tensor uploaded");
    DYN_CHECK(reject_streamed(c, "dyn_gpu_adjoint_dev"));
    if (mode < 0 || mode >= c->N) return set_err(c, DYN_ERR_INVALID, "bad mode %d", mode);
    if (!rp_ok(ld)) return set_err(c, DYN_ERR_INVALID, "ld must be 16, 32, 64 or 128 (got %d)", ld);
    DYN_HIP(c, hipSetDevice(c->device));
    if (!c->copy[mode].built) DYN_CHECK(remap_mode(c, mode, mode == 0 ? 1 : 0));
    return adjoint_mode(c, mode, A, ld, g, grad);
}

extern "C" int dyn_gpu_completion_grad_dev(dyn_gpu_ctx* c, int loss, const float* const* A,
                                           int ld, float* const* grad, double* loss_out) {
    if (!c || !A || !grad) return DYN_ERR_INVALID;
    if (c->N == 0) return set_err(c, DYN_ERR_STATE, "no tensor uploaded");
    DYN_CHECK(reject_streamed(c, "dyn_gpu_completion_grad_dev"));
    if (loss != DYN_LOSS_LS && loss != DYN_LOSS_BERNOULLI)
        return set_err(c, DYN_ERR_INVALID, "bad loss %d", loss);
    if (!rp_ok(ld)) return set_err(c, DYN_ERR_INVALID, "ld must be 16, 32, 64 or 128 (got %d)", ld);
    DYN_HIP(c, hipSetDevice(c->device));
    for (int m = 0; m < c->N; ++m)
        if (!c->copy[m].built) DYN_CHECK(remap_mode(c, m, m == 0 ? 1 : 0));
    return completion_grad(c, loss, A, ld, grad, loss_out);
}
