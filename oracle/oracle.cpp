// CPU oracle for the sparse CP-ALS path -- TEST INFRASTRUCTURE ONLY.
// Restates the reference algorithms (see oracle.h for the file:line map) in
// plain single-threaded C++ so the device engine can be checked against
// them.  Compiled with -std=c++17 (ISO mode: no floating-point contraction),
// the same as the reference build, so the FP32 algebra rounds identically.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
//  MTTKRP by definition (src/dynasor_reference.cpp:18-50): nonzeros in input
//  order, zero values skipped, tmp = v, tmp *= Y_w[i_w] for w ascending
//  (w != mode), then Yhat[i_mode] += tmp.
// ---------------------------------------------------------------------------
template <typename T>
void mttkrp_def(int N, uint64_t nnz, const uint32_t* dims, const float* vals,
                const uint32_t* const* idx, int R, int ld, const T* const* Y, int mode, T* out) {
    std::fill(out, out + (size_t)dims[mode] * ld, T(0));
    std::vector<T> tmp(R);
    for (uint64_t k = 0; k < nnz; ++k) {
        const T v = (T)vals[k];
        if (vals[k] == 0.0f) continue;
        for (int r = 0; r < R; ++r) tmp[r] = v;
        for (int w = 0; w < N; ++w) {
            if (w == mode) continue;
            const T* row = Y[w] + (size_t)idx[w][k] * ld;
            for (int r = 0; r < R; ++r) tmp[r] *= row[r];
        }
        T* o = out + (size_t)idx[mode][k] * ld;
        for (int r = 0; r < R; ++r) o[r] += tmp[r];
    }
}

// ---------------------------------------------------------------------------
//  Small dense algebra, single thread (src/dynasor_linalg.cpp).  Matrices are
//  R x R with leading dimension R.
// ---------------------------------------------------------------------------
template <typename T>
void gram(const T* Y, size_t rows, int R, T* G) {          // :47-99
    std::fill(G, G + (size_t)R * R, T(0));
    for (size_t i = 0; i < rows; ++i) {
        const T* yi = Y + i * R;
        for (int r = 0; r < R; ++r) {
            const T yir = yi[r];
            if (yir == T(0)) continue;
            T* Gr = G + (size_t)r * R;
            for (int c = 0; c < R; ++c) Gr[c] += yir * yi[c];
        }
    }
}

template <typename T>
void cholesky_lower(T* V, int R, T ridge) {                  // :160-200
    T dmax = 0;
    for (int r = 0; r < R; ++r) dmax = std::max(dmax, V[(size_t)r * R + r]);
    const T eps = std::max(ridge * dmax, T(0));
    for (int r = 0; r < R; ++r) V[(size_t)r * R + r] += eps;
    for (int j = 0; j < R; ++j) {
        T* Vj = V + (size_t)j * R;
        T s = Vj[j];
        for (int k = 0; k < j; ++k) s -= Vj[k] * Vj[k];
        if (s < T(0)) s = T(0);
        const T d = std::sqrt(s);
        Vj[j] = d;
        const T inv = (d > T(0)) ? (T(1) / d) : T(0);
        for (int i = j + 1; i < R; ++i) {
            T* Vi = V + (size_t)i * R;
            T t = Vi[j];
            for (int k = 0; k < j; ++k) t -= Vi[k] * Vj[k];
            Vi[j] = t * inv;
        }
    }
}

template <typename T>
void right_solve_chol(const T* L, int R, const T* B, size_t rows, T* X) {   // :210-246
    for (size_t i = 0; i < rows; ++i) {
        const T* bi = B + i * R;
        T* xi = X + i * R;
        for (int c = 0; c < R; ++c) xi[c] = bi[c];
        for (int r = 0; r < R; ++r) {
            const T* Lr = L + (size_t)r * R;
            T s = xi[r];
            for (int k = 0; k < r; ++k) s -= Lr[k] * xi[k];
            const T d = Lr[r];
            xi[r] = (d > T(0)) ? (s / d) : T(0);
        }
        for (int r = R - 1; r >= 0; --r) {
            T s = xi[r];
            for (int k = r + 1; k < R; ++k) s -= L[(size_t)k * R + r] * xi[k];
            const T d = L[(size_t)r * R + r];
            xi[r] = (d > T(0)) ? (s / d) : T(0);
        }
    }
}

template <typename T>
void col_norms(const T* Y, size_t rows, int R, T* norms) {   // :251-289
    std::vector<T> acc(R, T(0));
    for (size_t i = 0; i < rows; ++i)
        for (int c = 0; c < R; ++c) acc[c] += Y[i * R + c] * Y[i * R + c];
    for (int c = 0; c < R; ++c) {
        T s = 0;
        s += acc[c];
        norms[c] = std::sqrt(s);
    }
}

template <typename T>
void col_scale_inv(T* Y, size_t rows, int R, const T* norms) {   // :294-316
    std::vector<T> inv(R);
    const T eps = (T)1e-20f;
    for (int c = 0; c < R; ++c) inv[c] = (norms[c] > eps) ? (T(1) / norms[c]) : T(0);
    for (size_t i = 0; i < rows; ++i)
        for (int c = 0; c < R; ++c) Y[i * R + c] *= inv[c];
}

template <typename T>
double weighted_col_dot(const T* A, const T* B, size_t rows, int R, const T* w) {   // :324-366
    std::vector<double> acc(R, 0.0);
    for (size_t i = 0; i < rows; ++i)
        for (int c = 0; c < R; ++c) acc[c] += (double)A[i * R + c] * (double)B[i * R + c];
    double s = 0;
    for (int c = 0; c < R; ++c) {
        double colsum = 0;
        colsum += acc[c];
        s += (double)w[c] * colsum;
    }
    return s;
}

template <typename T>
double model_norm_sq(const std::vector<std::vector<T>>& G, const T* lambda, int R) {
    // src/dynasor_cpals.cpp:79-113: H = ones .* G_n in double, lambda^T H lambda
    std::vector<double> H((size_t)R * R, 1.0);
    for (const auto& Gn : G)
        for (size_t e = 0; e < H.size(); ++e) H[e] *= (double)Gn[e];
    double s = 0;
    for (int r = 0; r < R; ++r) {
        double rowsum = 0;
        for (int c = 0; c < R; ++c) rowsum += (double)lambda[c] * H[(size_t)r * R + c];
        s += (double)lambda[r] * rowsum;
    }
    return std::max(0.0, s);
}

// bfloat16 rounding of an FP32 value (round to nearest, ties to even, on the
// 16 dropped mantissa bits), returned as the FP32 value it represents.
float bf16_round(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    float y;
    std::memcpy(&y, &u, 4);
    return y;
}

template <typename T>
int cpals_t(int N, uint64_t nnz, const uint32_t* dims, const float* vals,
            const uint32_t* const* idx, int R, float* const* Yio, const orc_cp_opts* o,
            double* hist, double* lambda_out, double* final_fit) {
    std::vector<std::vector<T>> Y(N), Yh(N), G(N, std::vector<T>((size_t)R * R));
    for (int n = 0; n < N; ++n) {
        Y[n].resize((size_t)dims[n] * R);
        Yh[n].resize((size_t)dims[n] * R);
        for (size_t e = 0; e < Y[n].size(); ++e) Y[n][e] = (T)Yio[n][e];
        gram(Y[n].data(), dims[n], R, G[n].data());
    }
    std::vector<T> lambda(R, T(1));
    double tn2 = 0.0;
    for (uint64_t k = 0; k < nnz; ++k) tn2 += (double)vals[k] * (double)vals[k];
    const double tnorm = std::sqrt(tn2);
    std::vector<T> V((size_t)R * R);
    std::vector<const T*> Yp(N);
    double fit_prev = 0.0, fit = 0.0;
    int done = 0;
    std::vector<std::vector<T>> Yb(o->bf16_gather ? N : 0);
    // the engine's mixed mode gathers the mode whose Yhat carries the fit
    // (Gauss-Seidel: the last mode; Jacobi: the smallest mode) in FP32, so
    // the reported fit is that of the FP32 model
    int fit_mode = N - 1;
    if (o->jacobi) {
        fit_mode = 0;
        for (int m = 1; m < N; ++m)
            if (dims[m] < dims[fit_mode]) fit_mode = m;
    }
    auto mttkrp = [&](int n) {
        for (int w = 0; w < N; ++w) {
            Yp[w] = Y[w].data();
            if (o->bf16_gather && w != n && n != fit_mode) {
                Yb[w].resize(Y[w].size());
                for (size_t e = 0; e < Y[w].size(); ++e) Yb[w][e] = (T)bf16_round((float)Y[w][e]);
                Yp[w] = Yb[w].data();
            }
        }
        mttkrp_def<T>(N, nnz, dims, vals, idx, R, R, Yp.data(), n, Yh[n].data());
    };
    for (int iter = 0; iter < o->max_iters; ++iter) {
        if (o->jacobi)
            for (int n = 0; n < N; ++n) mttkrp(n);
        for (int n = 0; n < N; ++n) {
            if (!o->jacobi) mttkrp(n);
            for (int e = 0; e < R * R; ++e) V[e] = T(1);
            for (int m = 0; m < N; ++m)
                if (m != n)
                    for (int e = 0; e < R * R; ++e) V[e] *= G[m][e];
            cholesky_lower(V.data(), R, (T)o->ridge);
            right_solve_chol(V.data(), R, Yh[n].data(), dims[n], Y[n].data());
            col_norms(Y[n].data(), dims[n], R, lambda.data());
            col_scale_inv(Y[n].data(), dims[n], R, lambda.data());
            gram(Y[n].data(), dims[n], R, G[n].data());
        }
        double inner;
        if (o->jacobi && !o->stale_jacobi_fit) {
            inner = 0.0;
            for (uint64_t k = 0; k < nnz; ++k) {
                double m = 0.0;
                for (int r = 0; r < R; ++r) {
                    double p = (double)lambda[r];
                    for (int w = 0; w < N; ++w) p *= (double)Y[w][(size_t)idx[w][k] * R + r];
                    m += p;
                }
                inner += (double)vals[k] * m;
            }
        } else {
            inner = weighted_col_dot(Y[N - 1].data(), Yh[N - 1].data(), dims[N - 1], R,
                                     lambda.data());
            inner = (double)(T)inner;   // the reference returns value_t
        }
        const double mn = model_norm_sq(G, lambda.data(), R);
        double r2 = tn2 - 2.0 * inner + mn;
        if (r2 < 0.0) r2 = 0.0;
        fit = 1.0 - std::sqrt(r2) / std::max(tnorm, 1e-30);
        hist[iter] = fit;
        done = iter + 1;
        if (iter > 0 && std::fabs(fit - fit_prev) < o->tol) break;
        fit_prev = fit;
    }
    for (int n = 0; n < N; ++n)
        for (size_t e = 0; e < Y[n].size(); ++e) Yio[n][e] = (float)Y[n][e];
    for (int r = 0; r < R; ++r) lambda_out[r] = (double)lambda[r];
    *final_fit = fit;
    return done;
}

// splitmix64 finaliser of (seed + golden * (n + 1)): the per-mode seed of
// the CP-ALS initialisation (src/dynasor_cpals.cpp:52-57)
uint64_t mode_seed(uint64_t seed, int n) {
    uint64_t z = seed + 0x9E3779B97F4A7C15ULL * (uint64_t)(n + 1);
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ULL;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return z;
}

}  // namespace

extern "C" {

int orc_mttkrp(int N, uint64_t nnz, const uint32_t* dims, const float* vals,
               const uint32_t* const* idx, int R, int ld, const float* const* Y, int mode,
               float* out) {
    if (N < 2 || mode < 0 || mode >= N || R <= 0 || ld < R) return 1;
    mttkrp_def<float>(N, nnz, dims, vals, idx, R, ld, Y, mode, out);
    return 0;
}

int orc_mttkrp_f64(int N, uint64_t nnz, const uint32_t* dims, const float* vals,
                   const uint32_t* const* idx, int R, int ld, const double* const* Y, int mode,
                   double* out) {
    if (N < 2 || mode < 0 || mode >= N || R <= 0 || ld < R) return 1;
    mttkrp_def<double>(N, nnz, dims, vals, idx, R, ld, Y, mode, out);
    return 0;
}

void orc_init_normal(int N, const uint32_t* dims, int R, int ld, uint64_t seed,
                     float* const* Y) {
    for (int n = 0; n < N; ++n) {
        std::mt19937_64 rng(mode_seed(seed, n));
        std::normal_distribution<float> nd(0.0f, 1.0f);
        for (uint32_t i = 0; i < dims[n]; ++i) {
            float* row = Y[n] + (size_t)i * ld;
            for (int r = 0; r < R; ++r) row[r] = nd(rng);
            for (int r = R; r < ld; ++r) row[r] = 0.0f;
        }
    }
}

void orc_init_uniform(int N, const uint32_t* dims, int R, int ld, uint64_t seed,
                      float* const* Y) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<float> ud(0.0f, 1.0f);
    for (int n = 0; n < N; ++n) {
        for (uint32_t i = 0; i < dims[n]; ++i) {
            float* row = Y[n] + (size_t)i * ld;
            for (int r = 0; r < R; ++r) row[r] = ud(rng);
            for (int r = R; r < ld; ++r) row[r] = 0.0f;
        }
    }
}

int orc_remap(int N, uint64_t nnz, const uint32_t* dims, const uint32_t* const* idx, int mode,
              int secondary, uint32_t* perm, uint64_t* rowptr) {
    if (mode < 0 || mode >= N || secondary >= N || secondary == mode) return 1;
    std::vector<uint32_t> cur(nnz), nxt(nnz);
    for (uint64_t k = 0; k < nnz; ++k) cur[k] = (uint32_t)k;
    auto counting_pass = [&](const uint32_t* key, uint32_t dim) {
        std::vector<uint64_t> start((size_t)dim + 1, 0);
        for (uint64_t k = 0; k < nnz; ++k) ++start[(size_t)key[cur[k]] + 1];
        for (uint32_t r = 0; r < dim; ++r) start[r + 1] += start[r];
        for (uint64_t k = 0; k < nnz; ++k) nxt[start[key[cur[k]]]++] = cur[k];
        cur.swap(nxt);
    };
    if (secondary >= 0) counting_pass(idx[secondary], dims[secondary]);
    counting_pass(idx[mode], dims[mode]);
    std::memcpy(perm, cur.data(), nnz * sizeof(uint32_t));
    std::fill(rowptr, rowptr + (size_t)dims[mode] + 1, 0);
    for (uint64_t k = 0; k < nnz; ++k) ++rowptr[(size_t)idx[mode][k] + 1];
    for (uint32_t r = 0; r < dims[mode]; ++r) rowptr[r + 1] += rowptr[r];
    return 0;
}

int orc_cpals(int N, uint64_t nnz, const uint32_t* dims, const float* vals,
              const uint32_t* const* idx, int R, float* const* Y, const orc_cp_opts* opt,
              double* fit_hist, double* lambda_out, double* final_fit) {
    if (N < 2 || R <= 0 || !opt) return 0;
    if (opt->fp64)
        return cpals_t<double>(N, nnz, dims, vals, idx, R, Y, opt, fit_hist, lambda_out,
                               final_fit);
    return cpals_t<float>(N, nnz, dims, vals, idx, R, Y, opt, fit_hist, lambda_out, final_fit);
}

void orc_forward(int N, uint64_t nnz, const uint32_t* const* idx, int R, int ld,
                 const double* const* A, const double* lambda, double* m_out) {
    // every m_out[k] is formed by one thread in the same order, so the
    // result does not depend on the thread count
#pragma omp parallel for schedule(static)
    for (int64_t kk = 0; kk < (int64_t)nnz; ++kk) {
        const uint64_t k = (uint64_t)kk;
        double m = 0.0;
        for (int r = 0; r < R; ++r) {
            double p = lambda ? lambda[r] : 1.0;
            for (int n = 0; n < N; ++n) p *= A[n][(size_t)idx[n][k] * ld + r];
            m += p;
        }
        m_out[k] = m;
    }
}

double orc_completion_loss(int N, uint64_t nnz, const uint32_t* const* idx, const float* y,
                           int R, int ld, const double* const* A, int loss) {
    double total = 0.0;
    for (uint64_t k = 0; k < nnz; ++k) {
        double m = 0.0;
        for (int r = 0; r < R; ++r) {
            double p = 1.0;
            for (int n = 0; n < N; ++n) p *= A[n][(size_t)idx[n][k] * ld + r];
            m += p;
        }
        const double yk = y[k];
        if (loss == 1) {
            const double sp = std::max(m, 0.0) + std::log1p(std::exp(-std::fabs(m)));
            total += sp - yk * m;
        } else {
            total += 0.5 * (m - yk) * (m - yk);
        }
    }
    return total;
}

}  // extern "C"
