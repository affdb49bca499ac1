// Host helpers: CP-ALS initial factors and the synthetic tensor generator.
#include "dynasor_host.h"

#include <omp.h>
#include <parallel/algorithm>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

namespace {

// splitmix64 finaliser (Steele, Lea, Flood 2014)
inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// counter-based stream: independent per (seed, purpose, mode)
struct Stream {
    uint64_t key;
    Stream(uint64_t seed, uint64_t purpose, uint64_t mode)
        : key(mix64(seed * 0x9E3779B97F4A7C15ULL + purpose * 0xD1B54A32D192ED03ULL +
                    mode * 0xABC98388FB8FAC03ULL + 0x2545F4914F6CDD1DULL)) {}
    uint64_t u64(uint64_t ctr) const { return mix64(key + (ctr + 1) * 0x9E3779B97F4A7C15ULL); }
    double u01(uint64_t ctr) const { return (double)(u64(ctr) >> 11) * 0x1.0p-53; }
    // standard normal from two uniforms (Box-Muller)
    double normal(uint64_t ctr) const {
        const double u1 = 1.0 - u01(2 * ctr);
        const double u2 = u01(2 * ctr + 1);
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

enum Purpose : uint64_t { kCoord = 1, kPerm = 2, kFactor = 3, kNoise = 4, kLabel = 5, kValue = 6 };

struct ModeSampler {
    uint32_t dim = 1;
    bool zipf = false;
    std::vector<double> cdf;
    std::vector<uint32_t> perm;
    Stream st{0, 0, 0};

    void init(uint32_t d, bool z, double alpha, uint64_t seed, int mode) {
        dim = d;
        zipf = z;
        st = Stream(seed, kCoord, (uint64_t)mode);
        if (!zipf) return;
        cdf.resize(d);
        double s = 0.0;
        for (uint32_t k = 0; k < d; ++k) {
            s += std::pow((double)(k + 1), -alpha);
            cdf[k] = s;
        }
        for (uint32_t k = 0; k < d; ++k) cdf[k] /= s;
        perm.resize(d);
        for (uint32_t k = 0; k < d; ++k) perm[k] = k;
        Stream ps(seed, kPerm, (uint64_t)mode);
        for (uint32_t i = d - 1; i > 0; --i) {
            const uint64_t j = (uint64_t)(((unsigned __int128)ps.u64(i) * (i + 1)) >> 64);
            std::swap(perm[i], perm[j]);
        }
    }
    uint32_t draw(uint64_t ctr) const {
        if (!zipf) {
            return (uint32_t)(((unsigned __int128)st.u64(ctr) * dim) >> 64);
        }
        const double u = st.u01(ctr);
        uint32_t k = (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
        if (k >= dim) k = dim - 1;
        return perm[k];
    }
};

template <typename K>
struct Packer {
    int N;
    int shift[8];
    int bits[8];
    explicit Packer(const dyn_gen_spec& s) : N(s.num_modes) {
        int total = 0;
        for (int n = 0; n < N; ++n) {
            int b = 0;
            while (b < 32 && ((uint64_t)1 << b) < s.dims[n]) ++b;
            bits[n] = b;
            total += b;
        }
        int sh = total;
        for (int n = 0; n < N; ++n) {
            sh -= bits[n];
            shift[n] = sh;
        }
    }
    K pack(const uint32_t* c) const {
        K k = 0;
        for (int n = 0; n < N; ++n) k |= (K)c[n] << shift[n];
        return k;
    }
    uint32_t unpack(K k, int n) const {
        if (bits[n] == 0) return 0;
        return (uint32_t)((k >> shift[n]) & (K)(((uint64_t)1 << bits[n]) - 1));
    }
};

template <typename K>
int generate(const dyn_gen_spec& s, uint32_t* const* idx_out, float* vals_out, double* vals64,
             double* stats) {
    const int N = s.num_modes;
    const uint64_t nnz = s.nnz;
    Packer<K> pk(s);
    std::vector<ModeSampler> ms(N);
    for (int n = 0; n < N; ++n)
        ms[n].init(s.dims[n], s.coord_dist == DYN_GEN_COORD_ZIPF, s.alpha, s.seed, n);

    struct Cand {
        K key;
        uint64_t draw;
    };
    std::vector<K> S;
    uint64_t D = 0;
    double f_est = 1.0;
    int rounds = 0;
    std::vector<Cand> cand, fresh;
    while (S.size() < nnz) {
        ++rounds;
        const uint64_t deficit = nnz - S.size();
        double want = (double)deficit / f_est * 1.15 + 4096.0;
        const double cap = 4.0 * (double)nnz + 1e6;
        uint64_t M = (uint64_t)std::min(want, cap);
        if (M < deficit) M = deficit;
        cand.resize(M);
#pragma omp parallel for schedule(static)
        for (int64_t j = 0; j < (int64_t)M; ++j) {
            const uint64_t d = D + (uint64_t)j;
            uint32_t c[8];
            for (int n = 0; n < N; ++n) c[n] = ms[n].draw(d);
            cand[j].key = pk.pack(c);
            cand[j].draw = d;
        }
        D += M;
        __gnu_parallel::sort(cand.begin(), cand.end(), [](const Cand& a, const Cand& b) {
            return a.key < b.key || (a.key == b.key && a.draw < b.draw);
        });
        // first draw of every key that is not already kept
        std::vector<uint8_t> keep(M);
#pragma omp parallel for schedule(static)
        for (int64_t j = 0; j < (int64_t)M; ++j) {
            const bool first = (j == 0) || cand[j].key != cand[j - 1].key;
            bool ok = first;
            if (ok && !S.empty()) ok = !std::binary_search(S.begin(), S.end(), cand[j].key);
            keep[j] = ok ? 1 : 0;
        }
        fresh.clear();
        for (uint64_t j = 0; j < M; ++j)
            if (keep[j]) fresh.push_back(cand[j]);
        f_est = std::max((double)fresh.size() / (double)M, 1e-3);
        if (fresh.size() > deficit) {
            // keep the earliest draws: sequential-draw semantics
            std::nth_element(fresh.begin(), fresh.begin() + deficit, fresh.end(),
                             [](const Cand& a, const Cand& b) { return a.draw < b.draw; });
            fresh.resize(deficit);
            __gnu_parallel::sort(fresh.begin(), fresh.end(),
                                 [](const Cand& a, const Cand& b) { return a.key < b.key; });
        }
        std::vector<K> merged(S.size() + fresh.size());
        std::vector<K> fk(fresh.size());
        for (size_t j = 0; j < fresh.size(); ++j) fk[j] = fresh[j].key;
        std::merge(S.begin(), S.end(), fk.begin(), fk.end(), merged.begin());
        S.swap(merged);
        if (rounds > 200) return 2;   // cannot reach nnz distinct coordinates
    }
    cand.clear();
    cand.shrink_to_fit();

#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < (int64_t)nnz; ++k)
        for (int n = 0; n < N; ++n) idx_out[n][k] = pk.unpack(S[k], n);

    // ---- values ----
    double rms = 0.0;
    if (s.value_kind == DYN_GEN_VAL_LOGNORMAL || s.value_kind == DYN_GEN_VAL_UNIFORM) {
        Stream vs(s.seed, kValue, 0);
#pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < (int64_t)nnz; ++k) {
            const double v = (s.value_kind == DYN_GEN_VAL_UNIFORM)
                                 ? vs.u01((uint64_t)k)
                                 : std::exp(s.mu + s.sigma * vs.normal((uint64_t)k));
            vals_out[k] = (float)v;
            if (vals64) vals64[k] = v;
        }
    } else {
        const int Rt = std::max(1, s.planted_rank);
        std::vector<std::vector<double>> A(N);
        for (int n = 0; n < N; ++n) {
            Stream fs(s.seed, kFactor, (uint64_t)n);
            A[n].resize((size_t)s.dims[n] * Rt);
            for (size_t e = 0; e < A[n].size(); ++e)
                A[n][e] = (s.planted_factor_dist == DYN_GEN_FACTOR_NORMAL)
                              ? std::sqrt(s.planted_var) * fs.normal(e)
                              : fs.u01(e);
        }
        std::vector<double> x(nnz);
        double ss = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : ss)
        for (int64_t k = 0; k < (int64_t)nnz; ++k) {
            double m = 0.0;
            for (int r = 0; r < Rt; ++r) {
                double p = 1.0;
                for (int n = 0; n < N; ++n) p *= A[n][(size_t)idx_out[n][k] * Rt + r];
                m += p;
            }
            x[k] = m;
            ss += m * m;
        }
        rms = nnz ? std::sqrt(ss / (double)nnz) : 0.0;
        Stream ns(s.seed, kNoise, 0), ls(s.seed, kLabel, 0);
        const double sd = s.noise_frac * rms;
#pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < (int64_t)nnz; ++k) {
            double v;
            if (s.value_kind == DYN_GEN_VAL_BERNOULLI) {
                const double p = 1.0 / (1.0 + std::exp(-x[k]));
                v = (ls.u01((uint64_t)k) < p) ? 1.0 : 0.0;
            } else {
                v = x[k] + (sd > 0.0 ? sd * ns.normal((uint64_t)k) : 0.0);
            }
            vals_out[k] = (float)v;
            if (vals64) vals64[k] = v;
        }
    }
    if (stats) {
        stats[0] = rounds;
        stats[1] = (double)D;
        std::vector<uint64_t> cnt(s.dims[0], 0);
        for (uint64_t k = 0; k < nnz; ++k) ++cnt[idx_out[0][k]];
        stats[2] = nnz ? (double)*std::max_element(cnt.begin(), cnt.end()) : 0.0;
        stats[3] = rms;
    }
    return 0;
}

// per-mode seed of the CP-ALS initialisation (src/dynasor_cpals.cpp:52-57)
uint64_t init_mode_seed(uint64_t seed, int n) {
    return mix64(seed + 0x9E3779B97F4A7C15ULL * (uint64_t)(n + 1));
}

}  // namespace

extern "C" void dyn_host_init_normal(int N, const uint32_t* dims, int R, int ld, uint64_t seed,
                                     float* const* Y) {
    for (int n = 0; n < N; ++n) {
        std::mt19937_64 rng(init_mode_seed(seed, n));
        std::normal_distribution<float> nd(0.0f, 1.0f);
        for (uint32_t i = 0; i < dims[n]; ++i) {
            float* row = Y[n] + (size_t)i * ld;
            for (int r = 0; r < R; ++r) row[r] = nd(rng);
            for (int r = R; r < ld; ++r) row[r] = 0.0f;
        }
    }
}

extern "C" void dyn_host_init_uniform(int N, const uint32_t* dims, int R, int ld, uint64_t seed,
                                      float* const* Y) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<float> ud(0.0f, 1.0f);
    for (int n = 0; n < N; ++n)
        for (uint32_t i = 0; i < dims[n]; ++i) {
            float* row = Y[n] + (size_t)i * ld;
            for (int r = 0; r < R; ++r) row[r] = ud(rng);
            for (int r = R; r < ld; ++r) row[r] = 0.0f;
        }
}

extern "C" int dyn_gen_tensor(const dyn_gen_spec* s, uint32_t* const* idx_out, float* vals_out,
                              double* vals64_out, double* stats_out, int threads) {
    if (!s || !idx_out || !vals_out) return 1;
    if (s->num_modes < 2 || s->num_modes > 8) return 1;
    long double cells = 1.0L;
    int total_bits = 0;
    for (int n = 0; n < s->num_modes; ++n) {
        if (s->dims[n] == 0) return 1;
        cells *= (long double)s->dims[n];
        int b = 0;
        while (b < 32 && ((uint64_t)1 << b) < s->dims[n]) ++b;
        total_bits += b;
    }
    if ((long double)s->nnz > cells) return 1;
    if (total_bits > 128) return 1;
    if (threads > 0) omp_set_num_threads(threads);
    if (total_bits <= 64) return generate<uint64_t>(*s, idx_out, vals_out, vals64_out, stats_out);
    return generate<unsigned __int128>(*s, idx_out, vals_out, vals64_out, stats_out);
}
