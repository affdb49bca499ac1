// Host containers and file formats: Tensor / FactorMatrices storage, the
// FROSTT .tns parser, the .dnb binary cache and DYFM factor dumps.  Formats
// are byte-compatible with the reference (src/dynasor_io.cpp:544-1348).
#include "dynasor_common.h"
#include "dynasor_host.h"
#include "dynasor_ooc.h"

#include <sys/stat.h>

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace dynasor {

namespace {

void* aligned_alloc64(size_t bytes) {
    if (bytes == 0) bytes = 64;
    bytes = (bytes + 63) & ~size_t(63);
    return std::aligned_alloc(64, bytes);
}

int round_up(int x, int m) { return (x + m - 1) / m * m; }

constexpr char kDnbMagic[4] = {'D', 'N', 'B', '1'};
constexpr size_t kDnbHeader = 4 + 4 + 4 + 4 + 4 + 8 + 4 * DYN_MAX_MODES + 4 * 7;   // 88 bytes

}  // namespace

// ---------------------------------------------------------------------------
Tensor::~Tensor() { free_elements(); }

bool Tensor::try_allocate_elements(uint64_t n, int nmodes) {
    free_elements();
    num_modes = nmodes;
    nnz = n;
    bool ok = (vals = (value_t*)aligned_alloc64(std::max<uint64_t>(n, 1) * sizeof(value_t)));
    for (int m = 0; m < nmodes; ++m)
        ok = (idx_buf[m] = (idx_t*)aligned_alloc64(std::max<uint64_t>(n, 1) * sizeof(idx_t))) && ok;
    if (!ok) {
        free_elements();
        num_modes = 0;
        nnz = 0;
        std::fprintf(stderr, "Tensor::allocate_elements: out of host memory for %llu nonzeros\n",
                     (unsigned long long)n);
    }
    return ok;
}

void Tensor::allocate_elements(uint64_t n, int nmodes) {
    // the reference ends the process on a failed allocation
    // (src/dynasor_io.cpp:149-212); the loaders and the C-ABI use
    // try_allocate_elements and report the failure instead
    if (!try_allocate_elements(n, nmodes)) std::exit(1);
}

void Tensor::free_elements() {
    std::free(vals);
    vals = nullptr;
    for (int m = 0; m < DYN_MAX_MODES; ++m) {
        std::free(idx_buf[m]);
        idx_buf[m] = nullptr;
    }
}

void FactorMatrices::allocate(int nmodes, const idx_t* mode_sizes, int R) {
    free_all();
    num_modes = nmodes;
    rank = R;
    rank_padded = std::max(round_up(R, 8), round_up(R, 16));
    Y.assign(nmodes, nullptr);
    Yhat.assign(nmodes, nullptr);
    for (int n = 0; n < nmodes; ++n) {
        const size_t bytes = (size_t)mode_sizes[n] * rank_padded * sizeof(value_t);
        Y[n] = (value_t*)aligned_alloc64(bytes);
        Yhat[n] = (value_t*)aligned_alloc64(bytes);
        if (!Y[n] || !Yhat[n]) {
            std::fprintf(stderr, "FactorMatrices::allocate: OOM for mode %d\n", n);
            std::exit(1);
        }
        std::memset(Y[n], 0, bytes);
        std::memset(Yhat[n], 0, bytes);
    }
}

void FactorMatrices::free_all() {
    for (auto* p : Y) std::free(p);
    for (auto* p : Yhat) std::free(p);
    Y.clear();
    Yhat.clear();
    num_modes = rank = rank_padded = 0;
}

void FactorMatrices::zero_output() {
    (void)0;   // sizes are not kept; callers re-allocate (reference semantics)
}

// ---------------------------------------------------------------------------
//  FROSTT text
// ---------------------------------------------------------------------------
static bool load_tns_text(const std::string& path, Tensor& T) {
    std::FILE* fp = std::fopen(path.c_str(), "rb");
    if (!fp) {
        std::fprintf(stderr, "ERROR: cannot open tensor file '%s' (%s)\n", path.c_str(),
                     std::strerror(errno));
        return false;
    }
    std::vector<float> vals;
    std::vector<std::vector<idx_t>> idx;
    idx_t sizes[DYN_MAX_MODES] = {0};
    int nm = 0, header_expect = 0;
    bool header_done = false;
    std::vector<char> buf(1 << 16);
    double tok[DYN_MAX_MODES + 1];
    bool ok = true;
    while (std::fgets(buf.data(), (int)buf.size(), fp)) {
        char* p = buf.data();
        while (*p && std::isspace((unsigned char)*p)) ++p;
        if (!*p || *p == '#') continue;
        int ntok = 0;
        while (*p) {
            while (*p && std::isspace((unsigned char)*p)) ++p;
            if (!*p) break;
            if (ntok >= DYN_MAX_MODES + 1) {
                std::fprintf(stderr, "ERROR: line has more than %d tokens\n", DYN_MAX_MODES + 1);
                ok = false;
                break;
            }
            char* end = nullptr;
            const double v = std::strtod(p, &end);
            if (end == p) {
                std::fprintf(stderr, "ERROR: cannot parse token near: %s\n", p);
                ok = false;
                break;
            }
            tok[ntok++] = v;
            p = end;
        }
        if (!ok) break;
        if (!header_done) {
            // optional SPLATT header: "N" then the N mode sizes
            if (ntok == 1 && nm == 0 && tok[0] == (double)(int)tok[0] && tok[0] >= 2 &&
                tok[0] <= DYN_MAX_MODES) {
                nm = (int)tok[0];
                header_expect = nm;
                idx.assign(nm, {});
                continue;
            }
            if (header_expect > 0 && ntok == header_expect) {
                for (int n = 0; n < ntok; ++n) sizes[n] = (idx_t)tok[n];
                header_expect = 0;
                header_done = true;
                continue;
            }
            header_done = true;
        }
        if (ntok < 2) continue;
        const int lm = ntok - 1;
        if (nm == 0) {
            if (lm > DYN_MAX_MODES) {
                std::fprintf(stderr, "ERROR: tensor has %d modes (max %d)\n", lm, DYN_MAX_MODES);
                ok = false;
                break;
            }
            nm = lm;
            idx.assign(nm, {});
        } else if (lm != nm) {
            std::fprintf(stderr, "ERROR: inconsistent mode count (line expects %d, got %d)\n", nm,
                         lm);
            ok = false;
            break;
        }
        for (int n = 0; n < nm; ++n) {
            const long long ix = (long long)tok[n];
            if (ix <= 0) {
                std::fprintf(stderr, "ERROR: non-positive index %lld\n", ix);
                ok = false;
                break;
            }
            idx[n].push_back((idx_t)(ix - 1));
            if ((idx_t)ix > sizes[n]) sizes[n] = (idx_t)ix;
        }
        if (!ok) break;
        vals.push_back((float)tok[nm]);
    }
    std::fclose(fp);
    if (!ok) return false;
    if (nm == 0 || vals.empty()) {
        std::fprintf(stderr, "ERROR: '%s' holds no nonzeros\n", path.c_str());
        return false;
    }
    if (!T.try_allocate_elements(vals.size(), nm)) return false;
    std::memcpy(T.vals, vals.data(), vals.size() * sizeof(float));
    for (int n = 0; n < nm; ++n) {
        std::memcpy(T.idx_buf[n], idx[n].data(), vals.size() * sizeof(idx_t));
        T.mode_size[n] = sizes[n];
    }
    return true;
}

// ---------------------------------------------------------------------------
//  .dnb cache
// ---------------------------------------------------------------------------
bool save_dnb_tensor(const std::string& path, const Tensor& T) {
    if (T.nnz == 0 || T.num_modes <= 0 || T.num_modes > DYN_MAX_MODES) return false;
    const std::string tmp = path + ".tmp";
    std::FILE* fp = std::fopen(tmp.c_str(), "wb");
    if (!fp) return false;
    const uint32_t version = 1, vb = 4, ib = 4, nm = (uint32_t)T.num_modes;
    const uint64_t nnz = T.nnz;
    uint32_t sizes[DYN_MAX_MODES] = {0};
    for (int n = 0; n < T.num_modes; ++n) sizes[n] = T.mode_size[n];
    const uint32_t reserved[7] = {0};
    bool ok = std::fwrite(kDnbMagic, 4, 1, fp) == 1 && std::fwrite(&version, 4, 1, fp) == 1 &&
              std::fwrite(&vb, 4, 1, fp) == 1 && std::fwrite(&ib, 4, 1, fp) == 1 &&
              std::fwrite(&nm, 4, 1, fp) == 1 && std::fwrite(&nnz, 8, 1, fp) == 1 &&
              std::fwrite(sizes, sizeof(sizes), 1, fp) == 1 &&
              std::fwrite(reserved, sizeof(reserved), 1, fp) == 1 &&
              std::fwrite(T.vals, 4, nnz, fp) == nnz;
    for (int n = 0; ok && n < T.num_modes; ++n) ok = std::fwrite(T.idx_buf[n], 4, nnz, fp) == nnz;
    std::fclose(fp);
    if (!ok) {
        std::remove(tmp.c_str());
        return false;
    }
    std::remove(path.c_str());
    if (std::rename(tmp.c_str(), path.c_str()) != 0) {
        std::remove(tmp.c_str());
        return false;
    }
    return true;
}

static bool load_dnb(const std::string& path, Tensor& T) {
    std::FILE* fp = std::fopen(path.c_str(), "rb");
    if (!fp) return false;
    char magic[4];
    uint32_t version = 0, vb = 0, ib = 0, nm = 0, sizes[DYN_MAX_MODES], reserved[7];
    uint64_t nnz = 0;
    bool ok = std::fread(magic, 4, 1, fp) == 1 && std::fread(&version, 4, 1, fp) == 1 &&
              std::fread(&vb, 4, 1, fp) == 1 && std::fread(&ib, 4, 1, fp) == 1 &&
              std::fread(&nm, 4, 1, fp) == 1 && std::fread(&nnz, 8, 1, fp) == 1 &&
              std::fread(sizes, sizeof(sizes), 1, fp) == 1 &&
              std::fread(reserved, sizeof(reserved), 1, fp) == 1;
    if (!ok || std::memcmp(magic, kDnbMagic, 4) != 0 || version != 1 || vb != 4 || ib != 4 ||
        nm == 0 || nm > (uint32_t)DYN_MAX_MODES || nnz == 0) {
        std::fclose(fp);
        return false;
    }
    if (!T.try_allocate_elements(nnz, (int)nm)) {
        std::fclose(fp);
        return false;
    }
    for (uint32_t n = 0; n < nm; ++n) T.mode_size[n] = sizes[n];
    ok = std::fread(T.vals, 4, nnz, fp) == nnz;
    for (uint32_t n = 0; ok && n < nm; ++n) ok = std::fread(T.idx_buf[n], 4, nnz, fp) == nnz;
    std::fclose(fp);
    if (!ok) T.free_elements();
    return ok;
}

static bool dnb_fresh(const std::string& tns, const std::string& dnb) {
    struct stat b{};
    if (::stat(dnb.c_str(), &b) != 0 || (uint64_t)b.st_size < kDnbHeader) return false;
    struct stat a{};
    if (::stat(tns.c_str(), &a) != 0) return true;   // no source: trust the cache
    return b.st_mtime >= a.st_mtime;
}

bool ooc_load_header(const std::string& path, Tensor& T) {
    std::FILE* fp = std::fopen(path.c_str(), "rb");
    if (!fp) return false;
    char magic[4];
    uint32_t version = 0, vb = 0, ib = 0, nm = 0, sizes[DYN_MAX_MODES], reserved[7];
    uint64_t nnz = 0;
    bool ok = std::fread(magic, 4, 1, fp) == 1 && std::fread(&version, 4, 1, fp) == 1 &&
              std::fread(&vb, 4, 1, fp) == 1 && std::fread(&ib, 4, 1, fp) == 1 &&
              std::fread(&nm, 4, 1, fp) == 1 && std::fread(&nnz, 8, 1, fp) == 1 &&
              std::fread(sizes, sizeof(sizes), 1, fp) == 1 &&
              std::fread(reserved, sizeof(reserved), 1, fp) == 1;
    ok = ok && std::memcmp(magic, kDnbMagic, 4) == 0 && version == 1 && vb == 4 && ib == 4 &&
         nm >= 1 && nm <= (uint32_t)DYN_MAX_MODES && nnz > 0;
    long long size = -1;
    if (ok && std::fseek(fp, 0, SEEK_END) == 0) size = std::ftell(fp);
    std::fclose(fp);
    if (!ok || size < 0 || (uint64_t)size != kDnbHeader + nnz * 4ull * (1 + nm)) return false;
    T.free_elements();
    T.num_modes = (int)nm;
    T.nnz = nnz;
    for (uint32_t n = 0; n < nm; ++n) T.mode_size[n] = sizes[n];
    return true;
}

uint64_t ooc_default_chunk_nnz(const Tensor& T) {
    if (T.num_modes <= 0 || T.nnz == 0) return 0;
    uint64_t chunk = 1ull << 27;
    if (const char* env = std::getenv("DYN_OOC_CHUNK_BYTES"); env && env[0]) {
        const long long v = std::atoll(env);
        if (v > 0) chunk = std::max<uint64_t>((uint64_t)v / (4 + 4 * (uint64_t)T.num_modes), 1);
    }
    return std::min<uint64_t>(chunk, T.nnz);
}

bool load_coo_tensor(const std::string& path, Tensor& T, bool cache_on_success) {
    const std::string dnb = path + ".dnb";
    if (cache_on_success && dnb_fresh(path, dnb)) {
        if (load_dnb(dnb, T)) {
            std::printf("[io] loaded %llu nonzeros, %d modes, shape =", (unsigned long long)T.nnz,
                        T.num_modes);
            for (int n = 0; n < T.num_modes; ++n) std::printf(" %u", (unsigned)T.mode_size[n]);
            std::printf("  (from %s)\n", dnb.c_str());
            return true;
        }
        std::fprintf(stderr, "[io] WARN: stale or invalid binary cache '%s', reparsing text\n",
                     dnb.c_str());
    }
    if (!load_tns_text(path, T)) return false;
    std::printf("[io] loaded %llu nonzeros, %d modes, shape =", (unsigned long long)T.nnz,
                T.num_modes);
    for (int n = 0; n < T.num_modes; ++n) std::printf(" %u", (unsigned)T.mode_size[n]);
    std::printf("  (text)\n");
    if (cache_on_success && save_dnb_tensor(dnb, T))
        std::printf("[io] wrote binary cache '%s'\n", dnb.c_str());
    return true;
}

// ---------------------------------------------------------------------------
//  factor dumps
// ---------------------------------------------------------------------------
bool dump_factor_matrix(const std::string& path, const value_t* data, idx_t rows, int rank,
                        int rank_padded) {
    std::FILE* fp = std::fopen(path.c_str(), "w");
    if (!fp) return false;
    for (idx_t i = 0; i < rows; ++i) {
        const value_t* row = data + (size_t)i * rank_padded;
        for (int r = 0; r < rank; ++r) {
            std::fprintf(fp, "%.7g", (double)row[r]);
            if (r + 1 < rank) std::fputc(' ', fp);
        }
        std::fputc('\n', fp);
    }
    std::fclose(fp);
    return true;
}

bool dump_factor_matrix_bin(const std::string& path, const value_t* data, idx_t rows, int rank,
                            int rank_padded) {
    std::FILE* fp = std::fopen(path.c_str(), "wb");
    if (!fp) return false;
    const char magic[4] = {'D', 'Y', 'F', 'M'};
    const uint32_t version = 1, vb = 4, r = (uint32_t)rank;
    const uint64_t nr = rows;
    bool ok = std::fwrite(magic, 4, 1, fp) == 1 && std::fwrite(&version, 4, 1, fp) == 1 &&
              std::fwrite(&vb, 4, 1, fp) == 1 && std::fwrite(&r, 4, 1, fp) == 1 &&
              std::fwrite(&nr, 8, 1, fp) == 1;
    for (idx_t i = 0; ok && i < rows; ++i)
        ok = std::fwrite(data + (size_t)i * rank_padded, 4, rank, fp) == (size_t)rank;
    std::fclose(fp);
    if (!ok) std::remove(path.c_str());
    return ok;
}

bool load_factor_matrix_bin(const std::string& path, value_t* data, idx_t rows, int rank,
                            int rank_padded) {
    std::FILE* fp = std::fopen(path.c_str(), "rb");
    if (!fp) return false;
    char magic[4];
    uint32_t version = 0, vb = 0, r = 0;
    uint64_t nr = 0;
    bool ok = std::fread(magic, 4, 1, fp) == 1 && std::fread(&version, 4, 1, fp) == 1 &&
              std::fread(&vb, 4, 1, fp) == 1 && std::fread(&r, 4, 1, fp) == 1 &&
              std::fread(&nr, 8, 1, fp) == 1;
    ok = ok && std::memcmp(magic, "DYFM", 4) == 0 && version == 1 && vb == 4 &&
         r == (uint32_t)rank && nr == (uint64_t)rows;
    for (idx_t i = 0; ok && i < rows; ++i) {
        value_t* row = data + (size_t)i * rank_padded;
        ok = std::fread(row, 4, rank, fp) == (size_t)rank;
        for (int c = rank; c < rank_padded; ++c) row[c] = 0.0f;
    }
    std::fclose(fp);
    return ok;
}

}  // namespace dynasor

// ---------------------------------------------------------------------------
//  C-ABI loader hooks (ctypes / other-language callers)
// ---------------------------------------------------------------------------
struct dyn_host_tensor {
    dynasor::Tensor T;
};

extern "C" int dyn_host_tensor_load(const char* path, int cache_on_success,
                                    dyn_host_tensor** out) {
    if (!path || !out) return 1;
    auto* h = new dyn_host_tensor();
    if (!dynasor::load_coo_tensor(path, h->T, cache_on_success != 0)) {
        delete h;
        *out = nullptr;
        return 2;
    }
    *out = h;
    return 0;
}

extern "C" int dyn_host_tensor_info(const dyn_host_tensor* h, int* num_modes, uint64_t* nnz,
                                    uint32_t* dims) {
    if (!h) return 1;
    if (num_modes) *num_modes = h->T.num_modes;
    if (nnz) *nnz = h->T.nnz;
    if (dims)
        for (int n = 0; n < h->T.num_modes; ++n) dims[n] = h->T.mode_size[n];
    return 0;
}

extern "C" int dyn_host_tensor_copy(const dyn_host_tensor* h, float* vals,
                                    uint32_t* const* idx) {
    if (!h) return 1;
    if (vals) std::memcpy(vals, h->T.vals, h->T.nnz * sizeof(float));
    if (idx)
        for (int n = 0; n < h->T.num_modes; ++n)
            if (idx[n]) std::memcpy(idx[n], h->T.idx_buf[n], h->T.nnz * sizeof(uint32_t));
    return 0;
}

extern "C" int dyn_host_tensor_save_dnb(const dyn_host_tensor* h, const char* path) {
    if (!h || !path) return 1;
    return dynasor::save_dnb_tensor(path, h->T) ? 0 : 2;
}

extern "C" void dyn_host_tensor_free(dyn_host_tensor* h) { delete h; }
