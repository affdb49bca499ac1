// Host containers and file formats: Tensor / FactorMatrices storage, the
// FROSTT .tns parser, the .dnb binary cache and DYFM factor dumps.  Formats
// are byte-compatible with the reference (src/dynasor_io.cpp:544-1348).
#include "dynasor_common.h"
#include "dynasor_host.h"
#include "dynasor_ooc.h"

#include <fcntl.h>
#include <omp.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

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
//  FROSTT text, parallel (the reference's load_tns_mmap,
//  src/dynasor_io.cpp:960-1176): the file is mapped, cut into one slice of
//  whole lines per thread, every slice's data lines are counted, and each
//  thread then parses its slice straight into the tensor at its offset.
//  Indices must be plain decimal integers and every data line must carry
//  N + 1 tokens; a file that is not of that shape (an index written as 3.0,
//  a ragged line) returns false so that the caller falls back to the serial
//  parser above, which accepts it or reports the error.  Results are those
//  of the serial parser bit for bit: the same double parse of each value
//  cast to float, and the mode sizes are the larger of a SPLATT header's
//  and the largest index seen.
// ---------------------------------------------------------------------------
namespace {

struct Mapped {
    const char* p = nullptr;
    size_t n = 0;
    ~Mapped() {
        if (p) ::munmap((void*)p, n);
    }
};

inline bool blank(char c) { return c == ' ' || c == '\t' || c == '\r'; }

// first byte of the line after `pos` (or n)
inline size_t next_line(const char* b, size_t pos, size_t n) {
    const void* q = pos < n ? std::memchr(b + pos, '\n', n - pos) : nullptr;
    return q ? (size_t)((const char*)q - b) + 1 : n;
}

// Tokens of one line as doubles (the serial parser's reading), at most
// max_tok; returns the count, or -1 if a token does not parse.
int line_tokens(const char* s, const char* e, double* tok, int max_tok) {
    char local[512];
    const size_t len = std::min<size_t>((size_t)(e - s), sizeof(local) - 1);
    std::memcpy(local, s, len);
    local[len] = 0;
    int nt = 0;
    char* p = local;
    while (*p) {
        while (*p && std::isspace((unsigned char)*p)) ++p;
        if (!*p) break;
        if (nt >= max_tok) return -1;
        char* end = nullptr;
        const double v = std::strtod(p, &end);
        if (end == p) return -1;
        tok[nt++] = v;
        p = end;
    }
    return nt;
}

// Whether [s, e) is a data line: not blank, not a comment.
inline bool data_line(const char* s, const char* e) {
    while (s < e && (blank(*s) || *s == '\n')) ++s;
    return s < e && *s != '#';
}

}  // namespace

static bool load_tns_parallel(const std::string& path, Tensor& T, int* threads_used) {
    const int fd = ::open(path.c_str(), O_RDONLY);
    if (fd < 0) return false;
    struct stat st {};
    if (::fstat(fd, &st) != 0 || st.st_size <= 0) {
        ::close(fd);
        return false;
    }
    Mapped m;
    m.n = (size_t)st.st_size;
    void* map = ::mmap(nullptr, m.n, PROT_READ, MAP_PRIVATE, fd, 0);
    ::close(fd);
    if (map == MAP_FAILED) return false;
    m.p = (const char*)map;
    (void)::madvise(map, m.n, MADV_SEQUENTIAL);
    const char* b = m.p;
    const size_t n = m.n;

    // leading blank / comment lines, an optional SPLATT header ("N", then the
    // N mode sizes), and the first data line's token count
    size_t pos = 0;
    double tok[DYN_MAX_MODES + 2];
    int nm = 0;
    idx_t hdr[DYN_MAX_MODES] = {0};
    auto next_data = [&](size_t& p) {
        while (p < n) {
            const size_t e = next_line(b, p, n);
            if (data_line(b + p, b + e)) return true;
            p = e;
        }
        return false;
    };
    if (!next_data(pos)) return false;
    size_t e0 = next_line(b, pos, n);
    int nt = line_tokens(b + pos, b + e0, tok, DYN_MAX_MODES + 1);
    if (nt == 1 && tok[0] == (double)(int)tok[0] && tok[0] >= 2 && tok[0] <= DYN_MAX_MODES) {
        nm = (int)tok[0];
        pos = e0;
        if (!next_data(pos)) return false;
        e0 = next_line(b, pos, n);
        if (line_tokens(b + pos, b + e0, tok, DYN_MAX_MODES + 1) != nm) return false;
        for (int k = 0; k < nm; ++k) hdr[k] = (idx_t)tok[k];
        pos = e0;
        if (!next_data(pos)) return false;
        e0 = next_line(b, pos, n);
        nt = line_tokens(b + pos, b + e0, tok, DYN_MAX_MODES + 1);
        if (nt != nm + 1) return false;
    } else {
        if (nt < 2 || nt > DYN_MAX_MODES + 1) return false;
        nm = nt - 1;
    }
    const size_t start = pos;

    // one slice of whole lines per thread (at least 1 MiB each)
    int nth = std::max(1, omp_get_max_threads());
    nth = (int)std::min<size_t>((size_t)nth, std::max<size_t>(1, (n - start) >> 20));
    std::vector<size_t> cut(nth + 1, n);
    cut[0] = start;
    for (int t = 1; t < nth; ++t)
        cut[t] = std::max(cut[t - 1], next_line(b, start + (n - start) / nth * t - 1, n));
    std::vector<uint64_t> count(nth + 1, 0);
#pragma omp parallel for num_threads(nth) schedule(static, 1)
    for (int t = 0; t < nth; ++t) {
        uint64_t c = 0;
        for (size_t p = cut[t]; p < cut[t + 1];) {
            const size_t e = next_line(b, p, cut[t + 1]);
            c += data_line(b + p, b + e) ? 1 : 0;
            p = e;
        }
        count[t + 1] = c;
    }
    for (int t = 0; t < nth; ++t) count[t + 1] += count[t];
    const uint64_t total = count[nth];
    if (total == 0) return false;
    if (!T.try_allocate_elements(total, nm)) return false;

    std::vector<int> ok(nth, 1);
    std::vector<idx_t> mx((size_t)nth * DYN_MAX_MODES, 0);
#pragma omp parallel for num_threads(nth) schedule(static, 1)
    for (int t = 0; t < nth; ++t) {
        uint64_t k = count[t];
        idx_t* lmx = &mx[(size_t)t * DYN_MAX_MODES];
        for (size_t p = cut[t]; p < cut[t + 1] && ok[t];) {
            const size_t e = next_line(b, p, cut[t + 1]);
            const char* s = b + p;
            const char* le = b + e;
            p = e;
            if (!data_line(s, le)) continue;
            for (int w = 0; w < nm; ++w) {
                while (s < le && blank(*s)) ++s;
                uint64_t v = 0;
                const char* d = s;
                while (s < le && *s >= '0' && *s <= '9' && s - d < 12) v = v * 10 + (uint64_t)(*s++ - '0');
                // an index must be a whole decimal number in [1, 2^32]
                if (s == d || (s < le && !blank(*s) && *s != '\n') || v == 0 || v > 0x100000000ull) {
                    ok[t] = 0;
                    break;
                }
                const idx_t ix = (idx_t)(v - 1);
                T.idx_buf[w][k] = ix;
                if (ix + 1 > lmx[w]) lmx[w] = ix + 1;
            }
            if (!ok[t]) break;
            while (s < le && blank(*s)) ++s;
            // the value: the serial parser's strtod, on a terminated copy
            char local[128];
            size_t len = 0;
            while (s + len < le && !blank(s[len]) && s[len] != '\n' && len + 1 < sizeof(local)) {
                local[len] = s[len];
                ++len;
            }
            local[len] = 0;
            char* end = nullptr;
            const double val = std::strtod(local, &end);
            if (len == 0 || end != local + len) {
                ok[t] = 0;
                break;
            }
            // nothing but blanks may follow the value
            for (const char* q = s + len; q < le; ++q)
                if (!blank(*q) && *q != '\n') {
                    ok[t] = 0;
                    break;
                }
            if (!ok[t]) break;
            T.vals[k] = (value_t)val;
            ++k;
        }
        if (ok[t] && k != count[t + 1]) ok[t] = 0;
    }
    for (int t = 0; t < nth; ++t)
        if (!ok[t]) {
            T.free_elements();
            T.num_modes = 0;
            T.nnz = 0;
            return false;
        }
    for (int w = 0; w < nm; ++w) {
        idx_t s = hdr[w];
        for (int t = 0; t < nth; ++t) s = std::max(s, mx[(size_t)t * DYN_MAX_MODES + w]);
        T.mode_size[w] = s;
    }
    if (threads_used) *threads_used = nth;
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
    // the parallel parser first (src/dynasor_io.cpp:1215-1234), the serial
    // one when it declines the file; DYN_TNS_PARSER=serial skips the former
    int threads = 0;
    const char* env = std::getenv("DYN_TNS_PARSER");
    const bool parallel_ok = !(env && std::strcmp(env, "serial") == 0);
    bool parsed = parallel_ok && load_tns_parallel(path, T, &threads);
    if (!parsed) {
        threads = 0;
        if (!load_tns_text(path, T)) return false;
    }
    std::printf("[io] loaded %llu nonzeros, %d modes, shape =", (unsigned long long)T.nnz,
                T.num_modes);
    for (int n = 0; n < T.num_modes; ++n) std::printf(" %u", (unsigned)T.mode_size[n]);
    if (threads > 0)
        std::printf("  (text, parallel parser on %d thread%s)\n", threads, threads > 1 ? "s" : "");
    else
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
