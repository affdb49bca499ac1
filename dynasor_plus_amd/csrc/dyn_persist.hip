// Processed-state file of a device context (no sort on reload).
//
// The reference saves its post-preprocessing state (.dnp: FLYCOO schedule,
// NCopy slabs; include/dynasor_dnp.h:33-101, src/dynasor_dnp.cpp:638-805) so
// that `--runtime` runs skip the preprocessing.  The device counterpart of
// that preprocessing is the per-mode radix sort; what it produces that cannot
// be recomputed cheaply is each copy's permutation.  The file therefore holds
// the COO tensor in input order and one permutation per built copy; a reload
// uploads the tensor, checks every permutation (a permutation of [0, nnz)
// that sorts the mode's indices), and gathers the copies through it -- the
// frequency relabel, the row pointers and the work tables are recomputed from
// those exactly as after a sort, so a reloaded copy equals the saved one.
//
// Layout (little-endian), "DGP1":
//   char     magic[4]      "DGP1"
//   uint32_t version       1
//   uint32_t value_bytes   4
//   uint32_t index_bytes   4
//   uint32_t num_modes
//   uint32_t copies_mask   bit n: the mode-n permutation is stored
//   uint64_t nnz
//   uint32_t dims[8]
//   int32_t  secondary[8]  secondary sort mode of each stored copy (-1: none)
//   uint32_t reserved[8]
//   float    vals[nnz]                       input order
//   uint32_t idx[num_modes][nnz]             input order, 0-based
//   uint32_t perm[n][nnz] for each n in copies_mask, ascending n
#include "dyn_internal.hpp"

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace dyn {

namespace {

constexpr char kMagic[4] = {'D', 'G', 'P', '1'};
constexpr uint32_t kVersion = 1;
constexpr size_t kHeaderBytes = 4 + 5 * 4 + 8 + 8 * 4 + 8 * 4 + 8 * 4;

struct Header {
    uint32_t num_modes = 0;
    uint32_t mask = 0;
    uint64_t nnz = 0;
    uint32_t dims[kMaxModes] = {0};
    int32_t  secondary[kMaxModes] = {0};
};

struct File {
    std::FILE* f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

bool put(std::FILE* f, const void* p, size_t bytes) {
    return bytes == 0 || std::fwrite(p, 1, bytes, f) == bytes;
}
bool get(std::FILE* f, void* p, size_t bytes) {
    return bytes == 0 || std::fread(p, 1, bytes, f) == bytes;
}

// header fields in order; `why` names the first problem
bool read_header(std::FILE* f, Header* h, const char** why) {
    char magic[4];
    uint32_t version = 0, vb = 0, ib = 0, reserved[8];
    if (!get(f, magic, 4) || !get(f, &version, 4) || !get(f, &vb, 4) || !get(f, &ib, 4) ||
        !get(f, &h->num_modes, 4) || !get(f, &h->mask, 4) || !get(f, &h->nnz, 8) ||
        !get(f, h->dims, sizeof(h->dims)) || !get(f, h->secondary, sizeof(h->secondary)) ||
        !get(f, reserved, sizeof(reserved))) {
        *why = "truncated header";
        return false;
    }
    if (std::memcmp(magic, kMagic, 4) != 0) {
        *why = "not a processed-state file (bad magic)";
        return false;
    }
    if (version != kVersion || vb != 4 || ib != 4) {
        *why = "unsupported version or value/index width";
        return false;
    }
    const uint32_t N = h->num_modes;
    if (N < 2 || N > (uint32_t)kMaxModes) {
        *why = "mode count out of range";
        return false;
    }
    if (h->nnz >= kMaxNnz) {
        *why = "too many nonzeros for one device";
        return false;
    }
    if (h->mask >> N) {
        *why = "copy mask names a mode beyond the mode count";
        return false;
    }
    for (uint32_t n = 0; n < N; ++n) {
        if (h->dims[n] == 0) {
            *why = "a mode of size 0";
            return false;
        }
        const int32_t s = h->secondary[n];
        if (s < -1 || s >= (int32_t)N || s == (int32_t)n) {
            *why = "bad secondary sort mode";
            return false;
        }
    }
    return true;
}

uint64_t payload_bytes(const Header& h) {
    uint64_t copies = 0;
    for (uint32_t n = 0; n < h.num_modes; ++n) copies += (h.mask >> n) & 1u;
    return h.nnz * 4ull * (1ull + h.num_modes + copies);
}

template <typename T>
int d2h(dyn_gpu_ctx* c, std::vector<T>& v, const T* d, size_t n) {
    v.resize(n);
    if (n) {
        DYN_HIP(c, hipMemcpyAsync(v.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, c->stream));
        DYN_HIP(c, hipStreamSynchronize(c->stream));
    }
    return DYN_OK;
}

}  // namespace
}  // namespace dyn

using namespace dyn;

extern "C" int dyn_gpu_save_processed(dyn_gpu_ctx* c, const char* path) {
    if (!c || !path) return DYN_ERR_INVALID;
    if (c->N == 0) return set_err(c, DYN_ERR_STATE, "no tensor uploaded");
    DYN_CHECK(reject_streamed(c, "dyn_gpu_save_processed"));
    DYN_HIP(c, hipSetDevice(c->device));
    DYN_HIP(c, hipStreamSynchronize(c->stream));
    Header h;
    h.num_modes = (uint32_t)c->N;
    h.nnz = c->nnz;
    for (int n = 0; n < c->N; ++n) {
        h.dims[n] = c->dims[n];
        h.secondary[n] = -1;
        if (c->copy[n].built) {
            h.mask |= 1u << n;
            h.secondary[n] = c->copy[n].secondary;
        }
    }
    const std::string tmp = std::string(path) + ".tmp";
    {
        File out;
        out.f = std::fopen(tmp.c_str(), "wb");
        if (!out.f) return set_err(c, DYN_ERR_INVALID, "cannot write '%s'", tmp.c_str());
        const uint32_t version = kVersion, four = 4, reserved[8] = {0};
        bool ok = put(out.f, kMagic, 4) && put(out.f, &version, 4) && put(out.f, &four, 4) &&
                  put(out.f, &four, 4) && put(out.f, &h.num_modes, 4) && put(out.f, &h.mask, 4) &&
                  put(out.f, &h.nnz, 8) && put(out.f, h.dims, sizeof(h.dims)) &&
                  put(out.f, h.secondary, sizeof(h.secondary)) &&
                  put(out.f, reserved, sizeof(reserved));
        std::vector<float> vf;
        std::vector<uint32_t> vu;
        if (ok) {
            DYN_CHECK(d2h(c, vf, c->vals, c->nnz));
            ok = put(out.f, vf.data(), vf.size() * 4);
            std::vector<float>().swap(vf);
        }
        for (int n = 0; ok && n < c->N; ++n) {
            DYN_CHECK(d2h(c, vu, (const uint32_t*)c->idx[n], c->nnz));
            ok = put(out.f, vu.data(), vu.size() * 4);
        }
        for (int n = 0; ok && n < c->N; ++n) {
            if (!((h.mask >> n) & 1u)) continue;
            DYN_CHECK(d2h(c, vu, (const uint32_t*)c->copy[n].perm, c->nnz));
            ok = put(out.f, vu.data(), vu.size() * 4);
        }
        ok = ok && std::fflush(out.f) == 0;
        if (!ok) {
            std::fclose(out.f);
            out.f = nullptr;
            std::remove(tmp.c_str());
            return set_err(c, DYN_ERR_INVALID, "write to '%s' failed", tmp.c_str());
        }
    }
    if (std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return set_err(c, DYN_ERR_INVALID, "cannot rename '%s' to '%s'", tmp.c_str(), path);
    }
    return DYN_OK;
}

extern "C" int dyn_gpu_processed_info(const char* path, int* num_modes, uint64_t* nnz,
                                      uint32_t* dims, uint32_t* copies_mask) {
    if (!path) return DYN_ERR_INVALID;
    File in;
    in.f = std::fopen(path, "rb");
    if (!in.f) return DYN_ERR_INVALID;
    Header h;
    const char* why = nullptr;
    if (!read_header(in.f, &h, &why)) return DYN_ERR_INVALID;
    if (std::fseek(in.f, 0, SEEK_END) != 0) return DYN_ERR_INVALID;
    const long long size = std::ftell(in.f);
    if (size < 0 || (uint64_t)size != kHeaderBytes + payload_bytes(h)) return DYN_ERR_INVALID;
    if (num_modes) *num_modes = (int)h.num_modes;
    if (nnz) *nnz = h.nnz;
    if (dims)
        for (uint32_t n = 0; n < h.num_modes; ++n) dims[n] = h.dims[n];
    if (copies_mask) *copies_mask = h.mask;
    return DYN_OK;
}

extern "C" int dyn_gpu_load_processed(dyn_gpu_ctx* c, const char* path) {
    if (!c || !path) return DYN_ERR_INVALID;
    File in;
    in.f = std::fopen(path, "rb");
    if (!in.f) return set_err(c, DYN_ERR_INVALID, "cannot open '%s'", path);
    Header h;
    const char* why = nullptr;
    if (!read_header(in.f, &h, &why)) return set_err(c, DYN_ERR_INVALID, "'%s': %s", path, why);
    if (std::fseek(in.f, 0, SEEK_END) != 0)
        return set_err(c, DYN_ERR_INVALID, "'%s': cannot seek", path);
    const long long size = std::ftell(in.f);
    if (size < 0 || (uint64_t)size != kHeaderBytes + payload_bytes(h))
        return set_err(c, DYN_ERR_INVALID, "'%s': size %lld does not match its header (%llu)",
                       path, size, (unsigned long long)(kHeaderBytes + payload_bytes(h)));
    if (std::fseek(in.f, (long)kHeaderBytes, SEEK_SET) != 0)
        return set_err(c, DYN_ERR_INVALID, "'%s': cannot seek", path);
    const size_t n = (size_t)h.nnz;
    const int N = (int)h.num_modes;
    std::vector<float> vals(n);
    std::vector<std::vector<uint32_t>> idx(N, std::vector<uint32_t>(n));
    bool ok = get(in.f, vals.data(), n * 4);
    for (int m = 0; ok && m < N; ++m) ok = get(in.f, idx[m].data(), n * 4);
    if (!ok) return set_err(c, DYN_ERR_INVALID, "'%s': truncated tensor section", path);
    // the upload checks every index against its mode size
    const uint32_t* cols[kMaxModes];
    for (int m = 0; m < N; ++m) cols[m] = idx[m].data();
    DYN_CHECK(dyn_gpu_upload_coo(c, N, h.nnz, h.dims, vals.data(), cols));
    std::vector<float>().swap(vals);
    std::vector<std::vector<uint32_t>>().swap(idx);
    std::vector<uint32_t> perm(n);
    for (int m = 0; m < N; ++m) {
        if (!((h.mask >> m) & 1u)) continue;
        if (!get(in.f, perm.data(), n * 4))
            return set_err(c, DYN_ERR_INVALID, "'%s': truncated mode-%d permutation", path, m);
        const int rc = restore_copy(c, m, h.secondary[m], perm.data());
        if (rc != DYN_OK) {
            free_copy(c, m);
            return rc;
        }
    }
    return DYN_OK;
}
