// Processed-state files for runtime-only runs (dynasor_dnp.h), on top of the
// device C-ABI (dyn_gpu_save_processed / _load_processed / _processed_info).
#include "dynasor_dnp.h"
#include "dynasor_gpu.h"

#include <cstdio>
#include <cstring>

namespace dynasor {

namespace {

// header (bytes) and payload order of the "DGP1" file: see dyn_persist.hip
constexpr long kHeaderBytes = 4 + 5 * 4 + 8 + 8 * 4 + 8 * 4 + 8 * 4;

}  // namespace

bool save_dnp_tensor(const std::string& path, const Tensor& T, int num_threads, int device) {
    (void)num_threads;
    dyn_gpu_ctx* h = nullptr;
    int rc = dyn_gpu_create(device, &h);
    if (rc == DYN_OK)
        rc = dyn_gpu_upload_coo(h, T.num_modes, T.nnz, T.mode_size, T.vals, T.idx_buf);
    if (rc == DYN_OK) rc = dyn_gpu_remap_all(h);
    if (rc == DYN_OK) rc = dyn_gpu_save_processed(h, path.c_str());
    if (rc != DYN_OK)
        std::fprintf(stderr, "[dnp] cannot save '%s' (status %d): %s\n", path.c_str(), rc,
                     h ? dyn_gpu_last_error(h) : "no device context");
    if (h) dyn_gpu_destroy(h);
    return rc == DYN_OK;
}

bool load_dnp_header(const std::string& path, Tensor& T) {
    int N = 0;
    uint64_t nnz = 0;
    uint32_t dims[DYN_MAX_MODES] = {0}, mask = 0;
    if (dyn_gpu_processed_info(path.c_str(), &N, &nnz, dims, &mask) != DYN_OK) {
        std::fprintf(stderr, "[dnp] '%s' is not a processed-state file\n", path.c_str());
        return false;
    }
    T.free_elements();
    T.num_modes = N;
    T.nnz = nnz;
    for (int n = 0; n < DYN_MAX_MODES; ++n) T.mode_size[n] = n < N ? dims[n] : 0;
    T.layout = Layout::DeviceCopies;
    return true;
}

int load_dnp_num_threads(const std::string& path) {
    return dyn_gpu_processed_info(path.c_str(), nullptr, nullptr, nullptr, nullptr) == DYN_OK
               ? 0
               : -1;
}

Layout load_dnp_layout(const std::string& path) {
    if (dyn_gpu_processed_info(path.c_str(), nullptr, nullptr, nullptr, nullptr) == DYN_OK)
        return Layout::DeviceCopies;
    std::fprintf(stderr, "[dnp] '%s' is not a processed-state file\n", path.c_str());
    return Layout::PingPong;
}

bool load_dnp_tensor(const std::string& path, Tensor& T, int num_threads) {
    (void)num_threads;
    Tensor hdr;
    if (!load_dnp_header(path, hdr)) return false;
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    bool ok = std::fseek(f, kHeaderBytes, SEEK_SET) == 0 &&
              T.try_allocate_elements(hdr.nnz, hdr.num_modes);
    if (ok) {
        for (int n = 0; n < hdr.num_modes; ++n) T.mode_size[n] = hdr.mode_size[n];
        ok = std::fread(T.vals, 4, hdr.nnz, f) == hdr.nnz;
        for (int n = 0; ok && n < hdr.num_modes; ++n)
            ok = std::fread(T.idx_buf[n], 4, hdr.nnz, f) == hdr.nnz;
    }
    std::fclose(f);
    if (!ok) {
        T.free_elements();
        std::fprintf(stderr, "[dnp] cannot read the tensor of '%s'\n", path.c_str());
        return false;
    }
    T.layout = Layout::DeviceCopies;
    return true;
}

}  // namespace dynasor
