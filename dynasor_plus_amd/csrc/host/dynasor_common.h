// ============================================================================
//  dynasor_common.h -- C++ host types of the MI355X CP-ALS engine.
//
//  Source-compatible subset of the reference's include/dynasor_common.h for
//  the `--decompose cpals` path: the SoA COO Tensor, FactorMatrices with the
//  same padded-rank rule, the ComputePlan entry point and the .tns / .dnb /
//  DYFM I/O functions.  The CPU-side FLYCOO state of the reference (shards,
//  schedules, NCopy slabs, Morton / OOC state) has no counterpart: the device
//  builds its own per-mode copies (dyn_gpu_remap).
// ============================================================================
#ifndef DYNASOR_GPU_COMMON_H
#define DYNASOR_GPU_COMMON_H

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace dynasor {

#ifndef DYN_MAX_MODES
#define DYN_MAX_MODES 8
#endif

using idx_t = uint32_t;    // include/dynasor_common.h:28
using value_t = float;     // include/dynasor_common.h:30

// Storage layouts (include/dynasor_common.h:113-165).  The device engine
// holds the tensor as one sorted copy per mode in device memory
// (DeviceCopies), as one linearised Z-order copy (Morton: the reference's
// ALTO-style layout, single-pass all-modes, Jacobi only) or streams it from
// host memory (OutOfCore); the CPU-only layouts keep their values so callers
// that read, compare or assign a layout compile and behave unchanged.
enum class Layout : uint8_t {
    PingPong = 0,
    NCopy = 1,
    InPlace = 2,
    OutOfCore = 3,
    Morton = 4,         // one linearised (Z-order) copy in HBM (dyn_gpu_set_layout)
    DeviceCopies = 5,   // per-mode sorted copies in HBM (dyn_gpu_remap)
};

// Storage decision (include/dynasor_common.h:202-224): the same fields, with
// device memory in place of host RAM.  avail/total are the device's free and
// total bytes, ncopy_bytes the device tensor footprint (the COO upload plus
// one sorted copy per mode), overhead_bytes the factors, their BF16 shadow,
// Yhat and the sort scratch, and ncopy_need their sum.
struct StoragePlan {
    Layout      choice = Layout::DeviceCopies;
    size_t      avail_bytes = 0;
    size_t      total_bytes = 0;
    size_t      budget_bytes = 0;
    size_t      pingpong_bytes = 0;
    size_t      ncopy_bytes = 0;
    size_t      inplace_bytes = 0;
    size_t      overhead_bytes = 0;
    size_t      pingpong_need = 0;
    size_t      ncopy_need = 0;
    size_t      inplace_need = 0;
    size_t      ooc_need = 0;
    const char* reason = "";
    const char* override_source = "";
};

// Per-mode kernel implementation (include/dynasor_common.h:261-265).  The
// device spMTTKRP walks a row pointer over a sorted copy, the counterpart of
// FiberCSR, for every mode.
enum class KernelImpl : uint8_t {
    Element = 0,
    Fiber = 1,
    FiberCSR = 2,
};

// SoA COO tensor (include/dynasor_common.h:322-600, the fields cpals() and the
// loaders use).  Owns vals / idx_buf; non-copyable like the reference.
struct Tensor {
    int      num_modes = 0;
    uint64_t nnz = 0;
    idx_t    mode_size[DYN_MAX_MODES] = {0};
    value_t* vals = nullptr;
    idx_t*   idx_buf[DYN_MAX_MODES] = {nullptr};
    Layout   layout = Layout::DeviceCopies;   // include/dynasor_common.h:390

    void allocate_elements(uint64_t nnz, int nmodes);     // exits on host OOM (reference)
    bool try_allocate_elements(uint64_t nnz, int nmodes); // false on host OOM
    void free_elements();

    Tensor() = default;
    ~Tensor();
    Tensor(const Tensor&) = delete;
    Tensor& operator=(const Tensor&) = delete;
};

// Dense factors Y[n], Yhat[n]: I_n x rank_padded, row-major, padding zero
// (include/dynasor_common.h:604-615).  rank_padded = max(roundup(R, 8),
// roundup(R, 16)) as the reference's AVX2/AVX-512 builds compute it
// (src/dynasor_io.cpp:157-160).
struct FactorMatrices {
    int                   num_modes = 0;
    int                   rank = 0;
    int                   rank_padded = 0;
    std::vector<value_t*> Y;
    std::vector<value_t*> Yhat;

    void allocate(int nmodes, const idx_t* mode_sizes, int R);
    void free_all();
    void zero_output();
    ~FactorMatrices() { free_all(); }
};

// ---------------------------------------------------------------------------
//  ComputePlan (include/dynasor_common.h:271-303, src/dynasor_preprocess.cpp:
//  1179-1363).  Same entry point and kernel-selection surface; the storage
//  fields describe device memory instead of the CPU layouts.
// ---------------------------------------------------------------------------
struct ComputePlan {
    StoragePlan storage;                  // device storage decision
    int         kernel_enum = 0;          // CpKernel value (0 = GS, 1 = Jacobi)
    const char* kernel_reason = "";
    const char* kernel_source = "auto";   // "cli-kernel" | "env-kernel" | "auto"
    KernelImpl  per_mode[DYN_MAX_MODES] = {KernelImpl::Element};
    bool        pf_far = false;           // no host prefetch tier on the device path
    bool        use_jit = false;          // kernels are compiled ahead of time for gfx950
    size_t      ofibs_bytes = 0;          // all-modes pass: rank-ordered Yhat accumulators
    int         precision = 0;            // 0 = FP32 gathers, 1 = BF16 gathers
    const char* precision_source = "auto";
    int         num_modes = 0;
    int         rank_padded = 0;          // host padding (reference rule)
    int         device_rank_padded = 0;   // device row stride (power of two >= 16)
    int         num_threads = 0;
    size_t      factor_bytes = 0;         // host Y + Yhat
    size_t      device_tensor_bytes = 0;  // COO + N per-mode copies
    size_t      device_factor_bytes = 0;  // Y, BF16 shadow, Yhat on the device
    size_t      device_total_bytes = 0;
    size_t      device_free_bytes = 0;    // at planning time (0 if unknown)
    size_t      device_linear_bytes = 0;  // COO + one Z-order copy + accumulators + factors
    bool        fits = true;              // per-mode copies fit in free device memory
    const char* layout = "device-remap";
};

ComputePlan dyn_build_compute_plan(const Tensor& T, int rank_padded, int num_threads,
                                   const std::string& layout_cli, const std::string& ooc_cli,
                                   const std::string& kernel_cli);
// dyn_build_compute_plan with the device memory given instead of queried
// (device 0, capped by DYN_GPU_MEM_CAP_MB): `have_device` false plans as if
// no device were visible (auto then keeps Gauss-Seidel and `fits` stays true).
ComputePlan dyn_build_compute_plan_with_memory(const Tensor& T, int rank_padded, int num_threads,
                                               const std::string& layout_cli,
                                               const std::string& ooc_cli,
                                               const std::string& kernel_cli, bool have_device,
                                               uint64_t free_bytes, uint64_t total_bytes);
void dyn_print_compute_plan(const ComputePlan& p, const char* prefix = "[plan]");

// ---------------------------------------------------------------------------
//  I/O (src/dynasor_io.cpp)
// ---------------------------------------------------------------------------
// <path>.dnb when fresh (mtime >= the text's, or the text is absent), else
// parse the FROSTT text and, when cache_on_success, write <path>.dnb.
bool load_coo_tensor(const std::string& path, Tensor& T, bool cache_on_success = true);
bool save_dnb_tensor(const std::string& path, const Tensor& T);
bool dump_factor_matrix(const std::string& path, const value_t* data, idx_t rows, int rank,
                        int rank_padded);
bool dump_factor_matrix_bin(const std::string& path, const value_t* data, idx_t rows, int rank,
                            int rank_padded);
// Read a DYFM file into rows x rank_padded (padding zeroed).  Returns false on
// a bad header or a shape mismatch.
bool load_factor_matrix_bin(const std::string& path, value_t* data, idx_t rows, int rank,
                            int rank_padded);

}  // namespace dynasor

#endif  // DYNASOR_GPU_COMMON_H
