// ============================================================================
//  dynasor_dnp.h -- processed-state files for `--runtime` runs, the drop-in
//  counterpart of the reference's include/dynasor_dnp.h:110-145.
//
//  The reference stores its CPU preprocessing (FLYCOO schedule, NCopy slabs)
//  in a .dnp file.  The device engine's preprocessing is one radix sort per
//  mode; its file ("DGP1", dyn_gpu_save_processed in include/dynasor_gpu.h)
//  holds the COO tensor and one permutation per mode copy, so a runtime-only
//  run gathers the copies without sorting.  The same function names and
//  meanings; the thread count has no role on the device (the schedule is
//  rebuilt from the row pointers), so it is accepted and not checked.
// ============================================================================
#ifndef DYNASOR_GPU_DNP_H
#define DYNASOR_GPU_DNP_H

#include "dynasor_common.h"

#include <string>

namespace dynasor {

// Preprocess T on device `device` (upload + every mode's sorted copy) and
// write the processed state to `path`.  False on a device or I/O error
// (message on stderr).
bool save_dnp_tensor(const std::string& path, const Tensor& T, int num_threads, int device = 0);

// Header only: T.num_modes, T.nnz, T.mode_size[], T.layout (DeviceCopies);
// no elements are loaded.  False if `path` is not a processed-state file.
bool load_dnp_header(const std::string& path, Tensor& T);

// The file does not depend on a thread count: 0 for a valid file, -1 on error.
int load_dnp_num_threads(const std::string& path);

// Layout::DeviceCopies for a valid file, Layout::PingPong (and a message) on
// error, as the reference's peek reports errors.
Layout load_dnp_layout(const std::string& path);

// T gets the file's COO tensor (input order) on the host, ready for cpals()
// with CpOptions::processed_path = path, which then skips the device sort.
bool load_dnp_tensor(const std::string& path, Tensor& T, int num_threads);

}  // namespace dynasor

#endif  // DYNASOR_GPU_DNP_H
