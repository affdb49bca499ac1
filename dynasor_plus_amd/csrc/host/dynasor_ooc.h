// ============================================================================
//  dynasor_ooc.h -- the out-of-core helpers a caller uses before choosing to
//  stream (reference: include/dynasor_ooc.h:63-65, src/dynasor_ooc.cpp:196-216
//  and src/main.cpp:460-513): read only a .dnb file's header, and the default
//  streaming chunk.  The streaming itself is dyn_gpu_attach_dnb /
//  dyn_gpu_attach_host_coo (include/dynasor_gpu.h), selected through
//  CpOptions::stream_path / stream_from_host.
// ============================================================================
#ifndef DYNASOR_GPU_OOC_H
#define DYNASOR_GPU_OOC_H

#include "dynasor_common.h"

#include <cstdint>
#include <string>

namespace dynasor {

// num_modes, nnz and mode sizes from the .dnb header at `path` (no elements
// are read or allocated); false if the file is missing or its header or
// size is invalid.
bool ooc_load_header(const std::string& path, Tensor& T);

// Nonzeros per streamed chunk: DYN_OOC_CHUNK_BYTES / (4 + 4 N) when that is
// set (as in the reference), else 2^27 -- device chunks are large so that a
// chunk's sort and spMTTKRP launches fill the GPU; at most nnz.
uint64_t ooc_default_chunk_nnz(const Tensor& T);

}  // namespace dynasor

#endif  // DYNASOR_GPU_OOC_H
