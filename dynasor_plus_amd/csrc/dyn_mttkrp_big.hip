// The segment-kernel launches of tensors whose value / index columns reach
// 4 GiB (2^30 nonzeros or more on one device): the BIG instantiations of the
// kernel in dyn_seg.hpp, compiled apart from the ordinary ones.
#include "dyn_seg.hpp"

namespace dyn {

template int dispatch_rp<float, 0, true>(dyn_gpu_ctx*, int, int, SegArgs&, const uint32_t*);
template int dispatch_rp<bf16, 0, true>(dyn_gpu_ctx*, int, int, SegArgs&, const uint32_t*);
template int dispatch_rp<float, 1, true>(dyn_gpu_ctx*, int, int, SegArgs&, const uint32_t*);

}  // namespace dyn
