// Prints the device planner's decision for one tensor shape and memory
// budget, for tests/test_plan.py:
//   plan_cases <nnz> <rank> <kernel_cli> <have_device 0|1> <free_bytes> <total_bytes> <dims...>
// Output (tab-separated): "kernel=<0|1> reason=<...> source=<...> fits=<0|1> ofibs=<bytes>
//          need=<bytes> choice=<layout> linear=<bytes> layout=<name>"
// (environment PLAN_LAYOUT: the --layout value, default auto)
#include "dynasor_common.h"

#include <cstdio>
#include <cstdlib>

using namespace dynasor;

int main(int argc, char** argv) {
    if (argc < 9) return 1;
    Tensor T;
    T.nnz = std::strtoull(argv[1], nullptr, 10);
    const int rank = std::atoi(argv[2]);
    const std::string kernel = argv[3];
    const bool have = std::atoi(argv[4]) != 0;
    const uint64_t fr = std::strtoull(argv[5], nullptr, 10);
    const uint64_t tot = std::strtoull(argv[6], nullptr, 10);
    T.num_modes = argc - 7;
    for (int n = 0; n < T.num_modes; ++n) T.mode_size[n] = (idx_t)std::strtoul(argv[7 + n], nullptr, 10);
    const char* layout = std::getenv("PLAN_LAYOUT");
    ComputePlan p = dyn_build_compute_plan_with_memory(T, rank, 1, layout ? layout : "auto", "auto",
                                                       kernel, have, fr, tot);
    std::printf("kernel=%d\treason=%s\tsource=%s\tfits=%d\tofibs=%zu\tneed=%zu\tchoice=%d\t"
                "linear=%zu\tlayout=%s\n",
                p.kernel_enum, p.kernel_reason, p.kernel_source, (int)p.fits, p.ofibs_bytes,
                p.device_total_bytes, (int)p.storage.choice, p.device_linear_bytes, p.layout);
    return 0;
}
