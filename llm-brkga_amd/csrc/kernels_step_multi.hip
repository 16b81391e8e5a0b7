// kernels_step_multi.hip - instantiations of the multi-child step kernel
// (step_multi.hpp): CH = 2 or 4 children of BLOCK threads per workgroup.
#include "step_multi.hpp"

namespace brkga {

hipError_t launch_step_multi(int block, int children, const InstDesc& d, int cur, hipStream_t s, hipEvent_t start,
                             hipEvent_t stop)
{
    switch (block * 8 + children) {
    case 64 * 8 + 2: return launch_multi_tsp<64, 2>(d, cur, s, start, stop);
    case 64 * 8 + 4: return launch_multi_tsp<64, 4>(d, cur, s, start, stop);
    case 128 * 8 + 2: return launch_multi_tsp<128, 2>(d, cur, s, start, stop);
    case 128 * 8 + 4: return launch_multi_tsp<128, 4>(d, cur, s, start, stop);
    case 256 * 8 + 2: return launch_multi_tsp<256, 2>(d, cur, s, start, stop);
    case 256 * 8 + 4: return launch_multi_tsp<256, 4>(d, cur, s, start, stop);
    case 512 * 8 + 2: return launch_multi_tsp<512, 2>(d, cur, s, start, stop);
    default: return hipErrorInvalidValue;
    }
}

bool step_multi_ok(const InstDesc& d, int block, int children, size_t lds_limit)
{
    if (children != 2 && children != 4) return false;
    if (d.prob.kind != KIND_TSP_COORDS || block * children > 1024 || block < 64) return false;
    if (d.prob.n > kCompactItems * block || d.Pi % children != 0 || d.pop % children != 0) return false;
    if (!step_can_rank(d, block, lds_limit)) return false;
    return (size_t)multi_layout(d.prob.n, d.L, d.nb, d.Pi, children).bytes <= lds_limit;
}

}  // namespace brkga
