// kernels_step_knapsack.hip - step kernel instantiations for the knapsack kinds and TSPJ
// (k_step_one in step_kernel.hpp); one unit per group so they compile in
// parallel.
#include "step_kernel.hpp"

namespace brkga {

hipError_t launch_step_knapsack_tspj(int kind, int block, const InstDesc& d, int lds, int cur, int init, int rank,
                                     hipStream_t s, hipEvent_t start, hipEvent_t stop)
{
    switch (kind) {
    case KIND_KNAP_THRESHOLD: return step_kind<KIND_KNAP_THRESHOLD>(block, d, lds, cur, init, rank, s, start, stop);
    case KIND_KNAP_GREEDY: return step_kind<KIND_KNAP_GREEDY>(block, d, lds, cur, init, rank, s, start, stop);
    case KIND_TSPJ: return step_kind<KIND_TSPJ>(block, d, lds, cur, init, rank, s, start, stop);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace brkga
