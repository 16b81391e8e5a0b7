// kernels_queue_knapsack.hip - queue-step instantiations for the knapsack decoders.
#include "queue_step.hpp"

namespace brkga {

const void* queue_kernel_knapsack(int kind, int block, bool xys, int n_max)
{
    if (xys) return nullptr;
    if (kind == KIND_KNAP_THRESHOLD) return queue_kernel_kind<KIND_KNAP_THRESHOLD, false>(block, n_max);
    if (kind == KIND_KNAP_GREEDY) return queue_kernel_kind<KIND_KNAP_GREEDY, false>(block, n_max);
    return nullptr;
}

}  // namespace brkga
