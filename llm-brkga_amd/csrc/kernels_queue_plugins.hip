// kernels_queue_plugins.hip - queue-step instantiations for continuous functions, job scheduling and multi-day VRP.
#include "queue_step.hpp"

namespace brkga {

const void* queue_kernel_plugins(int kind, int block, bool xys, int n_max)
{
    if (xys) return nullptr;
    if (kind == KIND_CONTINUOUS) return queue_kernel_kind<KIND_CONTINUOUS, false>(block, n_max);
    if (kind == KIND_JOBSHOP) return queue_kernel_kind<KIND_JOBSHOP, false>(block, n_max);
    if (kind == KIND_VRP_DAYS) return queue_kernel_kind<KIND_VRP_DAYS, false>(block, n_max);
    return nullptr;
}

}  // namespace brkga
