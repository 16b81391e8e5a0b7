// kernels_queue_tsp.hip - queue-step instantiations for Euclidean TSP.
#include "queue_step.hpp"

namespace brkga {

const void* queue_kernel_tsp_coords(int block, bool xys, int n_max)
{
    return xys ? queue_kernel_kind<KIND_TSP_COORDS, true>(block, n_max)
               : queue_kernel_kind<KIND_TSP_COORDS, false>(block, n_max);
}

}  // namespace brkga
