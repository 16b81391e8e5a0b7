// kernels_queue_matrix.hip - queue-step instantiations for TSP over a matrix and TSPJ.
#include "queue_step.hpp"

namespace brkga {

const void* queue_kernel_matrix_kinds(int kind, int block, bool xys, int n_max)
{
    if (xys) return nullptr;
    if (kind == KIND_TSP_MATRIX) return queue_kernel_kind<KIND_TSP_MATRIX, false>(block, n_max);
    if (kind == KIND_TSPJ) return queue_kernel_kind<KIND_TSPJ, false>(block, n_max);
    return nullptr;
}

}  // namespace brkga
