// kernels_step_tsp.hip - step kernel instantiations for the two TSP kinds (coordinates and matrix)
// (k_step_one in step_kernel.hpp); one unit per group so they compile in
// parallel.
#include "step_kernel.hpp"

namespace brkga {

hipError_t launch_step_tsp_kinds(int kind, int block, const InstDesc& d, int lds, int cur, int init, int rank,
                                 hipStream_t s, hipEvent_t start, hipEvent_t stop)
{
    switch (kind) {
    case KIND_TSP_COORDS: return step_kind<KIND_TSP_COORDS>(block, d, lds, cur, init, rank, s, start, stop);
    case KIND_TSP_MATRIX: return step_kind<KIND_TSP_MATRIX>(block, d, lds, cur, init, rank, s, start, stop);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace brkga
