// kernels_step_plugins.hip - step kernel instantiations for the continuous, job-scheduling and VRP-days plug-ins
// (k_step_one in step_kernel.hpp); one unit per group so they compile in
// parallel.
#include "step_kernel.hpp"

namespace brkga {

hipError_t launch_step_plugins(int kind, int block, const InstDesc& d, int lds, int cur, int init, int rank,
                               hipStream_t s, hipEvent_t start, hipEvent_t stop)
{
    switch (kind) {
    case KIND_CONTINUOUS: return step_kind<KIND_CONTINUOUS>(block, d, lds, cur, init, rank, s, start, stop);
    case KIND_JOBSHOP: return step_kind<KIND_JOBSHOP>(block, d, lds, cur, init, rank, s, start, stop);
    case KIND_VRP_DAYS: return step_kind<KIND_VRP_DAYS>(block, d, lds, cur, init, rank, s, start, stop);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace brkga
