// kernels_step.hip - entry of the one-population generation step: picks the
// unit that holds the step kernel of the problem kind (k_step_one, defined
// in step_kernel.hpp and instantiated per group of kinds in
// kernels_step_{tsp,knapsack,plugins}.hip).  Replaces the fitness launch
// and the brkga_generation_kernel launch of the GPU-resident generation
// (core/solver.hpp:323, 346).
#include "step_kernel.hpp"

namespace brkga {

int step_lds_bytes(const InstDesc& d)
{
    return step_layout(d.prob.kind, d.prob.n, d.L, d.nb, d.Pi, false, d.prob.aux).bytes;
}

bool step_can_rank(const InstDesc& d, int block, size_t lds_limit)
{
    return d.ls_interval <= 0 && d.Pi <= block && (size_t)(step_lds_bytes(d) + island_rank_bytes(d.Pi)) <= lds_limit;
}

hipError_t launch_step_one(const InstDesc& d, int block, int cur, int init, int rank, hipStream_t s,
                           hipEvent_t start, hipEvent_t stop, int children)
{
    rank = rank && !init;
    // the multi-child form exists for ranking steps only; callers check
    // step_multi_ok
    if (children > 1 && rank) return launch_step_multi(block, children, d, cur, s, start, stop);
    const int lds = step_lds_bytes(d) + (rank ? island_rank_bytes(d.Pi) : 0);
    const int kind = d.prob.kind;
    switch (kind) {
    case KIND_TSP_COORDS:
    case KIND_TSP_MATRIX: return launch_step_tsp_kinds(kind, block, d, lds, cur, init, rank, s, start, stop);
    case KIND_KNAP_THRESHOLD:
    case KIND_KNAP_GREEDY:
    case KIND_TSPJ: return launch_step_knapsack_tspj(kind, block, d, lds, cur, init, rank, s, start, stop);
    case KIND_CONTINUOUS:
    case KIND_JOBSHOP:
    case KIND_VRP_DAYS: return launch_step_plugins(kind, block, d, lds, cur, init, rank, s, start, stop);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace brkga
