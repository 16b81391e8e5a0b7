// kernels_queue.hip - launchers of the persistent, queue-driven generation
// step of many resident populations (batched and served instances; the
// kernel template is in queue_step.hpp, instantiated per group of problem
// kinds in kernels_queue_*.hip), and the step-counter reset.
#include "launch_util.hpp"
#include "queue_step.hpp"

namespace brkga {

// Sets the step counter of every listed instance to `gen` and clears its
// partition arrival counter.
__global__ __launch_bounds__(256) void k_set_counters(const InstDesc* __restrict__ descs, int num, int gen)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= num) return;
    descs[i].gen[0] = gen;
    descs[i].gen[3] = gen;  // the odd-parity slot of steps that rank their island
    *descs[i].done = 0;
}

namespace {

const void* queue_kernel(int kind, int block, bool xys, int n_max)
{
    switch (kind) {
    case KIND_TSP_COORDS: return queue_kernel_tsp_coords(block, xys, n_max);
    case KIND_TSP_MATRIX:
    case KIND_TSPJ: return queue_kernel_matrix_kinds(kind, block, xys, n_max);
    case KIND_KNAP_THRESHOLD:
    case KIND_KNAP_GREEDY: return queue_kernel_knapsack(kind, block, xys, n_max);
    case KIND_CONTINUOUS:
    case KIND_JOBSHOP:
    case KIND_VRP_DAYS: return queue_kernel_plugins(kind, block, xys, n_max);
    default: return nullptr;
    }
}

}  // namespace

hipError_t launch_step_queue(int kind, int block, bool xys, int n_max, const InstDesc* d_descs, const int4* d_items,
                             int num_items, int* d_queue, int grid, int lds, hipStream_t s)
{
    const void* f = queue_kernel(kind, block, xys, n_max);
    if (!f) return hipErrorInvalidValue;
    if (num_items <= 0 || grid <= 0) return hipSuccess;
    ensure_lds(f, lds);
    void* args[] = {(void*)&d_descs, (void*)&d_items, (void*)&num_items, (void*)&d_queue};
    return hipLaunchKernel(f, dim3(grid), dim3(block), args, (size_t)lds, s);
}

int step_queue_blocks_per_cu(int kind, int block, bool xys, int n_max, int lds)
{
    const void* f = queue_kernel(kind, block, xys, n_max);
    if (!f) return 0;
    ensure_lds(f, lds);
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, block, (size_t)lds) != hipSuccess) return 0;
    return nb;
}

// A fixed-length chain of dependent ALU operations in one wave (no memory
// waits, so it always ends): timed on two streams at once it shows whether
// their work runs concurrently.
__global__ __launch_bounds__(64) void k_lane_probe(int iters, float* out)
{
    float x = (float)threadIdx.x;
    for (int i = 0; i < iters; ++i) x = x * 0.999f + 1.0f;
    out[threadIdx.x] = x;
}

hipError_t launch_lane_probe(int iters, float* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_lane_probe, dim3(1), dim3(64), 0, s, iters, out);
    return hipGetLastError();
}

hipError_t launch_set_counters(const InstDesc* d_descs, int num, int gen, hipStream_t s)
{
    if (num <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_set_counters, dim3((num + 255) / 256), dim3(256), 0, s, d_descs, num, gen);
    return hipGetLastError();
}

}  // namespace brkga
