// queue_step.hpp - the persistent, queue-driven generation step of many
// resident populations (batched and served instances), as a template
// instantiated per problem kind in the kernels_queue_*.hip units.
#pragma once

#include "launch.hpp"
#include "step_device.hpp"

namespace brkga {

// Persistent generation step over a work queue (batched and served
// populations).  A work item {instance, first slot, end slot, -} covers
// consecutive child slots of one island; items are handed out largest
// instance first.  The workgroup stages the island's rank order and, for
// Euclidean instances, the coordinates once per item, then builds and
// decodes the item's children one after another.  Each instance's step
// counter selects its buffers: -1 builds generation 0 into buffer 0,
// t >= gen_limit leaves the instance untouched.  Workgroup b takes item b
// first, with no atomic; items past the grid are handed out by queue[0]
// (item gridDim.x + queue[0]++), and queue[1] counts retired workgroups, the
// last of which re-arms both, so the kernel can be relaunched (or replayed
// from a graph) as is.  A launch with no more items than workgroups touches
// neither counter: atomics on one address serialise (about 12 ns each), so
// thousands of short items with one or two atomics each used to take longer
// than their work (4608 single-child items of n <= 126: 161 us).
// XYS: Euclidean coordinates are staged in LDS per item (otherwise the
// decoder reads them through L2, leaving LDS to more resident workgroups).
// Forms with at most 4 keys per thread ask for 8 waves per SIMD (at most 64
// VGPRs, a few spilled): their small children are latency-bound and their
// LDS leaves room for more resident workgroups (512 mixed-n instances,
// 3420 -> 3680 solves/s).
template <int KIND, int BLOCK, bool XYS, int ITEMS>
__global__ __launch_bounds__(BLOCK, ITEMS <= 4 ? 8 : 1) void k_step_queue(const InstDesc* __restrict__ descs,
                                                      const int4* __restrict__ items, int num_items,
                                                      int* __restrict__ queue)
{
    extern __shared__ __align__(16) char lds[];
    __shared__ int s_item;
    constexpr bool kXY = KIND == KIND_TSP_COORDS && XYS;
    const bool dynamic = num_items > (int)gridDim.x;
    for (int round = 0;; ++round) {
        __syncthreads();  // the previous item is done with LDS and s_item
        int it = blockIdx.x;
        if (round > 0) {
            if (!dynamic) break;
            if (threadIdx.x == 0) s_item = (int)gridDim.x + atomicAdd(queue, 1);
            __syncthreads();
            it = s_item;
        }
        if (it >= num_items) break;
        const int4 w = items[it];
        const InstDesc& d = descs[w.x];
        const int t = *d.gen;
        if (t >= d.gen_limit) continue;
        const int init = t < 0;
        const int cur = init ? 0 : (t & 1);
        const int nxt = init ? 0 : cur ^ 1;
        const int Pi = d.Pi, n = d.prob.n, M = d.prob.num_obj, ld = d.ld, nb = d.nb;
        const int island = w.y / Pi;
        const StepLayout sl = step_layout(KIND, n, d.L, nb, Pi, kXY, d.prob.aux);
        float4* kl = reinterpret_cast<float4*>(lds + sl.dec.keys_off);
        float2* xy = reinterpret_cast<float2*>(lds + sl.xy_off);
        int* order = reinterpret_cast<int*>(lds + sl.isl_off);
        const bool staged = Pi <= kMaxStagedOrder;
        const int* perm_cur = d.perm[cur] + (size_t)island * Pi;
        SortScratch ss;
        ss.cnt = reinterpret_cast<int*>(lds + sl.dec.cnt_off);
        ss.red = reinterpret_cast<long long*>(lds + sl.dec.red_off);
        if constexpr (kXY) {
            const float2* gxy = reinterpret_cast<const float2*>(d.prob.a);
            for (int i = threadIdx.x; i < n; i += BLOCK) xy[i] = gxy[i];
        }
        if (!init && staged)
            for (int i = threadIdx.x; i < Pi; i += BLOCK) order[i] = perm_cur[i];
        for (int gs = w.y; gs < w.z; ++gs) {
            const int s = gs - island * Pi;
            float4* out = reinterpret_cast<float4*>(d.keys[nxt] + (size_t)gs * ld);
            float* fit_out = d.fit[nxt] + (size_t)gs * M;
            __syncthreads();  // staging done; the previous child's decode is done with LDS
            argsort_clear<BLOCK>(ss, nb);
            const bool child =
                build_child<BLOCK>(d, island, s, cur, init, t, staged ? order : perm_cur, nullptr, out, kl,
                                   fit_out);
            if (!child) continue;
            __syncthreads();
            decode_block<KIND, BLOCK, true, kXY, ITEMS>(d.prob, lds, sl.dec, nb, fit_out, nullptr, kXY ? xy : nullptr,
                                                 nullptr);
        }
    }
    if (dynamic && threadIdx.x == 0) {
        const int retired = atomicAdd(queue + 1, 1);
        if (retired == (int)gridDim.x - 1) {
            atomicExch(queue, 0);
            atomicExch(queue + 1, 0);
        }
    }
}

// Kernel symbol of the queue step for (kind, block, coordinate staging,
// largest n); nullptr when there is no such instantiation.  The kernel
// holds the fewest keys per thread (1, 2, 4, 8 or 16) that the bucket's
// largest n allows: fewer keys, fewer registers and guards, more resident
// workgroups.
template <int KIND, bool XYS, int ITEMS>
inline const void* queue_kernel_items(int block)
{
    switch (block) {
    case 64: return reinterpret_cast<const void*>(k_step_queue<KIND, 64, XYS, ITEMS>);
    case 128: return reinterpret_cast<const void*>(k_step_queue<KIND, 128, XYS, ITEMS>);
    case 256: return reinterpret_cast<const void*>(k_step_queue<KIND, 256, XYS, ITEMS>);
    default: break;
    }
    if constexpr (ITEMS >= kCompactItems) {  // 1 and 2 keys per thread exist up to 256 threads
        if (block == 512) return reinterpret_cast<const void*>(k_step_queue<KIND, 512, XYS, ITEMS>);
        if (block == 1024) return reinterpret_cast<const void*>(k_step_queue<KIND, 1024, XYS, ITEMS>);
    }
    return nullptr;
}

template <int KIND, bool XYS>
inline const void* queue_kernel_kind(int block, int n_max)
{
    if (n_max > kSortItems * block) return nullptr;
    if constexpr (!kind_sorts(KIND)) {
        return queue_kernel_items<KIND, XYS, kSortItems>(block);  // no sort, no per-thread keys
    } else {
        switch (queue_items_for(KIND, block, n_max)) {
        case 1: return queue_kernel_items<KIND, XYS, 1>(block);
        case 2: return queue_kernel_items<KIND, XYS, 2>(block);
        case 4: return queue_kernel_items<KIND, XYS, 4>(block);
        case 8: return queue_kernel_items<KIND, XYS, 8>(block);
        default: return queue_kernel_items<KIND, XYS, kSortItems>(block);
        }
    }
}

// Kernel symbols per group of kinds (nullptr when there is no such
// instantiation), defined in the kernels_queue_*.hip units so that the
// instantiations compile in parallel.
const void* queue_kernel_tsp_coords(int block, bool xys, int n_max);
const void* queue_kernel_matrix_kinds(int kind, int block, bool xys, int n_max);
const void* queue_kernel_knapsack(int kind, int block, bool xys, int n_max);
const void* queue_kernel_plugins(int kind, int block, bool xys, int n_max);

}  // namespace brkga
