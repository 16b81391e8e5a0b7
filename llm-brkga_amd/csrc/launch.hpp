// launch.hpp - host-side launchers of the engine's kernels (defined in the
// kernels_*.hip translation units, called by the C ABI layer).
#pragma once

#include <hip/hip_runtime.h>

#include "engine_types.hpp"

namespace brkga {

// Decode + fitness of `count` caller-owned rows (row stride ld floats).
hipError_t launch_evaluate(const ProblemDesc& p, const float* keys, long long ld, int count, float* fit,
                           int* decoded, int block, hipStream_t s);

// One BRKGA generation step of one population, descriptor passed by value.
// cur = buffer parity of the parent generation; init=1 builds generation 0
// (Philox keys + evaluation) in buffer 0.  rank=1 (step_can_rank): each
// workgroup ranks its parent island itself and the step advances the
// parity-slot step counter (gen_slot), so no partition follows; a run of
// such steps ends with a partition in mode 3.
// start / stop (optional): events that take the kernel's own start and end
// times (hipExtLaunchKernel), for profiling without marker packets.
// children > 1 (step_multi_ok): each workgroup builds that many children
// of one island and ranks the island once for all of them (ranking steps,
// Euclidean TSP with keys in registers); results do not change.
hipError_t launch_step_one(const InstDesc& d, int block, int cur, int init, int rank, hipStream_t s,
                           hipEvent_t start = nullptr, hipEvent_t stop = nullptr, int children = 1);
// Whether launch_step_one can rank in the step at this workgroup size with
// `lds_limit` bytes of LDS per workgroup.
bool step_can_rank(const InstDesc& d, int block, size_t lds_limit);
// Whether the multi-child step runs this population at (block, children).
bool step_multi_ok(const InstDesc& d, int block, int children, size_t lds_limit);
hipError_t launch_step_multi(int block, int children, const InstDesc& d, int cur, hipStream_t s, hipEvent_t start,
                             hipEvent_t stop);

// Stable per-island sort of the fitness in buffer `buf` into the rank order.
// mode 0: sort only; 1: closes a generation step (bumps the step counter and
// raises the local-search flag when the new generation is due); 2: sorts the
// current generation again after a local search, only when the flag is
// raised, and lowers it (buf is then taken from the step counter); 3: closes
// a run of ranking steps whose last generation is in buffer `buf` (both
// step counter slots := that generation's counter).
// The set-ups of `count` new populations (list in device memory), one
// launch.
hipError_t launch_setup_pops(const PopSetup* d_list, int count, hipStream_t s);

hipError_t launch_partition_one(const InstDesc& d, int buf, int mode, hipStream_t s, hipEvent_t start = nullptr,
                                hipEvent_t stop = nullptr);
// buf < 0: buffer from each population's step counter.
hipError_t launch_partition_many(const InstDesc* d_descs, int num_inst, int blocks, int max_Pi, int buf, int mode,
                                 hipStream_t s);
// The same for the `count` listed {member, island} pairs only.
hipError_t launch_partition_list(const InstDesc* d_descs, const int2* d_segs, int count, int max_Pi, int buf,
                                 int mode, hipStream_t s);

// TSPJ local search of the current generation of every listed population
// whose local-search flag is raised; `items` = largest islands x ls_k.
hipError_t launch_tspj_local_search(const InstDesc* d_descs, int num_inst, int items, int block, int lds,
                                    hipStream_t s);
int tspj_ls_lds_bytes(int n, int nb);

// Persistent queue step over work items {instance, first slot, end slot, -}
// of one problem kind (batched / served populations) whose problem sizes are
// at most n_max.  xys: stage the Euclidean coordinates in LDS (coordinate
// kind only).  d_queue holds two counters that must be zero at the first
// launch; the kernel re-arms them.
hipError_t launch_step_queue(int kind, int block, bool xys, int n_max, const InstDesc* d_descs, const int4* d_items,
                             int num_items, int* d_queue, int grid, int lds, hipStream_t s);
// Keys per thread of the queue-step kernel for a bucket of problem sizes up
// to n_max on `block` threads: the fewest of 1, 2, 4, 8, 16 that holds n_max
// (at least 4 above 256 threads; 16 for kinds that do not sort).
inline int queue_items_for(int kind, int block, int n_max)
{
    if (!kind_sorts(kind)) return kSortItems;
    for (int items = block > 256 ? kCompactItems : 1; items < kSortItems; items <<= 1)
        if (n_max <= items * block) return items;
    return kSortItems;
}
// Resident workgroups of the queue step per CU (0: no such kernel).
int step_queue_blocks_per_cu(int kind, int block, bool xys, int n_max, int lds);
// One wave of `iters` dependent ALU steps writing out[0, 64): a probe of
// whether two streams run concurrently.
hipError_t launch_lane_probe(int iters, float* out, hipStream_t s);
// Step counter (both parity slots) := gen and arrival counter := 0 for `num`
// instances.
hipError_t launch_set_counters(const InstDesc* d_descs, int num, int gen, hipStream_t s);

// Keys of the next generation of a host-evaluated BRKGA population (one
// workgroup per slot, no decode): elite and migrant rows are copied with
// their fitness, the others are left for the caller to evaluate.  The
// partition that follows closes the generation.
hipError_t launch_host_children(const InstDesc& d, int cur, int init, hipStream_t s);

int step_lds_bytes(const InstDesc& d);
int partition_lds_bytes(int Pi);

// NSGA-II: offspring generation + evaluation (one workgroup per child).
hipError_t launch_nsga2_offspring(int kind, int block, const InstDesc* d_descs, int num_inst, int blocks,
                                  int lds, int init, hipStream_t s);
// NSGA-II: ranking, crowding and survivor selection (one workgroup per
// population).  mode 0: a generation step (parents + children); 1: rank
// generation 0 built by the host path (the counter stays); 2: rank the N
// members of the current generation again after a diversity injection
// (populations without one in progress are skipped).
hipError_t launch_nsga2_select(const InstDesc* d_descs, int num_inst, int mode, int max_pop, hipStream_t s);
// NSGA-II diversity injection after a step's gather (populations whose step
// is not due, or with the injection off, are skipped on the device): the
// unique members and the 70% test, the kept rows compacted, new rows drawn
// and decoded (`rows` = child blocks of the offspring launch, `lds` its
// LDS), then the members ranked and crowded again.
hipError_t launch_nsga2_diversity(int kind, int block, const InstDesc* d_descs, int num_inst, int rows, int lds,
                                  int max_pop, hipStream_t s);
// NSGA-II: gather of survivor rows into the next buffer.
hipError_t launch_nsga2_gather(const InstDesc* d_descs, int num_inst, int max_pop, hipStream_t s);

inline int default_block(int n)
{
    int b = 64;
    while (b < 1024 && b * 4 < n) b <<= 1;
    return b;
}

inline int partition_block(int Pi)
{
    int b = 64;
    while (b < 1024 && b * 2 < Pi) b <<= 1;
    return b;
}

}  // namespace brkga
