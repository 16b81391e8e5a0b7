// launch.hpp - host-side launchers of the engine's kernels (defined in the
// kernels_*.hip translation units, called by the C ABI layer).
#pragma once

#include <hip/hip_runtime.h>

#include "engine_types.hpp"

namespace brkga {

// Decode + fitness of `count` caller-owned rows (row stride ld floats).
hipError_t launch_evaluate(const ProblemDesc& p, const float* keys, long long ld, int count, float* fit,
                           int* decoded, int block, hipStream_t s);

// One BRKGA generation step (init=1: generation-0 keys + evaluation) for
// every instance in d_descs; `blocks` = total population slots, one
// workgroup each; `lds` = the largest decode layout among the instances.
hipError_t launch_brkga_step(int kind, int block, const InstDesc* d_descs, int num_inst, int blocks, int lds,
                             int init, hipStream_t s);

// Stable per-island sort of fitness (one workgroup per island) that closes
// a generation step and advances each instance's step counter.
hipError_t launch_partition(int block, const InstDesc* d_descs, int num_inst, int blocks, int lds, int init,
                            hipStream_t s);

// NSGA-II: offspring generation + evaluation (one workgroup per child).
hipError_t launch_nsga2_offspring(int kind, int block, const InstDesc* d_descs, int num_inst, int blocks,
                                  int lds, int init, hipStream_t s);
// NSGA-II: ranking, crowding and survivor selection (one workgroup per
// population).
hipError_t launch_nsga2_select(const InstDesc* d_descs, int num_inst, int init, int max_pop, hipStream_t s);
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
