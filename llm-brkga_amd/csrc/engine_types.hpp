// engine_types.hpp - plain-old-data descriptors shared by host code and
// kernels.  Kernels receive these by value or read them from device memory,
// so every field is a scalar or a raw device pointer.
#pragma once

#include <stdint.h>

namespace brkga {

// Problem kinds; the numbering is part of the C ABI (brkga_hip.h).
enum ProblemKind : int {
    KIND_TSP_COORDS = 0,      // configs/tsp_config.hpp (Euclidean, from coordinates)
    KIND_TSP_MATRIX = 1,      // configs/tsp_config.hpp (explicit n x n matrix)
    KIND_KNAP_THRESHOLD = 2,  // configs/knapsack_config.hpp:43-75
    KIND_KNAP_GREEDY = 3,     // greedy fill in key order (no reference counterpart)
    KIND_TSPJ = 4,            // configs/tspj_config.hpp:555-642
    KIND_BITSP_COORDS = 5,    // configs/multi_tsp_config.hpp, two coordinate sets
    KIND_BITSP_MATRIX = 6,    // configs/multi_tsp_config.hpp, two matrices
    KIND_CONTINUOUS = 7,      // configs/continuous_opt_gpu_config.hpp:116-160
    KIND_JOBSHOP = 8,         // configs/job_scheduling_gpu_config.hpp:138-158
    KIND_ZDT1 = 9,            // examples/zdt1.cu:43-63 (two objectives)
    KIND_VRP_DAYS = 10,       // configs/vrp_blazer_config.hpp:192-296
    KIND_MOVRP = 11,          // configs/multi_objective_vrp_config.hpp:208-300 (two objectives)
    KIND_HOST = 12,           // no device decoder: the caller evaluates on the host
                              // (BRKGAConfig::fitness_function, population.hpp:626-653)
    KIND_COUNT = 13
};

// Kinds whose decoder argsorts keys (and so has per-thread key registers).
__host__ __device__ constexpr bool kind_sorts(int kind)
{
    return !(kind == KIND_KNAP_THRESHOLD || kind == KIND_CONTINUOUS || kind == KIND_JOBSHOP || kind == KIND_ZDT1 ||
             kind == KIND_HOST);
}

// Continuous test functions (KIND_CONTINUOUS, ProblemDesc::ip0).
enum ContinuousFn : int { FN_SPHERE = 0, FN_RASTRIGIN = 1, FN_ROSENBROCK = 2 };

__host__ __device__ inline int kind_num_obj(int kind)
{
    return (kind == KIND_BITSP_COORDS || kind == KIND_BITSP_MATRIX || kind == KIND_ZDT1 || kind == KIND_MOVRP) ? 2
                                                                                                               : 1;
}

__host__ __device__ inline int kind_chrom_len(int kind, int n)
{
    return kind == KIND_TSPJ ? 2 * n : n;
}

enum RngTag : uint32_t {
    TAG_PARENTS = 0,
    TAG_COINS = 1,
    TAG_MUTANT = 2,
    TAG_INIT = 3,
    TAG_MUTCOIN = 4,
    TAG_LSEARCH = 5,
    TAG_DIVERSITY = 6
};

struct ProblemDesc {
    int kind;
    int n;          // cities / items (a depot, where there is one, is extra)
    int chrom_len;  // keys per individual (TSPJ: 2n)
    int num_obj;    // 1, or 2 for the bi-objective TSP
    int dim;        // rows of a / b (coords count or matrix dimension)
    int maximize;   // 1: larger fitness is better (knapsack)
    const float* a; // coords [dim][2] or matrix [dim][dim]
    const float* b; // second matrix / coordinate set
    const int* w;   // knapsack weights
    const int* v;   // knapsack values
    long long cap;  // knapsack capacity / vehicle capacity
    float scale0, inv0, scale1, inv1;  // fixed-point 2^S and 2^-S per objective
    int ip0, ip1;   // kind parameters: continuous function, machines, stops
                    // per day, vehicles
    float f0, f1;   // continuous bounds [f0, f1]
    double d0;      // VRP days: drive-time limit per day (minutes)
    int aux;        // extra LDS bytes the decoder needs (machine bins, ...)
};

// One resident population (a BRKGA run of `islands` islands, or one NSGA-II
// population) as seen by the generation kernels.  Buffers ping-pong on the
// parity of the step counter *gen: step t reads [t & 1] and writes
// [(t + 1) & 1].
struct InstDesc {
    ProblemDesc prob;
    float* keys[2];   // [pop][ld]
    float* fit[2];    // [pop][num_obj]
    int* perm[2];     // [islands][Pi]: local slot of each rank
    int* rank[2];     // NSGA-II front index per slot
    float* crowd[2];  // NSGA-II crowding distance per slot
    unsigned* dom;    // NSGA-II dominated-by bitsets [rows][rows / 32] (null: counted in the selection)
    int* gen;         // completed generation steps; gen[3] is the odd-parity slot of the ranking steps (gen_slot)
    int* done;        // arrival counter of the per-island partition blocks
    int* ls_flag;     // BRKGA: 1 while a local search of the current generation is due;
                      // NSGA-II: members kept by a diversity injection in progress (0: none)
    int L, ld;
    int pop, Pi, Ei, Mi, islands;
    int mig_interval, mig_k;
    uint32_t seed_lo, seed_hi;
    float rho;
    float mut_rate;   // NSGA-II per-gene resampling probability
    int ls_interval, ls_k, ls_moves;  // TSPJ local search (0 = off)
    int nb;           // argsort buckets for n keys
    int slot_begin;   // first block of this instance in a step launch
    int seg_begin;    // first block of this instance in a partition launch
    int gen_limit;    // steps are skipped once *gen >= gen_limit
    int div_interval; // NSGA-II diversity injection after steps t % div_interval == 0 (0 = off)
    double div_eps2;  // squared objective distance below which two members count as one
};

// NSGA-II populations of up to this many rows (parents + children) get
// dominated-by bitsets: one row per thread of the selection workgroup.
constexpr int kDomRows = 1024;

__host__ __device__ inline int round_up_int(int v, int m)
{
    return (v + m - 1) / m * m;
}

// LDS carve-up of one decode block.
struct DecodeLayout {
    int keys_off, perm_off, items_off, cnt_off, red_off, aux_off, bytes;
};

// LDS of a generation-step block: the decode layout, then the staged
// coordinates (when staged), then the staged rank order of the block's
// island.
struct StepLayout {
    DecodeLayout dec;
    int xy_off, isl_off, bytes;
};

// Bytes of coordinates a decoder of this kind can stage in LDS.
__host__ __device__ inline int coords_bytes(int kind, int n)
{
    if (kind == KIND_TSP_COORDS) return 8 * n;
    if (kind == KIND_BITSP_COORDS) return 16 * (n + 1);
    return 0;
}

// aux: the problem's extra decoder bytes (ProblemDesc::aux), placed last.
__host__ __device__ inline DecodeLayout decode_layout(int kind, int n, int L, int nb, int aux)
{
    DecodeLayout d;
    if (kind == KIND_HOST) {
        // no decode: the keys of the row being built, nothing else
        d.red_off = d.items_off = d.keys_off = d.perm_off = d.cnt_off = d.aux_off = 0;
        d.bytes = round_up_int(L, 32) * 4;
        return d;
    }
    int off = 0;
    d.red_off = off;
    off += 17 * 8;
    off = round_up_int(off, 16);
    d.items_off = off;
    off += (n + 4) * 8;  // four spare items: the ranking reads four from any bucket start
    off = round_up_int(off, 16);
    // the decoded order overlays the keys: the argsort reads each key for
    // the last time before it writes the first position
    d.keys_off = off;
    d.perm_off = off;
    const int key_bytes = round_up_int(L, 32) * 4;
    const int perm_bytes = round_up_int(kind == KIND_TSPJ ? 2 * n : n, 4) * 4 + 16;
    off += key_bytes > perm_bytes ? key_bytes : perm_bytes;
    d.cnt_off = off;
    off += (nb + 4) * 4;
    off = round_up_int(off, 16);
    d.aux_off = off;
    off += aux;
    d.bytes = round_up_int(off, 16);
    return d;
}

// Keys one decode thread holds: a workgroup of B threads decodes n <= 16 B.
// The step kernels also come in forms for n <= 4 B and n <= 8 B, which hold
// fewer registers.
constexpr int kSortItems = 16;
constexpr int kCompactItems = 4;

// Rank orders longer than this are read from global memory, not staged.
constexpr int kMaxStagedOrder = 2048;

// Device set-up of one new population, for set-ups batched into one launch
// (launch_setup_pops): both key buffers cleared (key_floats floats each,
// a multiple of 4), step counter slots gen[0..3] := {gen_edge, 0, 0,
// gen_edge}, and the descriptor copied to d_desc.
struct PopSetup {
    float* keys[2];
    size_t key_floats;
    int* gen;
    int gen_edge;
    InstDesc* d_desc;
    InstDesc desc;
};

// Step counter a ranking step reads when its parent buffer has parity
// `parity`: gen[0] (even) or gen[3] (odd).  The step writes t + 1 to the
// other slot, so no workgroup of a launch reads what that launch writes.
__host__ __device__ inline int* gen_slot(const InstDesc& d, int parity)
{
    return parity ? d.gen + 3 : d.gen;
}

// LDS a step that ranks its parent island itself adds to the step layout:
// two exchange buffers of pow2(Pi) 64-bit keys for the network stages that
// cross waves (none when the island fits one wave).
__host__ __device__ inline int island_rank_bytes(int Pi)
{
    int pn = 2;
    while (pn < Pi) pn <<= 1;
    return pn > 64 ? 16 * pn : 0;
}

__host__ __device__ inline StepLayout step_layout(int kind, int n, int L, int nb, int Pi, bool stage_xy, int aux)
{
    StepLayout s;
    s.dec = decode_layout(kind, n, L, nb, aux);
    s.xy_off = s.dec.bytes;
    s.isl_off = s.xy_off + (stage_xy ? round_up_int(coords_bytes(kind, n), 16) : 0);
    s.bytes = s.isl_off + round_up_int(4 * (Pi <= kMaxStagedOrder ? Pi : 0), 16);
    return s;
}

}  // namespace brkga
