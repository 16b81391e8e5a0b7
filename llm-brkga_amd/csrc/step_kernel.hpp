// step_kernel.hpp - one BRKGA generation step of one population, one
// workgroup per child slot: elite / migrant row copy, mutant keys from
// Philox, or biased uniform crossover, then decode + fitness of the new row
// straight from LDS.  Replaces brkga_generation_kernel
// (core/cuda_kernels.cuh:88-138) fused with the config fitness kernels, and
// the host-staged ring migration (core/solver.hpp:517-558).
//
// The population descriptor is a kernel argument.  `cur` is the buffer
// parity of the parent generation, known to the launcher; the generation
// counter in device memory only feeds the Philox counter and the migration
// schedule, so a captured graph can be replayed.  Key rows are [pop][ld]
// FP32 with ld a multiple of 32, so every row starts 128-byte aligned and is
// streamed as float4.
#pragma once

#include "launch.hpp"
#include "launch_util.hpp"
#include "step_device.hpp"

namespace brkga {

// A step counter slot read through the scalar cache: the slot a step reads
// was written by an earlier launch (kernel start invalidates the cache) and
// no workgroup of the same launch writes it.  A scalar load is waited for
// on its own counter, so loads issued after it (the island's fitness, the
// staged rank order) are in flight before the counter is needed.
__device__ __forceinline__ int load_step_counter(const int* slot)
{
    return *(const __attribute__((address_space(4))) int*)slot;
}

// One block of a generation step: slot gs of the population described by d.
// n <= ITEMS * BLOCK.  rank = 0: the parents' rank order comes from the
// partition kernel (perm[cur]) and the partition advances the step counter.
// rank = 1 (pow2(Pi) <= BLOCK, the LDS of island_rank_bytes(Pi) after the
// step layout): the block ranks the parent island itself (the preceding
// island for a migrant slot) and reads the step counter from its parity
// slot; block 0 writes t + 1 to the other slot.
template <int KIND, int BLOCK, int ITEMS>
__device__ __forceinline__ void step_body(const InstDesc& d, int gs, int island, int s, int cur, int init, int rank,
                                          char* lds)
{
    const int Pi = d.Pi, Ei = d.Ei, I = d.islands;
    const int nxt = init ? 0 : cur ^ 1;
    const int M = d.prob.num_obj;
    const int ld = d.ld, L = d.L, n = d.prob.n;
    const int nq = (L + 3) >> 2;
    float4* out = reinterpret_cast<float4*>(d.keys[nxt] + (size_t)gs * ld);
    float* fit_out = d.fit[nxt] + (size_t)gs * M;
    const int* perm_cur = d.perm[cur] + (size_t)island * Pi;
    const int k = d.mig_k;
    const bool maybe_migrant = !init && I > 1 && k > 0 && d.mig_interval > 0 && s >= Pi - k;

    // elite slot that no migration can claim: copy row + fitness, no decode
    if (!init && !rank && !maybe_migrant && s < Ei) {
        const int src = island * Pi + perm_cur[s];
        const float4* in = reinterpret_cast<const float4*>(d.keys[cur] + (size_t)src * ld);
        for (int c = threadIdx.x; c < nq; c += BLOCK) out[c] = in[c];
        if (threadIdx.x < M) fit_out[threadIdx.x] = d.fit[cur][(size_t)src * M + threadIdx.x];
        return;
    }

    const int nb = d.nb;
    const StepLayout sl = step_layout(KIND, n, L, nb, Pi, false, d.prob.aux);
    float4* kl = reinterpret_cast<float4*>(lds + sl.dec.keys_off);
    int* order = reinterpret_cast<int*>(lds + sl.isl_off);
    SortScratch ss;
    ss.cnt = reinterpret_cast<int*>(lds + sl.dec.cnt_off);
    ss.red = reinterpret_cast<long long*>(lds + sl.dec.red_off);
    int t = 0;
    const int* own_order = order;
    const int* mig_order = nullptr;
    if (init) {
        argsort_clear<BLOCK>(ss, nb);
    } else if (rank) {
        // the step counter is read while the island's fitness loads are in
        // flight; only a slot that migration may claim needs it first, to
        // choose the island it ranks.  The argsort counters (their own LDS)
        // are cleared first; one barrier below ends both.
        argsort_clear<BLOCK>(ss, nb);
        auto after_loads = [&]() {
            if (!maybe_migrant) t = load_step_counter(gen_slot(d, cur));
        };
        if (!maybe_migrant) {
            if (s < Ei || s < Pi - d.Mi)
                island_rank_order<BLOCK>(d, island, cur, order, lds + sl.bytes, after_loads);
            else
                t = load_step_counter(gen_slot(d, cur));
        } else {
            t = load_step_counter(gen_slot(d, cur));
            if (t > 0 && (t % d.mig_interval) == 0) {
                island_rank_order<BLOCK>(d, (island + I - 1) % I, cur, order, lds + sl.bytes, after_loads);
                mig_order = order;
            } else if (s < Ei || s < Pi - d.Mi) {
                island_rank_order<BLOCK>(d, island, cur, order, lds + sl.bytes, after_loads);
            }
        }
        if (gs == 0 && threadIdx.x == 0) *gen_slot(d, cur ^ 1) = t + 1;
    } else {
        t = load_step_counter(d.gen);
        // stage the island rank order, clear the argsort counters
        if (Pi <= kMaxStagedOrder)
            for (int i = threadIdx.x; i < Pi; i += BLOCK) order[i] = perm_cur[i];
        else
            own_order = perm_cur;
        argsort_clear<BLOCK>(ss, nb);
    }
    __syncthreads();

    if constexpr (KIND == KIND_TSP_COORDS && (ITEMS == kCompactItems || ITEMS == 2)) {
        // n <= ITEMS BLOCK: each thread ranks the one column of ITEMS keys it
        // built, from registers, without waiting for the others
        using V = typename KeyCol<ITEMS>::T;
        V mine{};
        V* out_w = reinterpret_cast<V*>(d.keys[nxt] + (size_t)gs * ld);
        // the decode ranks the keys from `mine`: no copy of the row in LDS
        if (!build_child<BLOCK, ITEMS, false>(d, island, s, cur, init, t, own_order, mig_order, out_w, nullptr,
                                              fit_out, &mine))
            return;
        decode_tsp_coords_reg<BLOCK>(d.prob, lds, sl.dec, nb, fit_out, mine);
    } else {
        const bool child = build_child<BLOCK>(d, island, s, cur, init, t, own_order, mig_order, out, kl, fit_out);
        if (!child) return;
        __syncthreads();
        // inlined here: a call would keep a stack frame in scratch for the widest decoders
        [[clang::always_inline]] decode_block<KIND, BLOCK, true, false, ITEMS>(d.prob, lds, sl.dec, nb, fit_out, nullptr);
    }
}

// Child slot gs of workgroup b, with its island and its slot s in the
// island.  Workgroups go to the 8 XCDs round-robin (XCD b % 8), so XCD x
// gets the contiguous slots [x P/8, (x+1) P/8): an island's children,
// whose rows and fitness the island's next generation reads, stay in the
// L2 of one or two XCDs.  A range that starts an island and ends inside it
// is rotated by the elite count, so that its workgroups of lower index
// (dispatched first, and older on their CU, whose wave arbiter favours
// older waves) build offspring and its elite copies, the shortest work,
// come last.  Identity when 8 does not divide P.  A bijection on [0, P)
// either way.  One integer division: a range no longer than an island
// crosses at most one island boundary.
struct SlotPos {
    int gs, island, s;
};
__device__ __forceinline__ SlotPos xcd_slot(int b, int P, int Pi, int Ei)
{
    const int R = P >> 3;
    if ((P & 7) == 0 && R <= Pi) {
        const int lo = (b & 7) * R;
        const int il = lo / Pi, sl = lo - il * Pi;
        int j = b >> 3;
        if (sl == 0 && Ei < R) {
            j += Ei;
            if (j >= R) j -= R;
        }
        const int s = sl + j;
        return s >= Pi ? SlotPos{lo + j, il + 1, s - Pi} : SlotPos{lo + j, il, s};
    }
    const int gs = (P & 7) ? b : (b & 7) * R + (b >> 3);
    const int island = gs / Pi;
    return SlotPos{gs, island, gs - island * Pi};
}

template <int KIND, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_step_one(InstDesc d, int cur, int init, int rank)
{
    extern __shared__ __align__(16) char lds[];
    const SlotPos q = xcd_slot(blockIdx.x, d.pop, d.Pi, d.Ei);
    step_body<KIND, BLOCK, ITEMS>(d, q.gs, q.island, q.s, cur, init, rank, lds);
}

template <int KIND, int BLOCK>
inline hipError_t step_single(const InstDesc& d, int lds, int cur, int init, int rank, hipStream_t s,
                              hipEvent_t start, hipEvent_t stop)
{
    if constexpr (KIND == KIND_TSP_COORDS && BLOCK >= 512) {
        // two keys per thread (float2 columns) when they fit
        if (d.prob.n <= 2 * BLOCK) {
            ensure_lds(k_step_one<KIND, BLOCK, 2>, lds);
            launch_maybe_timed(k_step_one<KIND, BLOCK, 2>, dim3(d.pop), dim3(BLOCK), lds, s, start, stop, d, cur, init,
                               rank);
            return hipGetLastError();
        }
    }
    if constexpr (kind_sorts(KIND)) {
        if (d.prob.n <= kCompactItems * BLOCK) {
            ensure_lds(k_step_one<KIND, BLOCK, kCompactItems>, lds);
            launch_maybe_timed(k_step_one<KIND, BLOCK, kCompactItems>, dim3(d.pop), dim3(BLOCK), lds, s, start, stop,
                               d, cur, init, rank);
            return hipGetLastError();
        }
        if (d.prob.n <= 2 * kCompactItems * BLOCK) {
            ensure_lds(k_step_one<KIND, BLOCK, 2 * kCompactItems>, lds);
            launch_maybe_timed(k_step_one<KIND, BLOCK, 2 * kCompactItems>, dim3(d.pop), dim3(BLOCK), lds, s, start,
                               stop, d, cur, init, rank);
            return hipGetLastError();
        }
    }
    ensure_lds(k_step_one<KIND, BLOCK, kSortItems>, lds);
    launch_maybe_timed(k_step_one<KIND, BLOCK, kSortItems>, dim3(d.pop), dim3(BLOCK), lds, s, start, stop, d, cur,
                       init, rank);
    return hipGetLastError();
}

template <int KIND>
inline hipError_t step_kind(int block, const InstDesc& d, int lds, int cur, int init, int rank, hipStream_t s,
                            hipEvent_t start, hipEvent_t stop)
{
    switch (block) {
    case 64: return step_single<KIND, 64>(d, lds, cur, init, rank, s, start, stop);
    case 128: return step_single<KIND, 128>(d, lds, cur, init, rank, s, start, stop);
    case 256: return step_single<KIND, 256>(d, lds, cur, init, rank, s, start, stop);
    case 512: return step_single<KIND, 512>(d, lds, cur, init, rank, s, start, stop);
    case 1024: return step_single<KIND, 1024>(d, lds, cur, init, rank, s, start, stop);
    default: return hipErrorInvalidValue;
    }
}

// Step launchers per group of kinds, defined in the kernels_step_*.hip
// units so that the instantiations compile in parallel.
hipError_t launch_step_tsp_kinds(int kind, int block, const InstDesc& d, int lds, int cur, int init, int rank,
                                 hipStream_t s, hipEvent_t start, hipEvent_t stop);
hipError_t launch_step_knapsack_tspj(int kind, int block, const InstDesc& d, int lds, int cur, int init, int rank,
                                     hipStream_t s, hipEvent_t start, hipEvent_t stop);
hipError_t launch_step_plugins(int kind, int block, const InstDesc& d, int lds, int cur, int init, int rank,
                               hipStream_t s, hipEvent_t start, hipEvent_t stop);

}  // namespace brkga
