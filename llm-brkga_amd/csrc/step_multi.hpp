// step_multi.hpp - the generation step with several children per workgroup
// (Euclidean TSP, keys held in registers, islands that fit the group).
//
// A workgroup of CH groups of BLOCK threads builds and decodes CH
// consecutive child slots of one island.  The island's rank order (the
// parents' ranking, which the one-child step computes in every workgroup) is
// computed once per workgroup by group 0 and shared through LDS; at a
// migration generation group 1 ranks the preceding island on the ring at
// the same time.  Each group then runs the one-child step's child
// construction and register-key decode on its own LDS; the groups pass the
// same barriers, so slots that are copies (elite, migrants) in a workgroup
// that also builds children decode their copied row too, without storing
// the (identical) fitness.  Workgroups whose slots are all copies copy and
// leave.  Results are those of k_step_one bit for bit.
#pragma once

#include "step_kernel.hpp"

namespace brkga {

struct MultiLayout {
    int dec_bytes, order_off, xch_off, flag_off, bytes;
};

// LDS of one multi-child workgroup: CH decode regions, the own and the
// preceding island's rank orders, two exchange buffers of the ranking
// network and the shared long-bucket flag.
__host__ __device__ inline MultiLayout multi_layout(int n, int L, int nb, int Pi, int ch)
{
    MultiLayout m;
    m.dec_bytes = round_up_int(decode_layout(KIND_TSP_COORDS, n, L, nb, 0).bytes, 16);
    int off = ch * m.dec_bytes;
    m.order_off = off;
    off += round_up_int(8 * Pi, 16);
    m.xch_off = off;
    off += 2 * island_rank_bytes(Pi);
    m.flag_off = off;
    off += 16;
    m.bytes = off;
    return m;
}

// First slot of workgroup w: XCD x = w mod 8 takes the contiguous slots
// [x P/8, (x+1) P/8) (as xcd_slot does for one child per workgroup) when
// 8 CH divides P; a bijection onto the CH-aligned slot groups either way.
template <int CH>
__device__ __forceinline__ int multi_base(int w, int P)
{
    return (P % (8 * CH)) ? w * CH : (w & 7) * (P >> 3) + (w >> 3) * CH;
}

template <int BLOCK, int CH>
__global__ __launch_bounds__(BLOCK * CH) void k_step_multi_tsp(InstDesc d, int cur)
{
    extern __shared__ __align__(16) char lds[];
    const int g = (int)threadIdx.x / BLOCK;
    const int lt = ltid<BLOCK>();
    const int P = d.pop, Pi = d.Pi, Ei = d.Ei, Mi = d.Mi, I = d.islands, k = d.mig_k;
    const int base = multi_base<CH>(blockIdx.x, P);
    const int island = base / Pi;
    const int s0 = base - island * Pi;
    const int n = d.prob.n, nb = d.nb;
    const MultiLayout ml = multi_layout(n, d.L, nb, Pi, CH);
    const DecodeLayout dl = decode_layout(KIND_TSP_COORDS, n, d.L, nb, 0);
    char* glds = lds + g * ml.dec_bytes;
    int* own_order = reinterpret_cast<int*>(lds + ml.order_off);
    int* prev_order = own_order + Pi;
    unsigned long long* xch = reinterpret_cast<unsigned long long*>(lds + ml.xch_off);
    int* long_flag = reinterpret_cast<int*>(lds + ml.flag_off);
    SortScratch ss;
    ss.cnt = reinterpret_cast<int*>(glds + dl.cnt_off);
    ss.red = reinterpret_cast<long long*>(glds + dl.red_off);
    argsort_clear<BLOCK>(ss, nb);
    if (threadIdx.x == 0) *long_flag = 0;

    // group 0 ranks this island, group 1 the preceding one when a migration
    // may be due; their fitness loads are issued before the step counter
    // (read through the scalar cache) is waited for
    const bool maybe_mig = I > 1 && k > 0 && d.mig_interval > 0 && s0 + CH > Pi - k;
    const bool loads = g == 0 || (g == 1 && maybe_mig);
    const int rank_island = g == 0 ? island : (island + I - 1) % I;
    float f = 0.0f;
    if (loads && lt < Pi) f = d.fit[cur][(size_t)(rank_island * Pi + lt) * d.prob.num_obj];
    const int t = load_step_counter(gen_slot(d, cur));
    const bool mig = maybe_mig && t > 0 && (t % d.mig_interval) == 0;
    bool all_copies = true, need_own = false;
#pragma unroll
    for (int j = 0; j < CH; ++j) {
        const int sj = s0 + j;
        const bool migrant = mig && sj >= Pi - k;
        all_copies = all_copies && (migrant || sj < Ei);
        need_own = need_own || (!migrant && sj < Pi - Mi);
    }
    const bool active = (g == 0 && need_own) || (g == 1 && mig);
    group_rank_order<BLOCK>(d, f, lt, active, g == 1 ? prev_order : own_order,
                            xch + (g == 1 ? island_rank_bytes(Pi) / 8 : 0));
    if (blockIdx.x == 0 && threadIdx.x == 0) *gen_slot(d, cur ^ 1) = t + 1;
    __syncthreads();

    const int gs = base + g, s = s0 + g;
    float4* out = reinterpret_cast<float4*>(d.keys[cur ^ 1] + (size_t)gs * d.ld);
    float* fit_out = d.fit[cur ^ 1] + (size_t)gs * d.prob.num_obj;
    float4 mine = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const bool built = build_child<BLOCK, 4, false>(d, island, s, cur, 0, t, own_order, prev_order, out, nullptr,
                                                    fit_out, &mine);
    if (all_copies) return;
    decode_tsp_coords_reg<BLOCK>(d.prob, glds, dl, nb, built ? fit_out : nullptr, mine, long_flag);
}

template <int BLOCK, int CH>
inline hipError_t launch_multi_tsp(const InstDesc& d, int cur, hipStream_t s, hipEvent_t start, hipEvent_t stop)
{
    const int lds = multi_layout(d.prob.n, d.L, d.nb, d.Pi, CH).bytes;
    ensure_lds(k_step_multi_tsp<BLOCK, CH>, lds);
    launch_maybe_timed(k_step_multi_tsp<BLOCK, CH>, dim3(d.pop / CH), dim3(BLOCK * CH), lds, s, start, stop, d, cur);
    return hipGetLastError();
}

}  // namespace brkga
