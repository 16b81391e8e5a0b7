// step_device.hpp - device helpers shared by the generation-step kernels
// (single-instance and queue-driven).
#pragma once

#include "decoders.hpp"

namespace brkga {

__device__ __forceinline__ int find_instance_by_slot(const InstDesc* __restrict__ d, int num, int b)
{
    int lo = 0, hi = num - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (d[mid].slot_begin <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int find_instance_by_seg(const InstDesc* __restrict__ d, int num, int b)
{
    int lo = 0, hi = num - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (d[mid].seg_begin <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Order-preserving FP32 -> u32 (smaller = better): -0 == +0, NaN last.
__device__ __forceinline__ uint32_t fitness_order_key(float f, int maximize)
{
    if (f != f) return 0xFFFFFFFFu;
    if (f == 0.0f) f = 0.0f;
    uint32_t u = __float_as_uint(f);
    u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    return maximize ? ~u : u;
}

// The 32-bit value of lane (lane ^ X) of the wave, 0 < X < 64: DPP quad
// permutes for 1, 2 and 3, row rotation by 8 for 8, the row half mirror for
// 7, the row mirror for 15, ds_swizzle's xor mode (within 32 lanes) for 4,
// 16 and 31, ds_bpermute from byte address `bp` = 4 (lane ^ X) otherwise.
// Every lane must be active.
template <int X>
__device__ __forceinline__ unsigned xor_lane_u32(unsigned v, int bp)
{
    if constexpr (X == 1) return (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xf, 0xf, true);  // [1,0,3,2]
    else if constexpr (X == 2) return (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xf, 0xf, true);  // [2,3,0,1]
    else if constexpr (X == 3) return (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x1B, 0xf, 0xf, true);  // [3,2,1,0]
    else if constexpr (X == 7) return (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x141, 0xf, 0xf, true);  // half mirror
    else if constexpr (X == 8) return (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x128, 0xf, 0xf, true);  // row_ror:8
    else if constexpr (X == 15) return (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x140, 0xf, 0xf, true);  // mirror
    else if constexpr (X == 4 || X == 16 || X == 31)
        return (unsigned)__builtin_amdgcn_ds_swizzle((int)v, 0x1F | (X << 10));
    else return (unsigned)__builtin_amdgcn_ds_bpermute(bp, (int)v);
}

// Lane (lane ^ X) of a 64-bit value, or through LDS (two buffers of PN
// keys, one barrier per exchange) when X reaches across waves.
template <int PN, int X>
__device__ __forceinline__ unsigned long long xor_exchange(unsigned long long v, unsigned long long* xch, int bp,
                                                           int& half)
{
    if constexpr (X >= kWave) {
        const int t = threadIdx.x;
        unsigned long long* b = xch + half * PN;
        half ^= 1;
        if (t < PN) b[t] = v;
        __syncthreads();
        return t < PN ? b[t ^ X] : v;
    } else {
        const unsigned lo = xor_lane_u32<X>((unsigned)v, bp);
        const unsigned hi = xor_lane_u32<X>((unsigned)(v >> 32), bp);
        return ((unsigned long long)hi << 32) | lo;
    }
}

// Half-cleaner steps J, J/2, .., 1 of an ascending bitonic network in flip
// form (every compare-exchange ascending): the thread with bit J clear
// keeps the minimum.  bit[i] = bit i of the thread index; bp32 / bp63 =
// ds_bpermute addresses of lanes lane ^ 32 and lane ^ 63.
template <int PN, int J>
__device__ __forceinline__ void bitonic_half_cleaners(unsigned long long& v, unsigned long long* xch, const bool* bit,
                                                      int bp32, int& half)
{
    const unsigned long long o = xor_exchange<PN, J>(v, xch, bp32, half);
    if ((o < v) != bit[__builtin_ctz(J)]) v = o;
    if constexpr (J > 1) bitonic_half_cleaners<PN, J / 2>(v, xch, bit, bp32, half);
}

// Merge level K (blocks of K keys): the flip step pairs t with t ^ (K - 1),
// the lower thread (bit K/2 clear) keeping the minimum, then the half
// cleaners from K/4 down.
template <int PN, int K>
__device__ __forceinline__ void bitonic_flip_levels(unsigned long long& v, unsigned long long* xch, const bool* bit,
                                                    int bp32, int bp63, int& half)
{
    const unsigned long long o = xor_exchange<PN, K - 1>(v, xch, K - 1 == 63 ? bp63 : bp32, half);
    if ((o < v) != bit[__builtin_ctz(K) - 1]) v = o;
    if constexpr (K >= 4) bitonic_half_cleaners<PN, K / 4>(v, xch, bit, bp32, half);
    if constexpr (K < PN) bitonic_flip_levels<PN, 2 * K>(v, xch, bit, bp32, bp63, half);
}

// Ascending bitonic sort of one 64-bit key per thread over threads [0, PN).
template <int PN>
__device__ __forceinline__ unsigned long long block_bitonic_sort(unsigned long long v, unsigned long long* xch)
{
    const int t = threadIdx.x;
    bool bit[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) bit[i] = (t >> i) & 1;
    const int lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    const int bp32 = (lane ^ 32) << 2, bp63 = (lane ^ 63) << 2;
    int half = 0;
    bitonic_flip_levels<PN, 2>(v, xch, bit, bp32, bp63, half);
    return v;
}

// Rank order of island `island`'s fitness in buffer `buf` into order[0, Pi)
// (LDS), computed by the whole workgroup: the keys (fitness order key,
// slot) of the partition kernels, which are unique, so the order is the
// same stable order.  (Converged populations hold many equal fitness
// values, so the slot has to travel with the key.)  pow2(Pi) <= BLOCK; the
// network spans max(pow2(Pi), 64) threads; xch holds island_rank_bytes(Pi)
// bytes.  order is complete after the caller's next barrier.
// while_loading() runs between issuing the fitness loads and their first
// use, so loads it issues overlap theirs.
template <int BLOCK, typename Hook = NoHook>
__device__ __forceinline__ void island_rank_order(const InstDesc& d, int island, int buf, int* order, void* xch,
                                                  Hook while_loading = Hook())
{
    const int Pi = d.Pi, M = d.prob.num_obj;
    const int t = threadIdx.x;
    const float* fit = d.fit[buf] + (size_t)island * Pi * M;
    float f = 0.0f;
    if (t < Pi) f = fit[(size_t)t * M];
    while_loading();
    unsigned long long v = ~0ull;
    if (t < Pi) v = ((unsigned long long)fitness_order_key(f, d.prob.maximize) << 32) | (unsigned)t;
    unsigned long long* x = static_cast<unsigned long long*>(xch);
    if (Pi <= 64) v = block_bitonic_sort<64>(v, x);
    else if constexpr (BLOCK >= 128) {
        if (Pi <= 128) v = block_bitonic_sort<128>(v, x);
        else if constexpr (BLOCK >= 256) {
            if (Pi <= 256) v = block_bitonic_sort<256>(v, x);
            else if constexpr (BLOCK >= 512) {
                if (Pi <= 512) v = block_bitonic_sort<512>(v, x);
                else if constexpr (BLOCK >= 1024) v = block_bitonic_sort<1024>(v, x);
            }
        }
    }
    if (t < Pi) order[t] = (int)(unsigned)v;
}

// ---- island rank order by one group of a multi-child workgroup ----------
//
// The same network as block_bitonic_sort over one (fitness key, slot) key
// per thread of a group of threads (local index lt), in flip form.  `active`
// is uniform per wave: inactive waves skip the arithmetic and the in-wave
// exchanges but pass the same barriers (exchanges that cross waves go
// through LDS, one barrier each), so groups that rank nothing and groups
// that rank different islands of the same size stay in step.

template <int PN, int X>
__device__ __forceinline__ unsigned long long xor_exchange_g(unsigned long long v, unsigned long long* xch, int lt,
                                                             int bp, int& half, bool active)
{
    if constexpr (X >= kWave) {
        unsigned long long* b = xch + half * PN;
        half ^= 1;
        if (active && lt < PN) b[lt] = v;
        __syncthreads();
        return (active && lt < PN) ? b[lt ^ X] : v;
    } else {
        if (!active) return v;
        const unsigned lo = xor_lane_u32<X>((unsigned)v, bp);
        const unsigned hi = xor_lane_u32<X>((unsigned)(v >> 32), bp);
        return ((unsigned long long)hi << 32) | lo;
    }
}

template <int PN, int J>
__device__ __forceinline__ void half_cleaners_g(unsigned long long& v, unsigned long long* xch, const bool* bit, int lt,
                                                int bp32, int& half, bool active)
{
    const unsigned long long o = xor_exchange_g<PN, J>(v, xch, lt, bp32, half, active);
    if (active && ((o < v) != bit[__builtin_ctz(J)])) v = o;
    if constexpr (J > 1) half_cleaners_g<PN, J / 2>(v, xch, bit, lt, bp32, half, active);
}

template <int PN, int K>
__device__ __forceinline__ void flip_levels_g(unsigned long long& v, unsigned long long* xch, const bool* bit, int lt,
                                              int bp32, int bp63, int& half, bool active)
{
    const unsigned long long o = xor_exchange_g<PN, K - 1>(v, xch, lt, K - 1 == 63 ? bp63 : bp32, half, active);
    if (active && ((o < v) != bit[__builtin_ctz(K) - 1])) v = o;
    if constexpr (K >= 4) half_cleaners_g<PN, K / 4>(v, xch, bit, lt, bp32, half, active);
    if constexpr (K < PN) flip_levels_g<PN, 2 * K>(v, xch, bit, lt, bp32, bp63, half, active);
}

template <int PN>
__device__ __forceinline__ unsigned long long group_bitonic_sort(unsigned long long v, unsigned long long* xch, int lt,
                                                                 bool active)
{
    bool bit[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) bit[i] = (lt >> i) & 1;
    const int lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    const int bp32 = (lane ^ 32) << 2, bp63 = (lane ^ 63) << 2;
    int half = 0;
    flip_levels_g<PN, 2>(v, xch, bit, lt, bp32, bp63, half, active);
    return v;
}

// Rank order of `island` (fitness f of its slot lt already loaded, for lt <
// Pi) into order[0, Pi) when `active`; every group of the workgroup calls
// it with the same Pi.  pow2(Pi) <= BLOCK; xch holds island_rank_bytes(Pi)
// bytes per active group.  order is complete after the caller's next
// barrier.
template <int BLOCK>
__device__ __forceinline__ void group_rank_order(const InstDesc& d, float f, int lt, bool active, int* order,
                                                 unsigned long long* xch)
{
    const int Pi = d.Pi;
    unsigned long long v = ~0ull;
    if (active && lt < Pi) v = ((unsigned long long)fitness_order_key(f, d.prob.maximize) << 32) | (unsigned)lt;
    if (Pi <= 64) v = group_bitonic_sort<64>(v, xch, lt, active);
    else if constexpr (BLOCK >= 128) {
        if (Pi <= 128) v = group_bitonic_sort<128>(v, xch, lt, active);
        else if constexpr (BLOCK >= 256) {
            if (Pi <= 256) v = group_bitonic_sort<256>(v, xch, lt, active);
            else if constexpr (BLOCK >= 512) {
                if (Pi <= 512) v = group_bitonic_sort<512>(v, xch, lt, active);
                else if constexpr (BLOCK >= 1024) v = group_bitonic_sort<1024>(v, xch, lt, active);
            }
        }
    }
    if (active && lt < Pi) order[lt] = (int)(unsigned)v;
}

__device__ __forceinline__ float4 unit4(const U4& r)
{
    return make_float4(unit_from_u32(r.x), unit_from_u32(r.y), unit_from_u32(r.z), unit_from_u32(r.w));
}

// A child row is built in columns of W consecutive keys (W = 4: float4, or
// W = 2: float2 for workgroups with two keys per thread).  Key j of a row
// always takes word j mod 4 of Philox block j / 4 of its stream, whatever W.
template <int W>
struct KeyCol;
template <>
struct KeyCol<4> {
    using T = float4;
    __device__ static T units(const U4& r, int) { return unit4(r); }
    __device__ static T pick(const U4& r, int, const T& a, const T& b, float rho)
    {
        T x;
        x.x = unit_from_u32(r.x) < rho ? a.x : b.x;
        x.y = unit_from_u32(r.y) < rho ? a.y : b.y;
        x.z = unit_from_u32(r.z) < rho ? a.z : b.z;
        x.w = unit_from_u32(r.w) < rho ? a.w : b.w;
        return x;
    }
};
template <>
struct KeyCol<2> {
    using T = float2;
    __device__ static T units(const U4& r, int c)
    {
        return (c & 1) ? make_float2(unit_from_u32(r.z), unit_from_u32(r.w))
                       : make_float2(unit_from_u32(r.x), unit_from_u32(r.y));
    }
    __device__ static T pick(const U4& r, int c, const T& a, const T& b, float rho)
    {
        const uint32_t u0 = (c & 1) ? r.z : r.x, u1 = (c & 1) ? r.w : r.y;
        T x;
        x.x = unit_from_u32(u0) < rho ? a.x : b.x;
        x.y = unit_from_u32(u1) < rho ? a.y : b.y;
        return x;
    }
};

// Builds child slot s of island `island` for the step from parent buffer
// `cur` with step index t (generation 0 when init): the row goes to `out`
// (global) and, with TO_LDS, to `kl` (LDS; callers that decode from `mine`
// need no copy there), in columns of W keys.  Elite and migrant slots
// are copied with their fitness instead; then it returns false and nothing
// is left to decode.  `order` is the island's rank order (LDS-staged or
// global); `mig_order` the rank order of the preceding island on the ring,
// or null to read it from the partition's output.  mine (optional, columns
// <= BLOCK): the column ltid this thread built (or copied), left unchanged
// for threads with none.
template <int BLOCK, int W = 4, bool TO_LDS = true>
__device__ __forceinline__ bool build_child(const InstDesc& d, int island, int s, int cur, int init, int t,
                                            const int* order, const int* mig_order, typename KeyCol<W>::T* out,
                                            typename KeyCol<W>::T* kl, float* fit_out,
                                            typename KeyCol<W>::T* mine = nullptr)
{
    using Col = KeyCol<W>;
    using V = typename Col::T;
    const int Pi = d.Pi, Ei = d.Ei, Mi = d.Mi, I = d.islands;
    const int M = d.prob.num_obj, ld = d.ld;
    const int nq = (d.L + W - 1) / W;
    const int tid = ltid<BLOCK>();
    constexpr int kShift = W == 4 ? 0 : 1;  // Philox block of column c: c >> kShift
    if (init) {
        for (int c = tid; c < nq; c += BLOCK) {
            const V x = Col::units(rng_block(d.seed_lo, d.seed_hi, island, 0, s, TAG_INIT, c >> kShift), c);
            out[c] = x;
            if constexpr (TO_LDS) kl[c] = x;
            if (mine) *mine = x;
        }
        return true;
    }
    const uint32_t g = (uint32_t)t + 1u;
    const int k = d.mig_k;
    const bool migrate = I > 1 && k > 0 && d.mig_interval > 0 && s >= Pi - k && t > 0 && (t % d.mig_interval) == 0;
    int src = -1;
    if (migrate) {
        const int si = (island + I - 1) % I;
        src = si * Pi + (mig_order ? mig_order : d.perm[cur] + si * Pi)[s - (Pi - k)];
    } else if (s < Ei) {
        src = island * Pi + order[s];
    }
    if (src >= 0) {
        const V* in = reinterpret_cast<const V*>(d.keys[cur] + (size_t)src * ld);
        for (int c = tid; c < nq; c += BLOCK) {
            const V x = in[c];
            out[c] = x;
            if (mine) *mine = x;
        }
        if (tid < M) fit_out[tid] = d.fit[cur][(size_t)src * M + tid];
        return false;
    }
    if (s >= Pi - Mi) {
        for (int c = tid; c < nq; c += BLOCK) {
            const V x = Col::units(rng_block(d.seed_lo, d.seed_hi, island, g, s, TAG_MUTANT, c >> kShift), c);
            out[c] = x;
            if constexpr (TO_LDS) kl[c] = x;
            if (mine) *mine = x;
        }
        return true;
    }
    const U4 pr = rng_block(d.seed_lo, d.seed_hi, island, g, s, TAG_PARENTS, 0);
    const int ep = island * Pi + order[(int)(pr.x % (uint32_t)Ei)];
    const int np = island * Pi + order[Ei + (int)(pr.y % (uint32_t)(Pi - Ei))];
    const V* pe = reinterpret_cast<const V*>(d.keys[cur] + (size_t)ep * ld);
    const V* pn = reinterpret_cast<const V*>(d.keys[cur] + (size_t)np * ld);
    const float rho = d.rho;
    for (int c = tid; c < nq; c += BLOCK) {
        const U4 r = rng_block(d.seed_lo, d.seed_hi, island, g, s, TAG_COINS, c >> kShift);
        const V x = Col::pick(r, c, pe[c], pn[c], rho);
        out[c] = x;
        if constexpr (TO_LDS) kl[c] = x;
        if (mine) *mine = x;
    }
    return true;
}

}  // namespace brkga
