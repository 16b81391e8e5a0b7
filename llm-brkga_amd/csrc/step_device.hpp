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

__device__ __forceinline__ float4 unit4(const U4& r)
{
    return make_float4(unit_from_u32(r.x), unit_from_u32(r.y), unit_from_u32(r.z), unit_from_u32(r.w));
}

// Builds child slot s of island `island` for the step from parent buffer
// `cur` with step index t (generation 0 when init): the row goes to `out`
// (global) and `kl` (LDS).  Elite and migrant slots are copied with their
// fitness instead; then it returns false and nothing is left to decode.
// `order` is the island's rank order (LDS-staged or global).
template <int BLOCK>
__device__ __forceinline__ bool build_child(const InstDesc& d, int island, int s, int cur, int init, int t,
                                            const int* order, float4* out, float4* kl, float* fit_out)
{
    const int Pi = d.Pi, Ei = d.Ei, Mi = d.Mi, I = d.islands;
    const int M = d.prob.num_obj, ld = d.ld;
    const int nq = (d.L + 3) >> 2;
    if (init) {
        for (int c = threadIdx.x; c < nq; c += BLOCK) {
            const float4 x = unit4(rng_block(d.seed_lo, d.seed_hi, island, 0, s, TAG_INIT, c));
            out[c] = x;
            kl[c] = x;
        }
        return true;
    }
    const uint32_t g = (uint32_t)t + 1u;
    const int k = d.mig_k;
    const bool migrate = I > 1 && k > 0 && d.mig_interval > 0 && s >= Pi - k && t > 0 && (t % d.mig_interval) == 0;
    int src = -1;
    if (migrate) {
        const int si = (island + I - 1) % I;
        src = si * Pi + d.perm[cur][si * Pi + (s - (Pi - k))];
    } else if (s < Ei) {
        src = island * Pi + order[s];
    }
    if (src >= 0) {
        const float4* in = reinterpret_cast<const float4*>(d.keys[cur] + (size_t)src * ld);
        for (int c = threadIdx.x; c < nq; c += BLOCK) out[c] = in[c];
        if (threadIdx.x < M) fit_out[threadIdx.x] = d.fit[cur][(size_t)src * M + threadIdx.x];
        return false;
    }
    if (s >= Pi - Mi) {
        for (int c = threadIdx.x; c < nq; c += BLOCK) {
            const float4 x = unit4(rng_block(d.seed_lo, d.seed_hi, island, g, s, TAG_MUTANT, c));
            out[c] = x;
            kl[c] = x;
        }
        return true;
    }
    const U4 pr = rng_block(d.seed_lo, d.seed_hi, island, g, s, TAG_PARENTS, 0);
    const int ep = island * Pi + order[(int)(pr.x % (uint32_t)Ei)];
    const int np = island * Pi + order[Ei + (int)(pr.y % (uint32_t)(Pi - Ei))];
    const float4* pe = reinterpret_cast<const float4*>(d.keys[cur] + (size_t)ep * ld);
    const float4* pn = reinterpret_cast<const float4*>(d.keys[cur] + (size_t)np * ld);
    const float rho = d.rho;
    for (int c = threadIdx.x; c < nq; c += BLOCK) {
        const U4 r = rng_block(d.seed_lo, d.seed_hi, island, g, s, TAG_COINS, c);
        const float4 a = pe[c], b = pn[c];
        float4 x;
        x.x = unit_from_u32(r.x) < rho ? a.x : b.x;
        x.y = unit_from_u32(r.y) < rho ? a.y : b.y;
        x.z = unit_from_u32(r.z) < rho ? a.z : b.z;
        x.w = unit_from_u32(r.w) < rho ? a.w : b.w;
        out[c] = x;
        kl[c] = x;
    }
    return true;
}

}  // namespace brkga
