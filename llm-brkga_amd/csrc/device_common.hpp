// device_common.hpp - device-side building blocks shared by every kernel of
// the BRKGA engine (gfx950, wave64).
//
//   * Philox4x32-10 counter-based generator (Salmon et al., SC'11); replaces
//     the reference's per-thread cuRAND XORWOW states
//     (core/cuda_kernels.cuh:10-15,117-129).
//   * 24-bit key conversion, fixed-point term accumulation.
//   * Block-wide reductions / scans over wave64 shuffles + LDS.
//   * A stable LDS argsort of float keys (bucket by value, rank inside the
//     bucket by (key, index)), the decode primitive that replaces the
//     reference's per-thread bubble/selection sorts
//     (configs/tsp_config.hpp:398-417, configs/tspj_config.hpp:906-944).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "engine_types.hpp"

namespace brkga {

constexpr int kWave = 64;

struct U4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                            uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t lo0 = 0xD2511F53u * c0;
        const uint32_t hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return U4{c0, c1, c2, c3};
}

// Four consecutive words (word block `blk`) of stream (stream, gen, slot, tag).
__device__ __forceinline__ U4 rng_block(uint32_t seed_lo, uint32_t seed_hi, uint32_t stream, uint32_t gen,
                                        uint32_t slot, uint32_t tag, uint32_t blk)
{
    return philox4x32_10(stream, gen, slot, (tag << 24) | blk, seed_lo, seed_hi);
}

__device__ __forceinline__ float unit_from_u32(uint32_t u)
{
    return (float)(u >> 8) * 0x1p-24f;
}

__device__ __forceinline__ uint32_t u4_get(const U4& r, int i)
{
    return i == 0 ? r.x : (i == 1 ? r.y : (i == 2 ? r.z : r.w));
}

// FP32 term -> int64 fixed point (round to nearest even); exact power-of-two
// scaling, so the sum of terms is independent of the reduction order.
__device__ __forceinline__ long long fx_term(float t, float scale)
{
    return __float2ll_rn(__fmul_rn(t, scale));
}

__device__ __forceinline__ float fx_to_float(long long acc, float inv_scale)
{
    return __fmul_rn(__ll2float_rn(acc), inv_scale);
}

// Correctly rounded FP32 square root of a finite x >= 0.  The hardware
// estimate is within one ulp; the exact nearest float is found by comparing
// x with the squared midpoints to the neighbouring floats in FP64 (a 25-bit
// midpoint squares exactly in 53 bits, and no midpoint square equals a
// float, so there are no ties).
__device__ __forceinline__ float sqrt_rn(float x)
{
    float y = __builtin_sqrtf(x);
    if (!(x > 0.0f) || !(x < 3.0e38f)) return y;
    const double xd = (double)x;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const float up = __uint_as_float(__float_as_uint(y) + 1u);
        const double mu = 0.5 * ((double)y + (double)up);
        if (mu * mu < xd) {
            y = up;
            continue;
        }
        const float dn = __uint_as_float(__float_as_uint(y) - 1u);
        const double md = 0.5 * ((double)y + (double)dn);
        if (md * md > xd) {
            y = dn;
            continue;
        }
        break;
    }
    return y;
}

// Euclidean edge in FP32 with separately rounded products and sum
// (configs/tsp_config.hpp:366-368), correctly rounded square root.
__device__ __forceinline__ float edge_xy(float xa, float ya, float xb, float yb)
{
#pragma clang fp contract(off)
    const float dx = xa - xb;
    const float dy = ya - yb;
    const float sx = dx * dx;
    const float sy = dy * dy;
    return sqrt_rn(sx + sy);
}

// ------------------------------------------------------------ wave / block --

__device__ __forceinline__ long long wave_sum_i64(long long v)
{
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

__device__ __forceinline__ long long wave_max_i64(long long v)
{
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        long long u = __shfl_xor(v, o, kWave);
        v = u > v ? u : v;
    }
    return v;
}

__device__ __forceinline__ int wave_inclusive_scan_i32(int v)
{
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        int u = __shfl_up(v, o, kWave);
        if (lane >= o) v += u;
    }
    return v;
}

__device__ __forceinline__ long long wave_inclusive_scan_i64(long long v)
{
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        long long u = __shfl_up(v, o, kWave);
        if (lane >= o) v += u;
    }
    return v;
}

// Block sum; every thread gets the total. `red` holds BLOCK/64 int64.
template <int BLOCK>
__device__ __forceinline__ long long block_sum_i64(long long v, long long* red)
{
    constexpr int W = BLOCK / kWave;
    v = wave_sum_i64(v);
    if (W == 1) return v;
    const int wid = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) red[wid] = v;
    __syncthreads();
    long long t = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) t += red[i];
    __syncthreads();
    return t;
}

template <int BLOCK>
__device__ __forceinline__ long long block_max_i64(long long v, long long* red)
{
    constexpr int W = BLOCK / kWave;
    v = wave_max_i64(v);
    if (W == 1) return v;
    const int wid = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) red[wid] = v;
    __syncthreads();
    long long t = red[0];
#pragma unroll
    for (int i = 1; i < W; ++i) t = red[i] > t ? red[i] : t;
    __syncthreads();
    return t;
}

// Block exclusive scan of one int per thread; `red` holds BLOCK/64 ints.
template <int BLOCK>
__device__ __forceinline__ int block_exclusive_scan_i32(int v, int* red, int* total)
{
    constexpr int W = BLOCK / kWave;
    const int inc = wave_inclusive_scan_i32(v);
    const int wid = threadIdx.x / kWave;
    if (W == 1) {
        *total = __shfl(inc, kWave - 1, kWave);
        return inc - v;
    }
    if ((threadIdx.x & (kWave - 1)) == kWave - 1) red[wid] = inc;
    __syncthreads();
    int base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) {
        const int r = red[i];
        base += i < wid ? r : 0;
        tot += r;
    }
    __syncthreads();
    *total = tot;
    return base + inc - v;
}

template <int BLOCK>
__device__ __forceinline__ long long block_inclusive_scan_i64(long long v, long long* red)
{
    constexpr int W = BLOCK / kWave;
    const long long inc = wave_inclusive_scan_i64(v);
    if (W == 1) return inc;
    const int wid = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == kWave - 1) red[wid] = inc;
    __syncthreads();
    long long base = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) base += i < wid ? red[i] : 0;
    __syncthreads();
    return base + inc;
}

// ------------------------------------------------------------- argsort ----

// Bucket of a key for nb (power of two) value buckets over [0,1); any other
// value is clamped to the first/last bucket, which keeps the map monotone.
__device__ __forceinline__ int key_bucket(float k, int nb)
{
    if (!(k >= 0.0f)) return 0;
    if (k >= 1.0f) return nb - 1;
    int b = (int)__fmul_rn(k, (float)nb);
    return b < nb - 1 ? b : nb - 1;
}

struct SortScratch {
    int* cnt;    // [nb + 1]
    int* items;  // [n]
    int* red;    // [BLOCK/64]
};

// perm[p] = index of the element at position p of the stable ascending order
// of keys[0..n) (equal keys keep ascending index order, i.e. std::sort on
// pair<float,int>(key, index) as configs/tsp_config.hpp:209-216 does).
// keys and perm are LDS arrays; the result is ready after the final barrier.
template <int BLOCK>
__device__ void block_argsort(const float* keys, int n, int nb, const SortScratch& s, int* perm)
{
    const int tid = threadIdx.x;
    for (int i = tid; i <= nb; i += BLOCK) s.cnt[i] = 0;
    __syncthreads();
    for (int e = tid; e < n; e += BLOCK) perm[e] = atomicAdd(&s.cnt[key_bucket(keys[e], nb)], 1);
    __syncthreads();
    // exclusive scan of cnt[0..nb) in place, chunked per thread
    const int chunk = (nb + BLOCK - 1) / BLOCK;
    const int lo = tid * chunk;
    int local = 0;
    for (int i = 0; i < chunk; ++i)
        if (lo + i < nb) local += s.cnt[lo + i];
    int total;
    int run = block_exclusive_scan_i32<BLOCK>(local, s.red, &total);
    for (int i = 0; i < chunk; ++i) {
        if (lo + i < nb) {
            const int c = s.cnt[lo + i];
            s.cnt[lo + i] = run;
            run += c;
        }
    }
    if (tid == 0) s.cnt[nb] = n;
    __syncthreads();
    for (int e = tid; e < n; e += BLOCK) s.items[s.cnt[key_bucket(keys[e], nb)] + perm[e]] = e;
    __syncthreads();
    for (int e = tid; e < n; e += BLOCK) {
        const float k = keys[e];
        const int b = key_bucket(k, nb);
        const int st = s.cnt[b], en = s.cnt[b + 1];
        int r = st;
        if (en - st > 1) {
            for (int q = st; q < en; ++q) {
                const int j = s.items[q];
                const float kj = keys[j];
                r += (kj < k) || (kj == k && j < e);
            }
        }
        perm[r] = e;
    }
    __syncthreads();
}

__host__ __device__ inline int pow2_at_least(int v)
{
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Buckets used to argsort n keys: the power of two >= n, at least 64.
__host__ __device__ inline int sort_buckets(int n)
{
    const int p = pow2_at_least(n);
    return p < 64 ? 64 : p;
}

}  // namespace brkga
