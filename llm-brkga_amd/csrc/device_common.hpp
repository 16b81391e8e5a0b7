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

// Thread index within a group of BLOCK threads (a power of two).  The
// block-level helpers below work on such groups: in a kernel of BLOCK
// threads this is threadIdx.x (launch bounds let the compiler drop the
// mask); in a workgroup that runs several groups side by side (one child
// each), every group has its own LDS and all groups pass the same
// sequence of barriers.
template <int BLOCK>
__device__ __forceinline__ int ltid()
{
    return (int)threadIdx.x & (BLOCK - 1);
}

struct U4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                            uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;  // one v_mad_u64_u32 each
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
        const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
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

// Correctly rounded FP32 square root of a finite x >= 0.
//
// y = hardware estimate (within one ulp).  With u = ulp(y) and u' the ulp
// below y, the exact residual d = x - y^2 is a multiple of u^2 (x and y^2
// both are) below 2^25 u^2 in magnitude, so fl(d) = fma(-y, y, x) is exact
// whenever |d| < 2^24 u^2 and keeps its order against y*u (< 2^24 u^2,
// exact) otherwise.  sqrt(x) lies above the midpoint y + u/2 iff d > y*u,
// and below y - u'/2 iff d <= -y*u' (midpoint squares are never floats, so
// there are no ties).  V_SQRT_F32 is accurate to one ulp, so one correction
// suffices; inputs whose ulp would be subnormal take the FP64 midpoint test.
// Zeros, negatives, infinities and NaN: the hardware result is already the
// IEEE one (the correction leaves infinities and NaN unchanged; zeros are
// selected back, since a zero y has no ulp to test against).
__device__ __forceinline__ float sqrt_rn(float x)
{
    float y = __builtin_amdgcn_sqrtf(x);
    if (__builtin_expect(x > 0.0f && x < 1.0e-30f, 0)) {
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
            if (md * md > xd) y = dn;
        }
        return y;
    }
    const uint32_t b = __float_as_uint(y);
    const float u = __uint_as_float((b & 0x7F800000u) - (23u << 23));  // ulp(y)
    const float ud = (b & 0x007FFFFFu) ? u : __uint_as_float((b & 0x7F800000u) - (24u << 23));
    const float d = __builtin_fmaf(-y, y, x);
    const float yc = d > y * u ? __uint_as_float(b + 1u) : (d <= -(y * ud) ? __uint_as_float(b - 1u) : y);
    return x == 0.0f ? x : yc;
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

// Deterministic cosine of an FP32 argument, correctly rounded in all but
// rare cases: Cody-Waite reduction by pi/2 in FP64 (pi/2 split into a
// 33-bit head and a tail, so k * head is exact for |k| < 2^20), then the
// Taylor series of sin / cos on [-pi/4, pi/4] to degree 17 / 16 (truncation
// below 3e-18), all in separately rounded FP64 operations.  The oracle
// evaluates the identical sequence on the CPU, so both agree bit for bit.
// |y| < 2^20 * pi/2; larger arguments return cos(0) of an inexact
// reduction and are not used by any decoder (|y| <= 2 pi * max bound).
__host__ __device__ inline float cos_det(float y)
{
#pragma clang fp contract(off)
    const double kTwoOverPi = 0.63661977236758134308;
    const double kPio2Hi = 1.57079632673412561417e+00;
    const double kPio2Lo = 6.07710050650619224932e-11;
    const double yd = (double)y;
    const double k = __builtin_rint(yd * kTwoOverPi);
    const double r = (yd - k * kPio2Hi) - k * kPio2Lo;
    const double z = r * r;
    double c = 1.0 / 20922789888000.0;        // 1/16!
    c = c * z - 1.0 / 87178291200.0;          // 1/14!
    c = c * z + 1.0 / 479001600.0;            // 1/12!
    c = c * z - 1.0 / 3628800.0;              // 1/10!
    c = c * z + 1.0 / 40320.0;                // 1/8!
    c = c * z - 1.0 / 720.0;                  // 1/6!
    c = c * z + 1.0 / 24.0;                   // 1/4!
    c = c * z - 0.5;
    c = c * z + 1.0;
    double sn = 1.0 / 355687428096000.0;      // 1/17!
    sn = sn * z - 1.0 / 1307674368000.0;      // 1/15!
    sn = sn * z + 1.0 / 6227020800.0;         // 1/13!
    sn = sn * z - 1.0 / 39916800.0;           // 1/11!
    sn = sn * z + 1.0 / 362880.0;             // 1/9!
    sn = sn * z - 1.0 / 5040.0;               // 1/7!
    sn = sn * z + 1.0 / 120.0;                // 1/5!
    sn = sn * z - 1.0 / 6.0;                  // 1/3!
    sn = (sn * z) * r + r;
    const int q = (int)((long long)k & 3);
    const double v = q == 0 ? c : (q == 1 ? -sn : (q == 2 ? -c : sn));
    return (float)v;
}

// Decision variable of a continuous gene: gene * (hi - lo) + lo in FP32
// (configs/continuous_opt_gpu_config.hpp:136-137).
__host__ __device__ inline float cont_var(float gene, float lo, float hi)
{
#pragma clang fp contract(off)
    const float w = hi - lo;
    const float t = gene * w;
    return t + lo;
}

// Machine of a job gene: int(gene * M) clamped to [0, M)
// (configs/job_scheduling_gpu_config.hpp:145-146).
__host__ __device__ inline int job_machine(float gene, int M)
{
    const float v = gene * (float)M;
    if (!(v >= 0.0f)) return 0;
    if (v >= (float)M) return M - 1;
    const int m = (int)v;
    return m < M ? m : M - 1;
}

// ------------------------------------------------------------ wave / block --

// Wave-wide scans and reductions on DPP row shifts and row broadcasts (no
// LDS round trips).  Within each 16-lane row a Hillis-Steele scan with
// row_shr 1, 2, 4, 8; then row_bcast:15 adds row 0's (row 2's) last lane to
// row 1 (row 3) and row_bcast:31 adds lane 31 to rows 2 and 3.  A lane whose
// source is outside its row, or in a row the mask leaves out, reads `old`.
// Every lane of the wave must be active.
enum : int {
    kDppShr1 = 0x111,
    kDppShr2 = 0x112,
    kDppShr4 = 0x114,
    kDppShr8 = 0x118,
    kDppBcast15 = 0x142,
    kDppBcast31 = 0x143
};

template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ unsigned dpp_u32(unsigned old, unsigned v)
{
    return (unsigned)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, ROW_MASK, 0xf, false);
}

// 64-bit lane value from the lane CTRL names (both halves from one lane).
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long old, unsigned long long v)
{
    const unsigned lo = dpp_u32<CTRL, ROW_MASK>((unsigned)old, (unsigned)v);
    const unsigned hi = dpp_u32<CTRL, ROW_MASK>((unsigned)(old >> 32), (unsigned)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

__device__ __forceinline__ long long wave_read_last(long long v)
{
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, kWave - 1);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)((unsigned long long)v >> 32), kWave - 1);
    return (long long)(((unsigned long long)hi << 32) | lo);
}

__device__ __forceinline__ int wave_inclusive_scan_i32(int v)
{
    unsigned x = (unsigned)v;
    x += dpp_u32<kDppShr1>(0u, x);
    x += dpp_u32<kDppShr2>(0u, x);
    x += dpp_u32<kDppShr4>(0u, x);
    x += dpp_u32<kDppShr8>(0u, x);
    x += dpp_u32<kDppBcast15, 0xa>(0u, x);
    x += dpp_u32<kDppBcast31, 0xc>(0u, x);
    return (int)x;
}

__device__ __forceinline__ long long wave_inclusive_scan_i64(long long v)
{
    unsigned long long x = (unsigned long long)v;
    x += dpp_u64<kDppShr1>(0ull, x);
    x += dpp_u64<kDppShr2>(0ull, x);
    x += dpp_u64<kDppShr4>(0ull, x);
    x += dpp_u64<kDppShr8>(0ull, x);
    x += dpp_u64<kDppBcast15, 0xa>(0ull, x);
    x += dpp_u64<kDppBcast31, 0xc>(0ull, x);
    return (long long)x;
}

// Sum over the wave, in every lane.
__device__ __forceinline__ long long wave_sum_i64(long long v)
{
    return wave_read_last(wave_inclusive_scan_i64(v));
}

// Maximum over the wave, in every lane (a lane with no source keeps its own
// value, the identity of max).
__device__ __forceinline__ long long wave_max_i64(long long v)
{
    long long x = v;
    auto step = [&](unsigned long long u) { x = (long long)u > x ? (long long)u : x; };
    step(dpp_u64<kDppShr1>((unsigned long long)x, (unsigned long long)x));
    step(dpp_u64<kDppShr2>((unsigned long long)x, (unsigned long long)x));
    step(dpp_u64<kDppShr4>((unsigned long long)x, (unsigned long long)x));
    step(dpp_u64<kDppShr8>((unsigned long long)x, (unsigned long long)x));
    step(dpp_u64<kDppBcast15, 0xa>((unsigned long long)x, (unsigned long long)x));
    step(dpp_u64<kDppBcast31, 0xc>((unsigned long long)x, (unsigned long long)x));
    return wave_read_last(x);
}

// Block sum; every thread gets the total. `red` holds BLOCK/64 int64.
template <int BLOCK>
__device__ __forceinline__ long long block_sum_i64(long long v, long long* red)
{
    constexpr int W = BLOCK / kWave;
    v = wave_sum_i64(v);
    if (W == 1) return v;
    const int wid = ltid<BLOCK>() / kWave;
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
    const int wid = ltid<BLOCK>() / kWave;
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
    const int wid = ltid<BLOCK>() / kWave;
    if (W == 1) {
        *total = __builtin_amdgcn_readlane(inc, kWave - 1);
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
    const int wid = ltid<BLOCK>() / kWave;
    if ((threadIdx.x & (kWave - 1)) == kWave - 1) red[wid] = inc;
    __syncthreads();
    long long base = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) base += i < wid ? red[i] : 0;
    __syncthreads();
    return base + inc;
}

// --------------------------------------------------------- fixed point ----

// A thread's sum of at most 16 terms llrint(t * 2^S), |t * 2^S| <= 2^48,
// held exactly in FP64: each term rounds to an integer in FP32 (a float of
// magnitude >= 2^23 already is one, a smaller one rounds to an integer below
// 2^24, which FP32 holds), and FP64 adds such integers exactly while the sum
// stays below 2^53 (16 x 2^48 = 2^52).
struct FxAcc {
    double s = 0.0;
};

__device__ __forceinline__ void fx_add(FxAcc& a, float t, float scale)
{
    a.s += (double)__builtin_rintf(t * scale);
}

__device__ __forceinline__ long long fx_value(const FxAcc& a)
{
    return (long long)a.s;
}

// ------------------------------------------------------------- argsort ----

// Bucket of a key for nb (power of two) value buckets over [0,1); any other
// value is clamped to the first/last bucket (NaN to the last, where its
// order key sorts after every number), which keeps the map monotone in the
// order of key_order.
__device__ __forceinline__ int key_bucket(float k, int nb)
{
    // clamped in FP32 to [0, nb - 1] first (fmaxf of NaN and 0 is 0, so the
    // conversion always gets a value in range); NaN goes to the last bucket
    const float x = fminf(fmaxf(__fmul_rn(k, (float)nb), 0.0f), (float)(nb - 1));
    return k != k ? nb - 1 : (int)x;
}

// Order-preserving FP32 -> u32 (-0 == +0, NaN last).
__device__ __forceinline__ uint32_t key_order(float k)
{
    // k + 0 turns -0 into +0 (IEEE addition) and keeps every other value;
    // negative numbers are flipped whole, others get the sign bit
    const uint32_t u = __float_as_uint(__fadd_rn(k, 0.0f));
    const uint32_t o = u ^ ((uint32_t)((int32_t)u >> 31) | 0x80000000u);
    return k != k ? 0xFFFFFFFFu : o;
}

struct SortScratch {
    int* cnt;                   // [nb + 1] bucket counts, then bucket starts;
                                // cnt[nb + 1]: some bucket is too long to scan
    unsigned long long* items;  // [n] (key order << 32 | index) in bucket order
    long long* red;             // [17] per-wave partials; red[16] = block sum
    int* flag = nullptr;        // long-bucket flag shared by every group of a
                                // multi-child workgroup (null: cnt[nb + 1])
};

// Clears the bucket counters (nb + 4 of them, 16-byte aligned) and the
// block-sum slot; a barrier must follow.
template <int BLOCK>
__device__ __forceinline__ void argsort_clear(const SortScratch& s, int nb)
{
    int4* c4 = reinterpret_cast<int4*>(s.cnt);
    for (int i = ltid<BLOCK>(); i <= (nb >> 2); i += BLOCK) c4[i] = make_int4(0, 0, 0, 0);
    if (ltid<BLOCK>() == 0) s.red[16] = 0;
}

// Exclusive scan of cnt[0..nb) in place (bucket starts), cnt[nb] = n.
template <int BLOCK>
__device__ __forceinline__ void scan_bucket_counts(int* cnt, int nb, int n, int* red)
{
    constexpr int W = BLOCK / kWave;
    const int tid = ltid<BLOCK>();
    const int per = nb >= BLOCK ? nb / BLOCK : 1;
    const int lo = tid * per;
    int local = 0;
    if (lo < nb) {
        if ((per & 3) == 0) {
            const int4* c4 = reinterpret_cast<const int4*>(cnt + lo);
            for (int k = 0; k < (per >> 2); ++k) {
                const int4 v = c4[k];
                local += v.x + v.y + v.z + v.w;
            }
        } else {
            for (int k = 0; k < per; ++k) local += cnt[lo + k];
        }
    }
    const int incl = wave_inclusive_scan_i32(local);
    int base = 0;
    if (W > 1) {
        const int wid = tid / kWave;
        if ((tid & (kWave - 1)) == kWave - 1) red[wid] = incl;
        __syncthreads();
#pragma unroll
        for (int w = 0; w < W; ++w) base += w < wid ? red[w] : 0;
    }
    int run = base + incl - local;
    if (lo < nb) {
        if ((per & 3) == 0) {
            int4* c4 = reinterpret_cast<int4*>(cnt + lo);
            for (int k = 0; k < (per >> 2); ++k) {
                const int4 v = c4[k];
                int4 o;
                o.x = run;
                o.y = run + v.x;
                o.z = o.y + v.y;
                o.w = o.z + v.z;
                run = o.w + v.w;
                c4[k] = o;
            }
        } else {
            for (int k = 0; k < per; ++k) {
                const int c = cnt[lo + k];
                cnt[lo + k] = run;
                run += c;
            }
        }
    }
    if (tid == 0) cnt[nb] = n;
    __syncthreads();
}

// Buckets longer than this are not ranked by scanning: a row with that many
// keys in one bucket (equal keys, keys outside [0, 1), written by a caller)
// is sorted instead (block_sort_items), so a row costs O(n log^2 n) however
// its keys are distributed.  Philox keys fill about half a bucket each, and
// a bucket over 64 keys has probability below 1e-30 per row.
constexpr int kBucketScanLimit = 64;

// Ascending sort of the n unique 64-bit items a[0..n) in LDS by the whole
// workgroup: a bitonic network in flip form (every compare-exchange puts the
// smaller item at the lower index) over pow2(n) positions, positions >= n
// standing for +infinity, so a compare-exchange that reaches one leaves both
// items in place and is skipped.  One barrier per stage; ends with one.
template <int BLOCK>
__device__ void block_sort_items(unsigned long long* a, int n)
{
    int pn = 1;
    while (pn < n) pn <<= 1;
    const int half = pn >> 1;
    for (int k = 2; k <= pn; k <<= 1) {
        for (int j = k >> 1; j >= 1; j >>= 1) {
            const int lj = __builtin_ctz(j);
            const bool flip = j == (k >> 1);
            for (int c = ltid<BLOCK>(); c < half; c += BLOCK) {
                const int base = (c >> lj) << (lj + 1);
                const int off = c & (j - 1);
                const int lo = base + off;
                const int hi = flip ? base + 2 * j - 1 - off : lo + j;
                if (hi < n) {
                    const unsigned long long x = a[lo], y = a[hi];
                    if (y < x) {
                        a[lo] = y;
                        a[hi] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Index of item v in the sorted unique items a[0..n).
__device__ __forceinline__ int sorted_items_find(const unsigned long long* a, int n, unsigned long long v)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Sorted positions of keys[0..n) in the stable ascending order (equal keys
// keep ascending index order, i.e. std::sort on pair<float,int>(key, index)
// as configs/tsp_config.hpp:209-216 does): on return pos[i] is the position
// of key tid + i * BLOCK, and perm[pos[i]] = tid + i * BLOCK when perm is
// given.  Thread tid holds the keys tid + i * BLOCK, i < ITEMS: n <= ITEMS *
// BLOCK.
//
// Bucket sort: keys are counted into nb value buckets (LDS atomics), the
// counts are scanned into bucket ranges and every key drops its
// (order key, index) item into its range.  Items are unique, so each key's
// final position is its bucket start plus the number of smaller items in
// its bucket (buckets hold about half a key on average); every key ranks
// itself, with no per-bucket leader and no divergent sorting networks.
// Each thread keeps its keys' order keys and bucket ranges in registers, so
// the ranking reads only the other items of the bucket.  keys are read once,
// before the first barrier, so perm may overlay them.  No barrier follows
// the ranking: callers synchronise before they read perm or reuse s.items.
// CLEARED: the caller already ran argsort_clear and a barrier.  pre_rank()
// runs between the placement and the ranking (loads issued there overlap
// the ranking without holding registers through the whole sort).
struct NoHook {
    __device__ void operator()() const {}
};

// Element index of a thread's i-th key: tid + i * BLOCK (strided), or
// tid * ITEMS + i (CONTIG: a thread's keys are consecutive).
template <int BLOCK, int ITEMS, bool CONTIG>
__device__ __forceinline__ int rank_elem(int tid, int i)
{
    return CONTIG ? tid * ITEMS + i : tid + i * BLOCK;
}

// block_rank with the keys supplied by key_at(i, e) (the thread's i-th key,
// element e) and the element mapping CONTIG; positions and perm follow the
// same mapping.
// KEYS_STAY: key_at(i, e) still returns the key during the placement and
// the ranking (the keys are not overwritten, perm is null); the order keys
// are then recomputed there instead of held in registers, for the widest
// decoders (16 keys per thread).
// BATCH: each thread issues the LDS operations of a phase together and
// waits once, which shortens a workgroup's chain of dependent LDS round
// trips at the cost of registers: for latency-bound launches (one
// population, a few workgroups per CU).  Otherwise each key's accesses are
// waited for in turn, which keeps the registers, and so the occupancy, of
// throughput-bound launches (batched and served fleets).
template <int BLOCK, bool CLEARED, int ITEMS, bool CONTIG, typename KeyAt, typename Hook = NoHook,
          bool KEYS_STAY = false, bool BATCH = false>
__device__ __forceinline__ void block_rank_by(KeyAt key_at, int n, int nb, const SortScratch& s,
                                              int (&pos)[ITEMS], int* perm = nullptr, Hook pre_rank = Hook())
{
    const int tid = ltid<BLOCK>();
    if (!CLEARED) {
        argsort_clear<BLOCK>(s, nb);
        __syncthreads();
    }
    // rng[i]: the bucket, then (bucket start | bucket size << 16); both are
    // below 2^15 since n <= ITEMS * BLOCK <= kSortItems * 1024
    // With 16 keys per thread the slot a key got in its bucket travels in
    // the upper half of rng (bucket < 2^11, slot < 2^14), which keeps the
    // widest decoders within the registers of a 1024-thread workgroup.
    constexpr bool kPack = ITEMS >= 16;
    // kSerial: one key at a time, each LDS access waited for where its
    // result is used (also the widest forms, which have no registers to
    // spare); otherwise a thread's LDS accesses of a phase go together.
    constexpr bool kSerial = !BATCH || kPack || (BLOCK >= 1024 && ITEMS >= 8);
    int rng[ITEMS], slot[kPack ? 1 : ITEMS];
    constexpr bool kReload = kPack && KEYS_STAY;
    uint32_t ok[kReload ? 1 : ITEMS];
    auto order_of = [&](int i, int e) -> uint32_t {
        if constexpr (kReload) return key_order(key_at(i, e));
        else return ok[i];
    };
    if constexpr (kSerial) {
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const int e = rank_elem<BLOCK, ITEMS, CONTIG>(tid, i);
            if (e < n) {
                const float k = key_at(i, e);
                if constexpr (!kReload) ok[i] = key_order(k);
                const int b = key_bucket(k, nb);
                const int sl = atomicAdd(&s.cnt[b], 1);
                if constexpr (kPack) {
                    rng[i] = b | (sl << 16);
                } else {
                    rng[i] = b;
                    slot[i] = sl;
                }
            }
        }
    } else {
        // Each thread issues a phase's LDS operations back to back and waits
        // once (no result is used inside the branch that issues it): the
        // keys are read first, then every key takes its slot.  An element
        // past n keeps bucket nb + 2, a cleared spare counter the placement
        // reads and the scan never touches.
        float kf[ITEMS];
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const int e = rank_elem<BLOCK, ITEMS, CONTIG>(tid, i);
            kf[i] = e < n ? key_at(i, e) : 0.0f;
        }
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const int e = rank_elem<BLOCK, ITEMS, CONTIG>(tid, i);
            ok[i] = key_order(kf[i]);
            rng[i] = e < n ? key_bucket(kf[i], nb) : nb + 2;
            // an element past n counts itself in the spare counter, which
            // nothing reads as a bucket (no per-key branch)
            slot[i] = atomicAdd(&s.cnt[rng[i]], 1);
        }
    }
    __syncthreads();
    scan_bucket_counts<BLOCK>(s.cnt, nb, n, reinterpret_cast<int*>(s.red));
    int longest = 0;
    if constexpr (kSerial) {
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const int e = rank_elem<BLOCK, ITEMS, CONTIG>(tid, i);
            if (e < n) {
                const int b = kPack ? (rng[i] & 0xFFFF) : rng[i];
                const int sl = kPack ? (rng[i] >> 16) : slot[kPack ? 0 : i];
                const int st = s.cnt[b];
                const int en = s.cnt[b + 1];
                s.items[st + sl] = ((unsigned long long)order_of(i, e) << 32) | (unsigned)e;
                rng[i] = st | ((en - st) << 16);
                longest = en - st > longest ? en - st : longest;
            }
        }
    } else {
        // every bucket start and end is read before the first item is
        // placed; elements past n read the spare counters and keep rng = 0
        int st[ITEMS], en[ITEMS];
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            st[i] = s.cnt[rng[i]];
            en[i] = s.cnt[rng[i] + 1];
        }
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const int e = rank_elem<BLOCK, ITEMS, CONTIG>(tid, i);
            const bool valid = e < n;
            // an element past n writes the first spare item, which the ranking
            // reads only as padding past a bucket's end
            s.items[valid ? st[i] + slot[i] : n] = ((unsigned long long)ok[i] << 32) | (unsigned)e;
            const int sz = en[i] - st[i];
            rng[i] = valid ? (st[i] | (sz << 16)) : 0;
            longest = valid && sz > longest ? sz : longest;
        }
    }
    int* const long_flag = s.flag ? s.flag : s.cnt + nb + 1;
    if (longest > kBucketScanLimit) *long_flag = 1;
    __syncthreads();
    pre_rank();
    if (__builtin_amdgcn_readfirstlane(*long_flag)) {
        // a bucket too long to scan: sort every item, then each key finds
        // its own (unique) item
        block_sort_items<BLOCK>(s.items, n);
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const int e = rank_elem<BLOCK, ITEMS, CONTIG>(tid, i);
            if (e < n) {
                const int p = sorted_items_find(s.items, n, ((unsigned long long)order_of(i, e) << 32) | (unsigned)e);
                pos[i] = p;
                if (perm) perm[p] = e;
            }
        }
        return;
    }
    if constexpr (kSerial) {
        // A key alone in its bucket reads nothing; otherwise the first four
        // items of the bucket are read as independent loads and longer
        // buckets finish in a loop.  A key compared with itself counts
        // nothing.
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const int e = rank_elem<BLOCK, ITEMS, CONTIG>(tid, i);
            if (e < n) {
                const int st = rng[i] & 0xFFFF;
                const int sz = rng[i] >> 16;
                int r = 0;
                if (sz > 1) {
                    const unsigned long long mine = ((unsigned long long)order_of(i, e) << 32) | (unsigned)e;
                    const unsigned long long* b = s.items + st;
                    const unsigned long long x0 = b[0];
                    const unsigned long long x1 = b[1];
                    const unsigned long long x2 = sz > 2 ? b[2] : ~0ull;
                    const unsigned long long x3 = sz > 3 ? b[3] : ~0ull;
                    r = (x0 < mine) + (x1 < mine) + (x2 < mine) + (x3 < mine);
                    for (int q = 4; q < sz; ++q) r += b[q] < mine ? 1 : 0;
                }
                pos[i] = st + r;
                if (perm) perm[st + r] = e;
            }
        }
        return;
    }
    // A key alone in its bucket reads nothing; every other key reads the
    // first four items from its bucket start, the loads of four keys issued
    // before any is used (the items array has four spare entries, so the
    // reads stay inside it).  An item past the bucket end counts nothing,
    // nor does the key compared with itself.  Buckets longer than four
    // finish in a loop afterwards.
    constexpr int G = ITEMS < 4 ? ITEMS : 4;
    int r[ITEMS];
#pragma unroll
    for (int i0 = 0; i0 < ITEMS; i0 += G) {
        unsigned long long x[G][4];
#pragma unroll
        for (int j = 0; j < G; ++j) {
            const unsigned long long* b = s.items + (rng[i0 + j] & 0xFFFF);
            // unread slots hold the largest item, which counts nothing
#pragma unroll
            for (int q = 0; q < 4; ++q) x[j][q] = ~0ull;
            if ((rng[i0 + j] >> 16) > 1) {
#pragma unroll
                for (int q = 0; q < 4; ++q) x[j][q] = b[q];
            }
        }
#pragma unroll
        for (int j = 0; j < G; ++j) {
            const int i = i0 + j;
            const int e = rank_elem<BLOCK, ITEMS, CONTIG>(tid, i);
            const int sz = rng[i] >> 16;
            const unsigned long long mine = ((unsigned long long)ok[i] << 32) | (unsigned)e;
            int c = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) c += (q < sz && x[j][q] < mine) ? 1 : 0;
            r[i] = c;
        }
    }
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const int e = rank_elem<BLOCK, ITEMS, CONTIG>(tid, i);
        const int st = rng[i] & 0xFFFF;
        const int sz = rng[i] >> 16;
        if (sz > 4) {
            const unsigned long long mine = ((unsigned long long)ok[i] << 32) | (unsigned)e;
            for (int q = 4; q < sz; ++q) r[i] += s.items[st + q] < mine ? 1 : 0;
        }
        pos[i] = st + r[i];
        if (perm && e < n) perm[st + r[i]] = e;
    }
}

template <int BLOCK, bool CLEARED = false, int ITEMS = kSortItems, typename Hook = NoHook, bool KEYS_STAY = false>
__device__ __forceinline__ void block_rank(const float* keys, int n, int nb, const SortScratch& s,
                                           int (&pos)[ITEMS], int* perm = nullptr, Hook pre_rank = Hook())
{
    auto key_at = [&](int, int e) { return keys[e]; };
    block_rank_by<BLOCK, CLEARED, ITEMS, false, decltype(key_at), Hook, KEYS_STAY>(key_at, n, nb, s, pos, perm,
                                                                                  pre_rank);
}

// perm[p] = index of the element at position p of the stable ascending
// order of keys[0..n) (block_rank).  keys and perm live in LDS and may
// overlay each other.  The result is ready after the final barrier.
template <int BLOCK, bool CLEARED = false, int ITEMS = kSortItems>
__device__ __forceinline__ void block_argsort(const float* keys, int n, int nb, const SortScratch& s, int* perm)
{
    int pos[ITEMS];
    block_rank<BLOCK, CLEARED, ITEMS>(keys, n, nb, s, pos, perm);
    __syncthreads();
}

// Block sum of one int64 per thread; the result is valid in thread 0.
// Uses s.red[16], cleared by argsort_clear.
template <int BLOCK>
__device__ __forceinline__ long long block_sum_fx(long long v, const SortScratch& s)
{
    v = wave_sum_i64(v);
    if (BLOCK == kWave) return v;
    if ((threadIdx.x & (kWave - 1)) == 0)
        atomicAdd(reinterpret_cast<unsigned long long*>(&s.red[16]), (unsigned long long)v);
    __syncthreads();
    return s.red[16];
}

__host__ __device__ inline int pow2_at_least(int v)
{
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Buckets used to argsort n keys: the power of two >= 2n, at least 64 and
// at most 2048.  With about half a key per bucket most keys rank without
// reading any other item.
__host__ __device__ inline int sort_buckets(int n)
{
    const int p = pow2_at_least(2 * n);
    return p < 64 ? 64 : (p > 2048 ? 2048 : p);
}

}  // namespace brkga
