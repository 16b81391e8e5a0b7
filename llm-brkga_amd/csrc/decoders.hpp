// decoders.hpp - per-individual decode + fitness, one workgroup per
// individual, keys already staged in LDS.
//
// Reference decoders replaced here (paths relative to the reference's brkga/):
//   TSP          configs/tsp_config.hpp:180-225 (+ GPU kernel 384-437, which
//                truncates at 128 cities; not reproduced)
//   knapsack     configs/knapsack_config.hpp:43-75, knapsack_gpu_config.hpp:242-268
//   TSPJ         configs/tspj_config.hpp:555-642 (+ GPU kernel 874-982)
//   bi-obj TSP   configs/multi_tsp_config.hpp:54-103
// Every edge / term is FP32 as in the reference; sums are order-independent
// int64 fixed point converted once to FP32, so the result does not depend on
// the block size.
#pragma once

#include "device_common.hpp"

namespace brkga {

// The TSP length over coordinates once the keys are ranked: each key writes
// its city's coordinates c[i] at its tour position pos[i] (element mapping
// CONTIG as in block_rank_by), then the edges are summed from consecutive
// positions.  Tour positions go to the argsort counters, dead once the keys
// are placed, when cnt[0, nb) holds n coordinate pairs (nb >= 2n); then no
// barrier is needed against the ranking's reads of the items, and cnt[nb + 1]
// (the long-bucket flag every thread reads after the placement) is never
// overwritten.  EARLY: c[] was loaded during the ranking; otherwise each
// coordinate pair is read from xy just before it is written (no array of
// ITEMS pairs held across the ranking: the widest decoders stay within
// their registers).
template <int BLOCK, int ITEMS, bool CONTIG, bool EARLY, bool BATCH = false>
__device__ __forceinline__ void tsp_coords_tail(const ProblemDesc& p, char* lds, const DecodeLayout& lay,
                                                const SortScratch& ss, int nb, const float2 (&c)[ITEMS],
                                                const int (&pos)[ITEMS], float* fit_out, int* decoded,
                                                const float2* xy = nullptr)
{
    const int tid = ltid<BLOCK>();
    const int n = p.n;
    const bool in_cnt = nb >= 2 * n;
    float2* txy = reinterpret_cast<float2*>(in_cnt ? reinterpret_cast<char*>(ss.cnt) : lds + lay.items_off);
    if (!in_cnt) __syncthreads();  // every ranking read of the items is done
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const int e = rank_elem<BLOCK, ITEMS, CONTIG>(tid, i);
        if (e < n) {
            txy[pos[i]] = EARLY ? c[i] : xy[e];
            if (decoded) decoded[pos[i]] = e;
        }
    }
    __syncthreads();
    // BATCH (as in block_rank_by): the two ends of four edges are read
    // before any is used; otherwise one edge at a time
    constexpr int G = BATCH ? (ITEMS < 4 ? ITEMS : 4) : 1;
    FxAcc acc;
#pragma unroll
    for (int i0 = 0; i0 < ITEMS; i0 += G) {
        float2 a[G], b[G];
#pragma unroll
        for (int j = 0; j < G; ++j) {
            const int q = tid + (i0 + j) * BLOCK;
            a[j] = txy[q < n ? q : 0];
            b[j] = txy[q + 1 < n ? q + 1 : 0];
        }
#pragma unroll
        for (int j = 0; j < G; ++j) {
            const bool in = tid + (i0 + j) * BLOCK < n;
            if constexpr (BATCH) {
                // few positions past n (a thread's four keys are its own
                // column): those add a zero edge, with no per-edge branch
                const float ej = edge_xy(a[j].x, a[j].y, b[j].x, b[j].y);
                fx_add(acc, in ? ej : 0.0f, p.scale0);
            } else if (in) {
                fx_add(acc, edge_xy(a[j].x, a[j].y, b[j].x, b[j].y), p.scale0);
            }
        }
    }
    const long long tot = block_sum_fx<BLOCK>(fx_value(acc), ss);
    if (tid == 0 && fit_out) fit_out[0] = fx_to_float(tot, p.inv0);
}

__device__ __forceinline__ void key_components(const float4& v, float (&k)[4])
{
    k[0] = v.x;
    k[1] = v.y;
    k[2] = v.z;
    k[3] = v.w;
}
__device__ __forceinline__ void key_components(const float2& v, float (&k)[2])
{
    k[0] = v.x;
    k[1] = v.y;
}

// TSP over coordinates with the child's keys in registers: thread t holds
// keys W t .. W t + W - 1 (the float4 or float2 column it built, so n <= W
// BLOCK) and ranks them straight away, with no barrier between building
// the child and ranking it.  The argsort counters were cleared before the
// caller's last barrier.
template <int BLOCK, typename V>
__device__ __forceinline__ void decode_tsp_coords_reg(const ProblemDesc& p, char* lds, const DecodeLayout& lay,
                                                      int nb, float* fit_out, V kv, int* long_flag = nullptr)
{
    constexpr int W = sizeof(V) / sizeof(float);
    const int tid = ltid<BLOCK>();
    const int n = p.n;
    SortScratch ss;
    ss.cnt = reinterpret_cast<int*>(lds + lay.cnt_off);
    ss.items = reinterpret_cast<unsigned long long*>(lds + lay.items_off);
    ss.red = reinterpret_cast<long long*>(lds + lay.red_off);
    ss.flag = long_flag;
    float kk[W];
    key_components(kv, kk);
    const float2* xy = reinterpret_cast<const float2*>(p.a);
    float2 c[W];
    int pos[W];
    auto load = [&]() {
#pragma unroll
        for (int i = 0; i < W; ++i) {
            const int e = W * tid + i;
            c[i] = xy[e < n ? e : n - 1];  // elements past n are never used
        }
    };
    auto key_at = [&](int i, int) { return kk[i]; };
    block_rank_by<BLOCK, true, W, true, decltype(key_at), decltype(load), false, true>(key_at, n, nb, ss, pos, nullptr,
                                                                                      load);
    tsp_coords_tail<BLOCK, W, true, true, true>(p, lds, lay, ss, nb, c, pos, fit_out, nullptr);
}

// CLEARED: the argsort counters were cleared (and a barrier passed) by the
// caller.  XY_LDS: coordinates were staged in LDS at xy0 / xy1 (coordinate
// kinds only); otherwise they are read from p.a / p.b.  ITEMS: keys per
// thread of the sorting decoders, n <= ITEMS * BLOCK.
template <int KIND, int BLOCK, bool CLEARED = false, bool XY_LDS = false, int ITEMS = kSortItems>
__device__ void decode_block(const ProblemDesc& p, char* lds, const DecodeLayout& lay, int nb,
                             float* fit_out, int* decoded, const float2* xy0 = nullptr,
                             const float2* xy1 = nullptr)
{
    const int tid = threadIdx.x;
    const int n = p.n;
    float* keys = reinterpret_cast<float*>(lds + lay.keys_off);
    int* perm = reinterpret_cast<int*>(lds + lay.perm_off);
    long long* red = reinterpret_cast<long long*>(lds + lay.red_off);
    SortScratch ss;
    ss.cnt = reinterpret_cast<int*>(lds + lay.cnt_off);
    ss.items = reinterpret_cast<unsigned long long*>(lds + lay.items_off);
    ss.red = red;

    if constexpr (KIND == KIND_TSP_COORDS) {
        // Each key writes its city's coordinates at its tour position, over
        // the sort items once the ranking is done; edges are then summed
        // from consecutive positions.
        // With few keys per thread the coordinate loads are issued before
        // the ranking; with many they would hold too many registers.
        constexpr bool kEarly = ITEMS <= 8;
        const float2* xy = XY_LDS ? xy0 : reinterpret_cast<const float2*>(p.a);
        float2 c[kEarly ? ITEMS : 1];
        int pos[ITEMS];
        if constexpr (kEarly) {
            auto load = [&]() {
#pragma unroll
                for (int i = 0; i < ITEMS; ++i) {
                    const int e = tid + i * BLOCK;
                    if (e < n) c[i] = xy[e];
                }
            };
            block_rank<BLOCK, CLEARED, ITEMS>(keys, n, nb, ss, pos, nullptr, load);
            tsp_coords_tail<BLOCK, ITEMS, false, true>(p, lds, lay, ss, nb, c, pos, fit_out, decoded);
        } else {
            // the keys stay in LDS until the tail (no perm is written)
            block_rank<BLOCK, CLEARED, ITEMS, NoHook, true>(keys, n, nb, ss, pos);
            const float2 none[ITEMS] = {};
            tsp_coords_tail<BLOCK, ITEMS, false, false>(p, lds, lay, ss, nb, none, pos, fit_out, decoded, xy);
        }
    } else if constexpr (KIND == KIND_TSP_MATRIX) {
        block_argsort<BLOCK, CLEARED, ITEMS>(keys, n, nb, ss, perm);
        // thread t sums the edges leaving positions [t*chunk, t*chunk+chunk)
        const int chunk = (n + BLOCK - 1) / BLOCK;
        const int p0 = tid * chunk;
        const int m = p0 < n ? (p0 + chunk < n ? chunk : n - p0) : 0;
        int pv[ITEMS + 1];
#pragma unroll
        for (int k = 0; k <= ITEMS; ++k) {
            const int p = p0 + k;
            if (k <= m && m > 0) pv[k] = perm[p < n ? p : p - n];
        }
        FxAcc acc;
        float dd[ITEMS];
#pragma unroll
        for (int k = 0; k < ITEMS; ++k)
            if (k < m) dd[k] = p.a[(size_t)pv[k] * n + pv[k + 1]];
#pragma unroll
        for (int k = 0; k < ITEMS; ++k)
            if (k < m) fx_add(acc, dd[k], p.scale0);
        const long long tot = block_sum_fx<BLOCK>(fx_value(acc), ss);
        if (tid == 0 && fit_out) fit_out[0] = fx_to_float(tot, p.inv0);
        if (decoded)
            for (int q = tid; q < n; q += BLOCK) decoded[q] = perm[q];
    } else if constexpr (KIND == KIND_KNAP_THRESHOLD) {
        long long tw = 0, tv = 0;
        for (int i = tid; i < n; i += BLOCK) {
            const int take = keys[i] > 0.5f;
            if (take) {
                tw += p.w[i];
                tv += p.v[i];
            }
            if (decoded) decoded[i] = take;
        }
        tw = block_sum_i64<BLOCK>(tw, red);
        tv = block_sum_i64<BLOCK>(tv, red);
        if (tid == 0 && fit_out) {
            const long long f = tw > p.cap ? tv - 1000LL * (tw - p.cap) : tv;
            fit_out[0] = __ll2float_rn(f);
        }
    } else if constexpr (KIND == KIND_KNAP_GREEDY) {
        block_argsort<BLOCK, CLEARED, ITEMS>(keys, n, nb, ss, perm);
        // With weights >= 0 the greedy fill takes every item up to the first
        // position f where the running weight in key order exceeds the
        // capacity (a block prefix scan finds f), rejects item f, and from
        // there on the remaining capacity is below w[f]: wave 0 walks the
        // rest 64 items at a time, packing the first open item that fits.
        // The weights and values are gathered in key order into LDS first
        // (over the argsort's items, free after its final barrier), every
        // thread's loads issued together, so neither the scan nor the walk
        // waits on global memory.
        int2* wv = reinterpret_cast<int2*>(ss.items);
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
            const int q = tid + k * BLOCK;
            if (q < n) {
                const int item = perm[q];
                wv[q] = make_int2(p.w[item], p.v[item]);
            }
        }
        __syncthreads();
        const int chunk = (n + BLOCK - 1) / BLOCK;
        const int lo = tid * chunk < n ? tid * chunk : n;
        const int hi = lo + chunk < n ? lo + chunk : n;
        long long wsum = 0;
        for (int q = lo; q < hi; ++q) wsum += wv[q].x;
        long long run = block_inclusive_scan_i64<BLOCK>(wsum, red) - wsum;
        int first = n;
        long long vsum = 0, rem_at = 0;
        for (int q = lo; q < hi; ++q) {
            const int2 e = wv[q];
            const long long w = e.x;
            if (run + w > p.cap) {
                first = q;
                rem_at = p.cap - run;
                break;
            }
            run += w;
            vsum += e.y;
        }
        const int f = (int)-block_max_i64<BLOCK>(-(long long)first, red);
        long long val = block_sum_i64<BLOCK>(vsum, red);
        // the thread that owns position f publishes the capacity left there
        if (first == f && f < n) red[16] = rem_at;
        __syncthreads();
        if (decoded)
            for (int q = tid; q < n && q <= f; q += BLOCK) decoded[perm[q]] = q < f ? 1 : 0;
        if (tid < kWave) {
            const int lane = tid;
            long long rem = f < n ? red[16] : 0;
            for (int base = f + 1; base < n; base += kWave) {
                const int q = base + lane;
                const bool valid = q < n;
                const int2 e = valid ? wv[q] : make_int2(0, 0);
                const long long w = e.x;
                const long long v = e.y;
                bool open = valid;
                bool taken = false;
                while (true) {
                    const unsigned long long fits = __ballot(open && w <= rem);
                    if (fits == 0ull) break;
                    const int fl = __builtin_ctzll(fits);
                    rem -= __shfl(w, fl, kWave);
                    val += __shfl(v, fl, kWave);
                    taken = taken || lane == fl;
                    open = open && lane > fl;
                }
                if (decoded && valid) decoded[perm[q]] = taken ? 1 : 0;
            }
            if (lane == 0 && fit_out) fit_out[0] = __ll2float_rn(val);
        }
    } else if constexpr (KIND == KIND_TSPJ) {
        int* tour = perm;
        int* jobs = perm + n;
        block_argsort<BLOCK, CLEARED, ITEMS>(keys, n, nb, ss, tour);
        block_argsort<BLOCK, false, ITEMS>(keys + n, n, nb, ss, jobs);
        const int dim = p.dim;
        const int chunk = (n + BLOCK - 1) / BLOCK;
        const int lo = tid * chunk;
        const int hi = lo + chunk < n ? lo + chunk : n;
        long long local = 0;
        for (int q = lo; q < hi; ++q) {
            const int prev = q == 0 ? 0 : tour[q - 1] + 1;
            local += fx_term(p.a[(size_t)prev * dim + tour[q] + 1], p.scale0);
        }
        long long t = block_inclusive_scan_i64<BLOCK>(local, red) - local;
        long long best = 0;
        for (int q = lo; q < hi; ++q) {
            const int city = tour[q] + 1;
            const int prev = q == 0 ? 0 : tour[q - 1] + 1;
            t += fx_term(p.a[(size_t)prev * dim + city], p.scale0);
            const long long done = t + fx_term(p.b[(size_t)city * dim + jobs[q] + 1], p.scale0);
            best = done > best ? done : best;
            if (q == n - 1) {
                const long long total = t + fx_term(p.a[(size_t)city * dim], p.scale0);
                best = total > best ? total : best;
            }
        }
        best = block_max_i64<BLOCK>(best, red);
        if (tid == 0 && fit_out) fit_out[0] = fx_to_float(best, p.inv0);
        if (decoded)
            for (int q = tid; q < n; q += BLOCK) {
                decoded[q] = tour[q] + 1;
                decoded[n + q] = jobs[q] + 1;
            }
    } else if constexpr (KIND == KIND_BITSP_COORDS || KIND == KIND_BITSP_MATRIX) {
        block_argsort<BLOCK, CLEARED, ITEMS>(keys, n, nb, ss, perm);
        const int dim = p.dim;
        long long acc0 = 0, acc1 = 0;
        for (int q = tid; q <= n; q += BLOCK) {
            const int from = q == 0 ? 0 : perm[q - 1] + 1;
            const int to = q == n ? 0 : perm[q] + 1;
            float d0, d1;
            if constexpr (KIND == KIND_BITSP_COORDS) {
                const float2* xa = XY_LDS ? xy0 : reinterpret_cast<const float2*>(p.a);
                const float2* xb = XY_LDS ? xy1 : reinterpret_cast<const float2*>(p.b);
                const float2 a0 = xa[from], a1 = xa[to], b0 = xb[from], b1 = xb[to];
                d0 = edge_xy(a0.x, a0.y, a1.x, a1.y);
                d1 = edge_xy(b0.x, b0.y, b1.x, b1.y);
            } else {
                d0 = p.a[(size_t)from * dim + to];
                d1 = p.b[(size_t)from * dim + to];
            }
            acc0 += fx_term(d0, p.scale0);
            acc1 += fx_term(d1, p.scale1);
        }
        acc0 = block_sum_i64<BLOCK>(acc0, red);
        acc1 = block_sum_i64<BLOCK>(acc1, red);
        if (tid == 0 && fit_out) {
            fit_out[0] = fx_to_float(acc0, p.inv0);
            fit_out[1] = fx_to_float(acc1, p.inv1);
        }
        if (decoded)
            for (int q = tid; q < n; q += BLOCK) decoded[q] = perm[q];
    } else if constexpr (KIND == KIND_CONTINUOUS) {
        // sum of per-variable terms (continuous_opt_gpu_config.hpp:132-160);
        // Rastrigin adds the constant 10 n exactly
        const float lo = p.f0, hi = p.f1;
        const int fn = p.ip0;
        long long acc = 0;
        for (int i = tid; i < n; i += BLOCK) {
#pragma clang fp contract(off)
            const float x = cont_var(keys[i], lo, hi);
            if (decoded) decoded[i] = __float_as_int(x);
            float t;
            if (fn == FN_SPHERE) {
                t = x * x;
            } else if (fn == FN_RASTRIGIN) {
                const float c = cos_det(6.2831854820251465f * x);
                t = x * x - 10.0f * c;
            } else {
                if (i + 1 >= n) continue;
                const float xn = cont_var(keys[i + 1], lo, hi);
                const float t1 = xn - x * x;
                const float t2 = 1.0f - x;
                t = (100.0f * t1) * t1 + t2 * t2;
            }
            acc += fx_term(t, p.scale0);
        }
        acc = block_sum_i64<BLOCK>(acc, red);
        if (tid == 0 && fit_out) {
            if (fn == FN_RASTRIGIN) acc += fx_term(10.0f * (float)n, p.scale0);
            fit_out[0] = fx_to_float(acc, p.inv0);
        }
    } else if constexpr (KIND == KIND_JOBSHOP) {
        // machine loads in LDS bins, makespan = largest load, empty machines
        // count as 0 (job_scheduling_gpu_config.hpp:138-158)
        long long* bins = reinterpret_cast<long long*>(lds + lay.aux_off);
        const int M = p.ip0;
        for (int m = tid; m < M; m += BLOCK) bins[m] = 0;
        __syncthreads();
        for (int i = tid; i < n; i += BLOCK) {
            const int m = job_machine(keys[i], M);
            atomicAdd(reinterpret_cast<unsigned long long*>(&bins[m]),
                      (unsigned long long)fx_term(p.a[i], p.scale0));
            if (decoded) decoded[i] = m;
        }
        __syncthreads();
        long long best = 0;
        for (int m = tid; m < M; m += BLOCK) best = bins[m] > best ? bins[m] : best;
        best = block_max_i64<BLOCK>(best, red);
        if (tid == 0 && fit_out) fit_out[0] = fx_to_float(best, p.inv0);
    } else if constexpr (KIND == KIND_ZDT1) {
        // f1 = x1, g = 1 + 9 sum(x2..xn) / (n - 1), f2 = g (1 - sqrt(f1 / g))
        // (examples/zdt1.cu:43-63); the key sum is exact fixed point
        long long acc = 0;
        for (int i = tid; i < n; i += BLOCK) {
            if (i > 0) acc += fx_term(keys[i], p.scale0);
            if (decoded) decoded[i] = __float_as_int(keys[i]);
        }
        acc = block_sum_i64<BLOCK>(acc, red);
        if (tid == 0 && fit_out) {
            const float f1 = keys[0];
            float g = 1.0f;
            if (n > 1) {
                const double sum = (double)acc * (double)p.inv0;
                g = (float)(1.0 + 9.0 * sum / (double)(n - 1));
            }
            const float r = sqrt_rn(__fdiv_rn(f1, g));
            fit_out[0] = f1;
            fit_out[1] = (float)((double)g * (1.0 - (double)r));
        }
    } else if constexpr (KIND == KIND_VRP_DAYS) {
        // Stops in key order fill day routes from the depot
        // (vrp_blazer_config.hpp:192-282): a stop joins the day while the day
        // has fewer than ip0 stops and cumulative drive + leg + return stays
        // within d0 minutes (FP32 sums, compared in FP64); a stop that fits
        // no empty day is skipped.  Every start position walks its own day
        // in parallel; one thread then chains the days from position 0.
        // Fitness = sum of day drive times + 1000 per entry of the
        // reference's unvisited list, which holds each skipped stop twice
        // (vrp_blazer_config.hpp:265-280, 285-296).
        block_argsort<BLOCK, CLEARED, ITEMS>(keys, n, nb, ss, perm);
        long long* dfx = reinterpret_cast<long long*>(lds + lay.items_off);
        int* nxt = reinterpret_cast<int*>(lds + lay.aux_off);
        int* dnum = nxt + n;
        const int dim = p.dim, maxs = p.ip0;
        const double lim = p.d0;
        for (int q = tid; q < n; q += BLOCK) {
#pragma clang fp contract(off)
            int cur = 0, k = 0;
            float cum = 0.0f;
            while (q + k < n && k < maxs) {
                const int st = perm[q + k] + 1;
                const float tts = p.a[(size_t)cur * dim + st];
                const float ttd = p.a[(size_t)st * dim];
                const float tot = (cum + tts) + ttd;
                if (!((double)tot <= lim)) break;
                cum = cum + tts;
                cur = st;
                ++k;
            }
            nxt[q] = k > 0 ? q + k : -(q + 1);
            dfx[q] = k > 0 ? fx_term(cum + p.a[(size_t)cur * dim], p.scale0) : 0;
            dnum[q] = 0;
        }
        __syncthreads();
        if (tid == 0) {
            long long acc = 0;
            int skipped = 0, day = 0;
            for (int q = 0; q < n;) {
                const int nx = nxt[q];
                if (nx < 0) {
                    ++skipped;
                    dnum[q] = -1;
                    q = -nx;
                } else {
                    acc += dfx[q];
                    dnum[q] = ++day;
                    q = nx;
                }
            }
            acc += fx_term(2000.0f * (float)skipped, p.scale0);
            if (fit_out) fit_out[0] = fx_to_float(acc, p.inv0);
        }
        if (decoded) {
            __syncthreads();
            // the day of a position is that of the nearest chained start at
            // or before it whose day reaches it
            for (int q = tid; q < n; q += BLOCK) {
                int dv = 0;
                if (dnum[q] >= 0) {
                    for (int s0 = q; s0 >= 0 && s0 > q - maxs; --s0)
                        if (dnum[s0] > 0 && nxt[s0] > q) {
                            dv = dnum[s0];
                            break;
                        }
                }
                decoded[perm[q]] = dv;
            }
        }
    } else if constexpr (KIND == KIND_MOVRP) {
        // Customers in biased key order (the bias adds i 1e-5 and a running
        // FP32 hash of the keys, multi_objective_vrp_config.hpp:216-228)
        // split greedily into routes by vehicle capacity (a route always
        // takes its first customer, 233-254).  Objectives: total distance +
        // 100 per unit of overload, and the longest route; routes beyond the
        // vehicle count cost 500 each in both (272-300).
        float* vals = keys;
        int* dem = reinterpret_cast<int*>(lds + lay.aux_off);
        int* nxt = dem + round_up_int(n + 1, 2);
        int* rid = nxt + n;
        long long* rload = reinterpret_cast<long long*>(rid + n);
        __shared__ float s_hash;
        for (int i = tid; i <= n; i += BLOCK) dem[i] = p.w[i];
        if (tid == 0) {
#pragma clang fp contract(off)
            float h = 0.0f;
            for (int i = 0; i < n; ++i) h = h + keys[i] * (float)(i + 1);
            s_hash = h;
        }
        __syncthreads();
        {
#pragma clang fp contract(off)
            const float hb = s_hash * 0.00001f;
            for (int i = tid; i < n; i += BLOCK) vals[i] = keys[i] + ((float)i * 0.00001f + hb);
        }
        if (CLEARED) __syncthreads();
        block_argsort<BLOCK, CLEARED, ITEMS>(vals, n, nb, ss, perm);
        const long long cap = p.cap;
        for (int q = tid; q < n; q += BLOCK) {
            long long load = 0;
            int k = 0;
            while (q + k < n) {
                const long long dd = dem[perm[q + k] + 1];
                if (k > 0 && load + dd > cap) break;
                load += dd;
                ++k;
            }
            nxt[q] = q + k;
            rload[q] = load;
            rid[q] = 0;
        }
        long long* rsum = reinterpret_cast<long long*>(lds + lay.items_off);
        for (int q = tid; q < n; q += BLOCK) rsum[q] = 0;
        __syncthreads();
        if (tid == 0) {
            int r = 0;
            for (int q = 0; q < n; q = nxt[q]) rid[q] = ++r;
            red[16] = r;
        }
        __syncthreads();
        const int routes = (int)red[16];
        // route of each position: count of route starts up to it
        const int chunk = (n + BLOCK - 1) / BLOCK;
        const int lo = tid * chunk < n ? tid * chunk : n;
        const int hi = lo + chunk < n ? lo + chunk : n;
        long long starts = 0;
        for (int q = lo; q < hi; ++q) starts += rid[q] > 0 ? 1 : 0;
        long long run = block_inclusive_scan_i64<BLOCK>(starts, red) - starts;
        const float2* xy = reinterpret_cast<const float2*>(p.a);
        long long tot = 0, pen = 0;
        for (int q = lo; q < hi; ++q) {
            if (rid[q] > 0) {
                ++run;
                if (rload[q] > cap) pen += fx_term((float)(rload[q] - cap) * 100.0f, p.scale0);
            }
            const int c = perm[q] + 1;
            const int prev = rid[q] > 0 ? 0 : perm[q - 1] + 1;
            long long leg = fx_term(edge_xy(xy[prev].x, xy[prev].y, xy[c].x, xy[c].y), p.scale0);
            if (q == n - 1 || rid[q + 1] > 0) leg += fx_term(edge_xy(xy[c].x, xy[c].y, xy[0].x, xy[0].y), p.scale0);
            tot += leg;
            atomicAdd(reinterpret_cast<unsigned long long*>(&rsum[run - 1]), (unsigned long long)leg);
            if (decoded) decoded[c - 1] = (int)run;
        }
        tot = block_sum_i64<BLOCK>(tot + pen, red);
        __syncthreads();
        long long best = 0;
        for (int r = tid; r < routes; r += BLOCK) best = rsum[r] > best ? rsum[r] : best;
        best = block_max_i64<BLOCK>(best, red);
        if (tid == 0 && fit_out) {
            const long long veh = routes > p.ip0 ? fx_term((float)(routes - p.ip0) * 500.0f, p.scale0) : 0;
            fit_out[0] = fx_to_float(tot + veh, p.inv0);
            fit_out[1] = fx_to_float(best + veh, p.inv0);
        }
    }
}

}  // namespace brkga
