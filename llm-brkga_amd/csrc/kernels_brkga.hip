// kernels_brkga.hip - BRKGA generation step for gfx950.
//
//   k_evaluate       decode + fitness of caller-owned key rows: the
//                    evaluate_population_gpu hook (core/config.hpp:50-53).
//   k_step_one       one workgroup per child slot: elite / migrant row copy,
//                    mutant keys from Philox, or biased uniform crossover,
//                    then decode + fitness of the new row straight from LDS.
//                    Replaces brkga_generation_kernel (core/cuda_kernels.cuh:88-138)
//                    fused with the config fitness kernels, and the host-staged
//                    ring migration (core/solver.hpp:517-558).
//   k_partition_*    stable sort of each island's fitness -> rank order;
//                    replaces thrust::sequence + sort_by_key (core/solver.hpp:416-419).
//
// The *_one kernels take the population descriptor by value (kernel
// arguments); k_partition_many reads one descriptor per instance from
// device memory (batched and served populations).  `cur` is the buffer
// parity of the parent generation, known to the launcher; the generation
// counter in device memory only feeds the Philox counter and the migration
// schedule, so a captured graph can be replayed.
//
// Key rows are [pop][ld] FP32 with ld a multiple of 32, so every row starts
// 128-byte aligned and is streamed as float4.
#include <hip/hip_ext.h>

#include "launch.hpp"
#include "step_device.hpp"

namespace brkga {

template <int KIND, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_evaluate(ProblemDesc p, const float* __restrict__ keys, long long ld,
                                                    int count, float* __restrict__ fit, int* __restrict__ decoded)
{
    extern __shared__ __align__(16) char lds[];
    const int row = blockIdx.x;
    if (row >= count) return;
    const int L = p.chrom_len;
    const int nb = sort_buckets(p.n);
    const DecodeLayout lay = decode_layout(KIND, p.n, L, nb, p.aux);
    float* k = reinterpret_cast<float*>(lds + lay.keys_off);
    const float* src = keys + (size_t)row * ld;
    for (int j = threadIdx.x; j < L; j += BLOCK) k[j] = src[j];
    __syncthreads();
    decode_block<KIND, BLOCK>(p, lds, lay, nb, fit ? fit + (size_t)row * p.num_obj : nullptr,
                              decoded ? decoded + (size_t)row * L : nullptr);
}

// One block of a generation step: slot gs of the population described by d.
// n <= ITEMS * BLOCK.
template <int KIND, int BLOCK, int ITEMS>
__device__ __forceinline__ void step_body(const InstDesc& d, int gs, int cur, int init, char* lds)
{
    const int Pi = d.Pi, Ei = d.Ei, I = d.islands;
    const int island = gs / Pi;
    const int s = gs - island * Pi;
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
    if (!init && !maybe_migrant && s < Ei) {
        const int src = island * Pi + perm_cur[s];
        const float4* in = reinterpret_cast<const float4*>(d.keys[cur] + (size_t)src * ld);
        for (int c = threadIdx.x; c < nq; c += BLOCK) out[c] = in[c];
        if (threadIdx.x < M) fit_out[threadIdx.x] = d.fit[cur][(size_t)src * M + threadIdx.x];
        return;
    }

    const int t = init ? 0 : *d.gen;
    const int nb = d.nb;
    const StepLayout sl = step_layout(KIND, n, L, nb, Pi, false, d.prob.aux);
    float4* kl = reinterpret_cast<float4*>(lds + sl.dec.keys_off);
    int* order = reinterpret_cast<int*>(lds + sl.isl_off);
    const bool staged = Pi <= kMaxStagedOrder;
    SortScratch ss;
    ss.cnt = reinterpret_cast<int*>(lds + sl.dec.cnt_off);
    ss.red = reinterpret_cast<long long*>(lds + sl.dec.red_off);

    // stage the island rank order, clear the argsort counters
    if (!init && staged)
        for (int i = threadIdx.x; i < Pi; i += BLOCK) order[i] = perm_cur[i];
    argsort_clear<BLOCK>(ss, nb);
    __syncthreads();

    const bool child = build_child<BLOCK>(d, island, s, cur, init, t, staged ? order : perm_cur, out, kl, fit_out);
    if (!child) return;
    __syncthreads();
    decode_block<KIND, BLOCK, true, false, ITEMS>(d.prob, lds, sl.dec, nb, fit_out, nullptr);
}

template <int KIND, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_step_one(InstDesc d, int cur, int init)
{
    extern __shared__ __align__(16) char lds[];
    step_body<KIND, BLOCK, ITEMS>(d, blockIdx.x, cur, init, lds);
}

constexpr int kPartBlock = 1024;

// Workgroups that rank one island together: islands of 128 to kPartBlock
// individuals are split into slices of 64 keys, one workgroup each.
__host__ __device__ inline int partition_parts(int Pi)
{
    const int Pn = pow2_at_least(Pi < 2 ? 2 : Pi);
    return (Pn < 128 || Pn > kPartBlock) ? 1 : Pn / 64;
}

// Stable rank order of island `island`'s fitness in buffer `buf`.  Up to
// kPartBlock individuals: every key (fitness order key, index) is unique,
// so the rank of a key is the number of smaller keys; workgroup `part` of
// `parts` ranks the keys of its slice, G = kPartBlock / slice threads per
// key counting over interleaved shares and combining with wave shuffles.
// Larger islands use a bitonic network (parts = 1).  mode 1: the last
// workgroup of the population bumps the generation counter to t + 1 and
// raises the local-search flag when generation t + 1 is due; mode 2: the
// last workgroup lowers the flag.
__device__ __forceinline__ void partition_body(const InstDesc& d, int island, int part, int parts, int buf,
                                               int mode, int t, char* lds)
{
    constexpr int B = kPartBlock;
    unsigned long long* sk = reinterpret_cast<unsigned long long*>(lds);
    const int Pi = d.Pi;
    const int M = d.prob.num_obj;
    const float* fit = d.fit[buf] + (size_t)island * Pi * M;
    int* perm = d.perm[buf] + (size_t)island * Pi;
    const int maximize = d.prob.maximize;
    const int Pn = pow2_at_least(Pi < 2 ? 2 : Pi);
    for (int i = threadIdx.x; i < Pn; i += B)
        sk[i] = i < Pi ? ((unsigned long long)fitness_order_key(fit[(size_t)i * M], maximize) << 32) | (unsigned)i
                       : ~0ull;
    __syncthreads();
    if (Pn <= B) {
        const int slice = Pn / parts;
        const int G = B / slice;  // threads per key, power of two <= 512
        const int e = part * slice + threadIdx.x / G, sub = threadIdx.x % G;
        const unsigned long long mine = sk[e];
        const ulonglong2* pairs = reinterpret_cast<const ulonglong2*>(sk);
        int cnt = 0;
        for (int q = sub; q < (Pn >> 1); q += G) {
            const ulonglong2 v = pairs[q];
            cnt += (v.x < mine) + (v.y < mine);
        }
        for (int o = 1; o < G && o < kWave; o <<= 1) cnt += __shfl_xor(cnt, o, kWave);
        if (G > kWave) {
            // more than a wave per key (Pn < 16): fold the waves through LDS
            __syncthreads();
            int* part = reinterpret_cast<int*>(sk + Pn);
            if ((threadIdx.x & (kWave - 1)) == 0) part[threadIdx.x / kWave] = cnt;
            __syncthreads();
            if (sub == 0) {
                cnt = 0;
                for (int w = 0; w < G / kWave; ++w) cnt += part[threadIdx.x / kWave + w];
            }
        }
        if (sub == 0 && e < Pi) perm[cnt] = e;
    } else {
        for (int k = 2; k <= Pn; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int i = threadIdx.x; i < (Pn >> 1); i += B) {
                    const int lo = 2 * j * (i / j) + (i % j);
                    const int hi = lo + j;
                    const bool up = (lo & k) == 0;
                    const unsigned long long x = sk[lo], y = sk[hi];
                    if ((x > y) == up) {
                        sk[lo] = y;
                        sk[hi] = x;
                    }
                }
                __syncthreads();
            }
        }
        for (int i = threadIdx.x; i < Pi; i += B) perm[i] = (int)(sk[i] & 0xFFFFFFFFull);
    }
    if (mode != 0) {
        // The counters only change once every island block is done; no
        // block reads what another one wrote (the next kernel sees the rank
        // orders through the kernel boundary), so no fence is needed here.
        __syncthreads();
        if (threadIdx.x == 0) {
            const int old = atomicAdd(d.done, 1);
            if (old == d.islands * parts - 1) {
                *d.done = 0;
                if (mode == 1) {
                    const int g = t + 1;
                    *d.ls_flag = d.ls_interval > 0 && g > 0 && g % d.ls_interval == 0 ? 1 : 0;
                    *d.gen = g;
                } else {
                    *d.ls_flag = 0;
                }
            }
        }
    }
}

// Grid: islands x partition_parts(Pi) workgroups.
__global__ __launch_bounds__(kPartBlock) void k_partition_one(InstDesc d, int buf, int mode)
{
    extern __shared__ __align__(16) char lds[];
    const int parts = partition_parts(d.Pi);
    int t = 0;
    if (mode == 1) {
        t = *d.gen;
    } else if (mode == 2) {
        if (*d.ls_flag == 0) return;
        buf = *d.gen & 1;
    }
    partition_body(d, blockIdx.x / parts, blockIdx.x % parts, parts, buf, mode, t, lds);
}

__global__ __launch_bounds__(kPartBlock) void k_partition_many(const InstDesc* __restrict__ descs, int num_inst,
                                                               int buf, int mode)
{
    extern __shared__ __align__(16) char lds[];
    const InstDesc& d = descs[find_instance_by_seg(descs, num_inst, blockIdx.x)];
    const int island = blockIdx.x - d.seg_begin;
    if (island >= d.islands) return;
    int t = 0, b = buf;
    if (mode == 1) {
        // t = -1: generation 0 was just built into buffer 0; the counter
        // moves to 0 like after any other step
        t = *d.gen;
        if (t >= d.gen_limit) return;
        if (b < 0) b = (t + 1) & 1;
    } else if (mode == 2) {
        if (*d.ls_flag == 0) return;
        b = *d.gen & 1;
    } else if (b < 0) {
        b = 0;
    }
    partition_body(d, island, 0, 1, b, mode, t, lds);
}

// TSPJ local search: one workgroup per (population, island, rank r < ls_k)
// of every population whose local-search flag is raised (its current
// generation g is due).  The individual's city tour (stable argsort of the
// city keys) takes ls_moves random 2-opt moves: Philox words 2m, 2m+1 of
// (island, g, r, LSEARCH) give positions x = u0 % n, y = u1 % n; when they
// are at least two apart the tour between them is reversed.  The city keys,
// sorted, are then handed to the new tour in order, and the candidate
// replaces the individual only if its makespan is strictly smaller.
// Reference: tspj_config.hpp:77-135, cuda_kernels.cuh:166-246 (which
// reverses key values by position and so does not reverse the decoded
// tour; the tour reversal is what is done here).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_tspj_local_search(const InstDesc* __restrict__ descs, int num_inst)
{
    extern __shared__ __align__(16) char lds[];
    __shared__ float s_fit;
    const InstDesc& d = descs[blockIdx.y];
    if (d.prob.kind != KIND_TSPJ || d.ls_k <= 0) return;
    const int item = blockIdx.x;
    if (item >= d.islands * d.ls_k) return;
    if (*d.ls_flag == 0) return;
    const int g = *d.gen;
    const int b = g & 1;
    const int island = item / d.ls_k, r = item - island * d.ls_k;
    const int n = d.prob.n, nb = d.nb;
    const int slot = island * d.Pi + d.perm[b][island * d.Pi + r];
    float* row = d.keys[b] + (size_t)slot * d.ld;
    const DecodeLayout lay = decode_layout(KIND_TSPJ, n, 2 * n, nb, 0);
    float* keys = reinterpret_cast<float*>(lds + lay.keys_off);
    int* tour = reinterpret_cast<int*>(lds + lay.bytes);
    float* vals = reinterpret_cast<float*>(tour + n);
    float* cand = vals + n;
    SortScratch ss;
    ss.cnt = reinterpret_cast<int*>(lds + lay.cnt_off);
    ss.items = reinterpret_cast<unsigned long long*>(lds + lay.items_off);
    ss.red = reinterpret_cast<long long*>(lds + lay.red_off);
    for (int i = threadIdx.x; i < n; i += BLOCK) {
        const float k = row[i];
        keys[i] = k;
        cand[i] = k;
    }
    __syncthreads();
    // decoded tour into the key area (the argsort overlays it)
    block_argsort<BLOCK>(keys, n, nb, ss, reinterpret_cast<int*>(keys));
    const int* order = reinterpret_cast<const int*>(keys);
    for (int p = threadIdx.x; p < n; p += BLOCK) {
        const int c = order[p];
        tour[p] = c;
        vals[p] = cand[c];
    }
    __syncthreads();
    for (int m = 0; m < d.ls_moves; ++m) {
        const U4 w = rng_block(d.seed_lo, d.seed_hi, (uint32_t)island, (uint32_t)g, (uint32_t)r, TAG_LSEARCH, m >> 1);
        int x = (int)(((m & 1) ? w.z : w.x) % (uint32_t)n);
        int y = (int)(((m & 1) ? w.w : w.y) % (uint32_t)n);
        if (x > y) {
            const int tmp = x;
            x = y;
            y = tmp;
        }
        if (y - x < 2) continue;  // uniform across the block
        const int half = (y - x + 1) >> 1;
        for (int q = threadIdx.x; q < half; q += BLOCK) {
            const int a = tour[x + q];
            tour[x + q] = tour[y - q];
            tour[y - q] = a;
        }
        __syncthreads();
    }
    for (int p = threadIdx.x; p < n; p += BLOCK) cand[tour[p]] = vals[p];
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += BLOCK) {
        keys[i] = cand[i];
        keys[n + i] = row[n + i];
    }
    __syncthreads();
    decode_block<KIND_TSPJ, BLOCK>(d.prob, lds, lay, nb, &s_fit, nullptr);
    __syncthreads();
    const float old = d.fit[b][slot];
    if (s_fit < old) {
        for (int i = threadIdx.x; i < n; i += BLOCK) row[i] = cand[i];
        if (threadIdx.x == 0) d.fit[b][slot] = s_fit;
    }
}

int partition_lds_bytes(int Pi)
{
    const int Pn = pow2_at_least(Pi < 2 ? 2 : Pi);
    return Pn * 8 + 64;
}

// ------------------------------------------------------------- launchers --

namespace {

template <typename F>
void ensure_lds(F func, int bytes)
{
    if (bytes > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(func), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  bytes);
}

template <int KIND, int BLOCK>
hipError_t eval_one(const ProblemDesc& p, const float* keys, long long ld, int count, float* fit, int* decoded,
                    int lds, hipStream_t s)
{
    ensure_lds(k_evaluate<KIND, BLOCK>, lds);
    hipLaunchKernelGGL((k_evaluate<KIND, BLOCK>), dim3(count), dim3(BLOCK), lds, s, p, keys, ld, count, fit,
                       decoded);
    return hipGetLastError();
}

template <int KIND>
hipError_t eval_kind(int block, const ProblemDesc& p, const float* keys, long long ld, int count, float* fit,
                     int* decoded, int lds, hipStream_t s)
{
    switch (block) {
    case 64: return eval_one<KIND, 64>(p, keys, ld, count, fit, decoded, lds, s);
    case 128: return eval_one<KIND, 128>(p, keys, ld, count, fit, decoded, lds, s);
    case 256: return eval_one<KIND, 256>(p, keys, ld, count, fit, decoded, lds, s);
    case 512: return eval_one<KIND, 512>(p, keys, ld, count, fit, decoded, lds, s);
    case 1024: return eval_one<KIND, 1024>(p, keys, ld, count, fit, decoded, lds, s);
    default: return hipErrorInvalidValue;
    }
}

template <typename K, typename... Args>
void launch_maybe_timed(K kernel, dim3 grid, dim3 block, int lds, hipStream_t s, hipEvent_t start, hipEvent_t stop,
                        Args... args)
{
    if (start || stop)
        hipExtLaunchKernelGGL(kernel, grid, block, (uint32_t)lds, s, start, stop, 0u, args...);
    else
        hipLaunchKernelGGL(kernel, grid, block, lds, s, args...);
}

template <int KIND, int BLOCK>
hipError_t step_single(const InstDesc& d, int lds, int cur, int init, hipStream_t s, hipEvent_t start,
                       hipEvent_t stop)
{
    if (kind_sorts(KIND) && d.prob.n <= kCompactItems * BLOCK) {
        ensure_lds(k_step_one<KIND, BLOCK, kCompactItems>, lds);
        launch_maybe_timed(k_step_one<KIND, BLOCK, kCompactItems>, dim3(d.pop), dim3(BLOCK), lds, s, start, stop, d,
                           cur, init);
    } else if (kind_sorts(KIND) && d.prob.n <= 2 * kCompactItems * BLOCK) {
        ensure_lds(k_step_one<KIND, BLOCK, 2 * kCompactItems>, lds);
        launch_maybe_timed(k_step_one<KIND, BLOCK, 2 * kCompactItems>, dim3(d.pop), dim3(BLOCK), lds, s, start, stop,
                           d, cur, init);
    } else {
        ensure_lds(k_step_one<KIND, BLOCK, kSortItems>, lds);
        launch_maybe_timed(k_step_one<KIND, BLOCK, kSortItems>, dim3(d.pop), dim3(BLOCK), lds, s, start, stop, d, cur,
                           init);
    }
    return hipGetLastError();
}

template <int KIND>
hipError_t step_kind(int block, const InstDesc& d, int lds, int cur, int init, hipStream_t s, hipEvent_t start,
                     hipEvent_t stop)
{
    switch (block) {
    case 64: return step_single<KIND, 64>(d, lds, cur, init, s, start, stop);
    case 128: return step_single<KIND, 128>(d, lds, cur, init, s, start, stop);
    case 256: return step_single<KIND, 256>(d, lds, cur, init, s, start, stop);
    case 512: return step_single<KIND, 512>(d, lds, cur, init, s, start, stop);
    case 1024: return step_single<KIND, 1024>(d, lds, cur, init, s, start, stop);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_evaluate(const ProblemDesc& p, const float* keys, long long ld, int count, float* fit,
                           int* decoded, int block, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    const int lds = decode_layout(p.kind, p.n, p.chrom_len, sort_buckets(p.n), p.aux).bytes;
    switch (p.kind) {
    case KIND_TSP_COORDS: return eval_kind<KIND_TSP_COORDS>(block, p, keys, ld, count, fit, decoded, lds, s);
    case KIND_TSP_MATRIX: return eval_kind<KIND_TSP_MATRIX>(block, p, keys, ld, count, fit, decoded, lds, s);
    case KIND_KNAP_THRESHOLD:
        return eval_kind<KIND_KNAP_THRESHOLD>(block, p, keys, ld, count, fit, decoded, lds, s);
    case KIND_KNAP_GREEDY: return eval_kind<KIND_KNAP_GREEDY>(block, p, keys, ld, count, fit, decoded, lds, s);
    case KIND_TSPJ: return eval_kind<KIND_TSPJ>(block, p, keys, ld, count, fit, decoded, lds, s);
    case KIND_BITSP_COORDS: return eval_kind<KIND_BITSP_COORDS>(block, p, keys, ld, count, fit, decoded, lds, s);
    case KIND_BITSP_MATRIX: return eval_kind<KIND_BITSP_MATRIX>(block, p, keys, ld, count, fit, decoded, lds, s);
    case KIND_CONTINUOUS: return eval_kind<KIND_CONTINUOUS>(block, p, keys, ld, count, fit, decoded, lds, s);
    case KIND_JOBSHOP: return eval_kind<KIND_JOBSHOP>(block, p, keys, ld, count, fit, decoded, lds, s);
    case KIND_ZDT1: return eval_kind<KIND_ZDT1>(block, p, keys, ld, count, fit, decoded, lds, s);
    case KIND_VRP_DAYS: return eval_kind<KIND_VRP_DAYS>(block, p, keys, ld, count, fit, decoded, lds, s);
    case KIND_MOVRP: return eval_kind<KIND_MOVRP>(block, p, keys, ld, count, fit, decoded, lds, s);
    default: return hipErrorInvalidValue;
    }
}

int step_lds_bytes(const InstDesc& d)
{
    return step_layout(d.prob.kind, d.prob.n, d.L, d.nb, d.Pi, false, d.prob.aux).bytes;
}

hipError_t launch_step_one(const InstDesc& d, int block, int cur, int init, hipStream_t s, hipEvent_t start,
                           hipEvent_t stop)
{
    const int lds = step_lds_bytes(d);
    switch (d.prob.kind) {
    case KIND_TSP_COORDS: return step_kind<KIND_TSP_COORDS>(block, d, lds, cur, init, s, start, stop);
    case KIND_TSP_MATRIX: return step_kind<KIND_TSP_MATRIX>(block, d, lds, cur, init, s, start, stop);
    case KIND_KNAP_THRESHOLD: return step_kind<KIND_KNAP_THRESHOLD>(block, d, lds, cur, init, s, start, stop);
    case KIND_KNAP_GREEDY: return step_kind<KIND_KNAP_GREEDY>(block, d, lds, cur, init, s, start, stop);
    case KIND_TSPJ: return step_kind<KIND_TSPJ>(block, d, lds, cur, init, s, start, stop);
    case KIND_CONTINUOUS: return step_kind<KIND_CONTINUOUS>(block, d, lds, cur, init, s, start, stop);
    case KIND_JOBSHOP: return step_kind<KIND_JOBSHOP>(block, d, lds, cur, init, s, start, stop);
    case KIND_VRP_DAYS: return step_kind<KIND_VRP_DAYS>(block, d, lds, cur, init, s, start, stop);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_partition_one(const InstDesc& d, int buf, int mode, hipStream_t s, hipEvent_t start,
                                hipEvent_t stop)
{
    const int lds = partition_lds_bytes(d.Pi);
    ensure_lds(k_partition_one, lds);
    launch_maybe_timed(k_partition_one, dim3(d.islands * partition_parts(d.Pi)), dim3(kPartBlock), lds, s, start,
                       stop, d, buf, mode);
    return hipGetLastError();
}

hipError_t launch_partition_many(const InstDesc* d_descs, int num_inst, int blocks, int max_Pi, int buf, int mode,
                                 hipStream_t s)
{
    if (blocks <= 0) return hipSuccess;
    const int lds = partition_lds_bytes(max_Pi);
    ensure_lds(k_partition_many, lds);
    hipLaunchKernelGGL(k_partition_many, dim3(blocks), dim3(kPartBlock), lds, s, d_descs, num_inst, buf, mode);
    return hipGetLastError();
}

int tspj_ls_lds_bytes(int n, int nb)
{
    return decode_layout(KIND_TSPJ, n, 2 * n, nb, 0).bytes + 12 * n + 16;
}

namespace {

template <int BLOCK>
hipError_t ls_one(const InstDesc* d, int num, int items, int lds, hipStream_t s)
{
    ensure_lds(k_tspj_local_search<BLOCK>, lds);
    hipLaunchKernelGGL((k_tspj_local_search<BLOCK>), dim3(items, num), dim3(BLOCK), lds, s, d, num);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_tspj_local_search(const InstDesc* d_descs, int num_inst, int items, int block, int lds,
                                    hipStream_t s)
{
    if (num_inst <= 0 || items <= 0) return hipSuccess;
    switch (block) {
    case 64: return ls_one<64>(d_descs, num_inst, items, lds, s);
    case 128: return ls_one<128>(d_descs, num_inst, items, lds, s);
    case 256: return ls_one<256>(d_descs, num_inst, items, lds, s);
    case 512: return ls_one<512>(d_descs, num_inst, items, lds, s);
    case 1024: return ls_one<1024>(d_descs, num_inst, items, lds, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace brkga
