// kernels_brkga.hip - BRKGA generation step for gfx950.
//
//   k_evaluate      decode + fitness of caller-owned key rows: the
//                   evaluate_population_gpu hook (core/config.hpp:50-53).
//   k_brkga_step    one workgroup per child slot: elite / migrant row copy,
//                   mutant keys from Philox, or biased uniform crossover,
//                   then decode + fitness of the new row straight from LDS.
//                   Replaces brkga_generation_kernel (core/cuda_kernels.cuh:88-138)
//                   fused with the config fitness kernels, and the host-staged
//                   ring migration (core/solver.hpp:517-558).
//   k_partition     stable sort of each island's fitness -> rank order;
//                   replaces thrust::sequence + sort_by_key (core/solver.hpp:416-419).
//
// Key rows are [pop][ld] FP32 with ld a multiple of 32, so every row starts
// 128-byte aligned and is streamed as float4.
#include "launch.hpp"
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

template <int KIND, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_evaluate(ProblemDesc p, const float* __restrict__ keys, long long ld,
                                                    int count, float* __restrict__ fit, int* __restrict__ decoded)
{
    extern __shared__ __align__(16) char lds[];
    const int row = blockIdx.x;
    if (row >= count) return;
    const int L = p.chrom_len;
    const int nb = sort_buckets(p.n);
    const DecodeLayout lay = decode_layout(KIND, p.n, L, nb);
    float* k = reinterpret_cast<float*>(lds + lay.keys_off);
    const float* src = keys + (size_t)row * ld;
    for (int j = threadIdx.x; j < L; j += BLOCK) k[j] = src[j];
    __syncthreads();
    decode_block<KIND, BLOCK>(p, lds, lay, nb, fit + (size_t)row * p.num_obj,
                              decoded ? decoded + (size_t)row * L : nullptr);
}

template <int KIND, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_brkga_step(const InstDesc* __restrict__ descs, int num_inst, int init)
{
    extern __shared__ __align__(16) char lds[];
    const int b = blockIdx.x;
    const InstDesc& d = descs[find_instance_by_slot(descs, num_inst, b)];
    const int gs = b - d.slot_begin;
    if (gs >= d.pop) return;
    const int t = init ? 0 : *d.gen;
    if (!init && t >= d.gen_limit) return;

    const int Pi = d.Pi, Ei = d.Ei, Mi = d.Mi, I = d.islands;
    const int island = gs / Pi;
    const int s = gs - island * Pi;
    const int cur = t & 1;
    const int nxt = init ? 0 : cur ^ 1;
    const int M = d.prob.num_obj;
    const int ld = d.ld, L = d.L;
    const int nq = (L + 3) >> 2;
    float4* out = reinterpret_cast<float4*>(d.keys[nxt] + (size_t)gs * ld);
    float* fit_out = d.fit[nxt] + (size_t)gs * M;
    const int nb = d.nb;
    const DecodeLayout lay = decode_layout(KIND, d.prob.n, L, nb);
    float4* kl = reinterpret_cast<float4*>(lds + lay.keys_off);

    if (init) {
        for (int c = threadIdx.x; c < nq; c += BLOCK) {
            const U4 r = rng_block(d.seed_lo, d.seed_hi, island, 0, s, TAG_INIT, c);
            const float4 x = make_float4(unit_from_u32(r.x), unit_from_u32(r.y), unit_from_u32(r.z),
                                         unit_from_u32(r.w));
            out[c] = x;
            kl[c] = x;
        }
    } else {
        const int* perm = d.perm[cur];
        const float* keys_in = d.keys[cur];
        const uint32_t g = (uint32_t)t + 1u;
        const int k = d.mig_k;
        const bool migrate = I > 1 && d.mig_interval > 0 && k > 0 && t > 0 && (t % d.mig_interval) == 0;
        int src = -1;
        if (migrate && s >= Pi - k) {
            const int si = (island + I - 1) % I;
            src = si * Pi + perm[si * Pi + (s - (Pi - k))];
        } else if (s < Ei) {
            src = island * Pi + perm[island * Pi + s];
        }
        if (src >= 0) {
            // elite / migrant: the row and its fitness are copied unchanged
            const float4* in = reinterpret_cast<const float4*>(keys_in + (size_t)src * ld);
            for (int c = threadIdx.x; c < nq; c += BLOCK) out[c] = in[c];
            if (threadIdx.x < M) fit_out[threadIdx.x] = d.fit[cur][(size_t)src * M + threadIdx.x];
            return;
        }
        if (s >= Pi - Mi) {
            for (int c = threadIdx.x; c < nq; c += BLOCK) {
                const U4 r = rng_block(d.seed_lo, d.seed_hi, island, g, s, TAG_MUTANT, c);
                const float4 x = make_float4(unit_from_u32(r.x), unit_from_u32(r.y), unit_from_u32(r.z),
                                             unit_from_u32(r.w));
                out[c] = x;
                kl[c] = x;
            }
        } else {
            const U4 pr = rng_block(d.seed_lo, d.seed_hi, island, g, s, TAG_PARENTS, 0);
            const int ep = island * Pi + perm[island * Pi + (int)(pr.x % (uint32_t)Ei)];
            const int np = island * Pi + perm[island * Pi + Ei + (int)(pr.y % (uint32_t)(Pi - Ei))];
            const float4* pe = reinterpret_cast<const float4*>(keys_in + (size_t)ep * ld);
            const float4* pn = reinterpret_cast<const float4*>(keys_in + (size_t)np * ld);
            const float rho = d.rho;
            for (int c = threadIdx.x; c < nq; c += BLOCK) {
                const U4 r = rng_block(d.seed_lo, d.seed_hi, island, g, s, TAG_COINS, c);
                const float4 a = pe[c], bb = pn[c];
                float4 x;
                x.x = unit_from_u32(r.x) < rho ? a.x : bb.x;
                x.y = unit_from_u32(r.y) < rho ? a.y : bb.y;
                x.z = unit_from_u32(r.z) < rho ? a.z : bb.z;
                x.w = unit_from_u32(r.w) < rho ? a.w : bb.w;
                out[c] = x;
                kl[c] = x;
            }
        }
    }
    __syncthreads();
    decode_block<KIND, BLOCK>(d.prob, lds, lay, nb, fit_out, nullptr);
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_partition(const InstDesc* __restrict__ descs, int num_inst, int init)
{
    extern __shared__ __align__(16) char lds[];
    unsigned long long* sk = reinterpret_cast<unsigned long long*>(lds);
    const int b = blockIdx.x;
    const InstDesc& d = descs[find_instance_by_seg(descs, num_inst, b)];
    const int island = b - d.seg_begin;
    if (island >= d.islands) return;
    const int t = init ? 0 : *d.gen;
    if (!init && t >= d.gen_limit) return;
    const int buf = init ? 0 : ((t + 1) & 1);
    const int Pi = d.Pi;
    const int Pn = pow2_at_least(Pi);
    const int M = d.prob.num_obj;
    const float* fit = d.fit[buf] + (size_t)island * Pi * M;
    const int maximize = d.prob.maximize;
    for (int i = threadIdx.x; i < Pn; i += BLOCK) {
        sk[i] = i < Pi ? ((unsigned long long)fitness_order_key(fit[(size_t)i * M], maximize) << 32) | (unsigned)i
                       : ~0ull;
    }
    __syncthreads();
    for (int k = 2; k <= Pn; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < (Pn >> 1); i += BLOCK) {
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
    int* perm = d.perm[buf] + (size_t)island * Pi;
    for (int i = threadIdx.x; i < Pi; i += BLOCK) perm[i] = (int)(sk[i] & 0xFFFFFFFFull);
    if (!init) {
        __syncthreads();
        if (threadIdx.x == 0) {
            __threadfence();
            const int old = atomicAdd(d.done, 1);
            if (old == d.islands - 1) {
                *d.done = 0;
                *d.gen = t + 1;
            }
        }
    }
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

template <int KIND, int BLOCK>
hipError_t step_one(const InstDesc* d, int num, int blocks, int lds, int init, hipStream_t s)
{
    ensure_lds(k_brkga_step<KIND, BLOCK>, lds);
    hipLaunchKernelGGL((k_brkga_step<KIND, BLOCK>), dim3(blocks), dim3(BLOCK), lds, s, d, num, init);
    return hipGetLastError();
}

template <int KIND>
hipError_t step_kind(int block, const InstDesc* d, int num, int blocks, int lds, int init, hipStream_t s)
{
    switch (block) {
    case 64: return step_one<KIND, 64>(d, num, blocks, lds, init, s);
    case 128: return step_one<KIND, 128>(d, num, blocks, lds, init, s);
    case 256: return step_one<KIND, 256>(d, num, blocks, lds, init, s);
    case 512: return step_one<KIND, 512>(d, num, blocks, lds, init, s);
    case 1024: return step_one<KIND, 1024>(d, num, blocks, lds, init, s);
    default: return hipErrorInvalidValue;
    }
}

template <int BLOCK>
hipError_t part_one(const InstDesc* d, int num, int blocks, int lds, int init, hipStream_t s)
{
    ensure_lds(k_partition<BLOCK>, lds);
    hipLaunchKernelGGL((k_partition<BLOCK>), dim3(blocks), dim3(BLOCK), lds, s, d, num, init);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_evaluate(const ProblemDesc& p, const float* keys, long long ld, int count, float* fit,
                           int* decoded, int block, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    const int lds = decode_layout(p.kind, p.n, p.chrom_len, sort_buckets(p.n)).bytes;
    switch (p.kind) {
    case KIND_TSP_COORDS: return eval_kind<KIND_TSP_COORDS>(block, p, keys, ld, count, fit, decoded, lds, s);
    case KIND_TSP_MATRIX: return eval_kind<KIND_TSP_MATRIX>(block, p, keys, ld, count, fit, decoded, lds, s);
    case KIND_KNAP_THRESHOLD:
        return eval_kind<KIND_KNAP_THRESHOLD>(block, p, keys, ld, count, fit, decoded, lds, s);
    case KIND_KNAP_GREEDY: return eval_kind<KIND_KNAP_GREEDY>(block, p, keys, ld, count, fit, decoded, lds, s);
    case KIND_TSPJ: return eval_kind<KIND_TSPJ>(block, p, keys, ld, count, fit, decoded, lds, s);
    case KIND_BITSP_COORDS: return eval_kind<KIND_BITSP_COORDS>(block, p, keys, ld, count, fit, decoded, lds, s);
    case KIND_BITSP_MATRIX: return eval_kind<KIND_BITSP_MATRIX>(block, p, keys, ld, count, fit, decoded, lds, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_brkga_step(int kind, int block, const InstDesc* d_descs, int num_inst, int blocks, int lds,
                             int init, hipStream_t s)
{
    if (blocks <= 0) return hipSuccess;
    switch (kind) {
    case KIND_TSP_COORDS: return step_kind<KIND_TSP_COORDS>(block, d_descs, num_inst, blocks, lds, init, s);
    case KIND_TSP_MATRIX: return step_kind<KIND_TSP_MATRIX>(block, d_descs, num_inst, blocks, lds, init, s);
    case KIND_KNAP_THRESHOLD:
        return step_kind<KIND_KNAP_THRESHOLD>(block, d_descs, num_inst, blocks, lds, init, s);
    case KIND_KNAP_GREEDY: return step_kind<KIND_KNAP_GREEDY>(block, d_descs, num_inst, blocks, lds, init, s);
    case KIND_TSPJ: return step_kind<KIND_TSPJ>(block, d_descs, num_inst, blocks, lds, init, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_partition(int block, const InstDesc* d_descs, int num_inst, int blocks, int lds, int init,
                            hipStream_t s)
{
    if (blocks <= 0) return hipSuccess;
    switch (block) {
    case 64: return part_one<64>(d_descs, num_inst, blocks, lds, init, s);
    case 128: return part_one<128>(d_descs, num_inst, blocks, lds, init, s);
    case 256: return part_one<256>(d_descs, num_inst, blocks, lds, init, s);
    case 512: return part_one<512>(d_descs, num_inst, blocks, lds, init, s);
    case 1024: return part_one<1024>(d_descs, num_inst, blocks, lds, init, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace brkga
