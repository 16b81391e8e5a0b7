// kernels_brkga.hip - per-island ranking and TSPJ local search for gfx950.
//
//   k_partition_*    stable sort of each island's fitness -> rank order;
//                    replaces thrust::sequence + sort_by_key (core/solver.hpp:416-419).
//   k_tspj_local_search  2-opt moves on each island's top elites
//                    (configs/tspj_config.hpp:77-135).
//
// k_partition_one takes the population descriptor by value;
// k_partition_many reads one descriptor per instance from device memory
// (batched and served populations).
#include "launch.hpp"
#include "launch_util.hpp"
#include "step_device.hpp"

namespace brkga {

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
// workgroup of the population bumps the generation counter (both parity
// slots) to t + 1 and raises the local-search flag when generation t + 1 is
// due; mode 2: the last workgroup lowers the flag; mode 3: the last
// workgroup sets both counter slots to t.
// B: workgroup size (kPartBlock, or 256 for islands of at most 256).
template <int B = kPartBlock>
__device__ __forceinline__ void partition_body(const InstDesc& d, int island, int part, int parts, int buf,
                                               int mode, int t, char* lds)
{
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
                    d.gen[0] = g;
                    d.gen[3] = g;
                } else if (mode == 3) {
                    d.gen[0] = t;
                    d.gen[3] = t;
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
    } else if (mode == 3) {
        t = *gen_slot(d, buf);
    }
    partition_body(d, blockIdx.x / parts, blockIdx.x % parts, parts, buf, mode, t, lds);
}

// Island `island` of member d of a fleet (k_partition_many / _list), ranked
// by a workgroup of B threads.
template <int B = kPartBlock>
__device__ __forceinline__ void partition_member(const InstDesc& d, int island, int buf, int mode, char* lds)
{
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
    partition_body<B>(d, island, 0, 1, b, mode, t, lds);
}

__global__ __launch_bounds__(kPartBlock) void k_partition_many(const InstDesc* __restrict__ descs, int num_inst,
                                                               int buf, int mode)
{
    extern __shared__ __align__(16) char lds[];
    const InstDesc& d = descs[find_instance_by_seg(descs, num_inst, blockIdx.x)];
    const int island = blockIdx.x - d.seg_begin;
    if (island >= d.islands) return;
    partition_member(d, island, buf, mode, lds);
}

// One workgroup per listed {member, island}: the islands of the members one
// stream lane of a fleet steps.  Islands of at most 256 are ranked by 256
// threads (one key each): a lane's ranking launch runs beside the other
// lanes' step launches, and four-wave workgroups find room on a CU sooner
// than sixteen-wave ones.
template <int B>
__global__ __launch_bounds__(B) void k_partition_list(const InstDesc* __restrict__ descs,
                                                      const int2* __restrict__ segs, int buf, int mode)
{
    extern __shared__ __align__(16) char lds[];
    const int2 sg = segs[blockIdx.x];
    partition_member<B>(descs[sg.x], sg.y, buf, mode, lds);
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
    [[clang::always_inline]] decode_block<KIND_TSPJ, BLOCK>(d.prob, lds, lay, nb, &s_fit, nullptr);
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

// Workgroups (x, p) clear population p's key buffers with float4 stores
// (grid-stride over x); workgroup (0, p) also sets its counters and
// copies its descriptor.
constexpr int kSetupBlock = 256;
constexpr int kSetupBlocksX = 32;

__global__ __launch_bounds__(kSetupBlock) void k_setup_pops(const PopSetup* __restrict__ list)
{
    const PopSetup& p = list[blockIdx.y];
    const size_t q = p.key_floats >> 2;
    const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (size_t i = (size_t)blockIdx.x * kSetupBlock + threadIdx.x; i < q; i += (size_t)gridDim.x * kSetupBlock) {
        reinterpret_cast<float4*>(p.keys[0])[i] = z;
        reinterpret_cast<float4*>(p.keys[1])[i] = z;
    }
    if (blockIdx.x == 0) {
        if (threadIdx.x < 4) p.gen[threadIdx.x] = (threadIdx.x == 0 || threadIdx.x == 3) ? p.gen_edge : 0;
        constexpr int kWords = (int)(sizeof(InstDesc) / sizeof(int));
        const int* src = reinterpret_cast<const int*>(&p.desc);
        int* dst = reinterpret_cast<int*>(p.d_desc);
        for (int w = threadIdx.x; w < kWords; w += kSetupBlock) dst[w] = src[w];
    }
}

hipError_t launch_setup_pops(const PopSetup* d_list, int count, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_setup_pops, dim3(kSetupBlocksX, count), dim3(kSetupBlock), 0, s, d_list);
    return hipGetLastError();
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

hipError_t launch_partition_list(const InstDesc* d_descs, const int2* d_segs, int count, int max_Pi, int buf,
                                 int mode, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    const int lds = partition_lds_bytes(max_Pi);
    if (max_Pi <= 256) {
        hipLaunchKernelGGL(k_partition_list<256>, dim3(count), dim3(256), lds, s, d_descs, d_segs, buf, mode);
    } else {
        ensure_lds(k_partition_list<kPartBlock>, lds);
        hipLaunchKernelGGL(k_partition_list<kPartBlock>, dim3(count), dim3(kPartBlock), lds, s, d_descs, d_segs, buf,
                           mode);
    }
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
