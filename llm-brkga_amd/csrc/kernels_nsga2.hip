// kernels_nsga2.hip - NSGA-II generation on the GPU (two objectives).
//
// The reference runs NSGA-II on the host only (core/population.hpp:334-517,
// core/solver.hpp:751-1091).  Per generation the engine runs:
//   k_nsga2_offspring  one workgroup per child: two binary tournaments,
//                      uniform crossover, per-gene resampling, decode of the
//                      child's two objectives from LDS;
//   k_nsga2_select     one workgroup per population over parents + children:
//                      domination counts and front peeling, crowding
//                      distance per front, survivor order;
//   k_nsga2_gather     survivor rows copied into the next buffer, four rows
//                      per workgroup.
// The selection advances the step counter (one workgroup per population, no
// atomics) and leaves d.done = 1 when the gather has rows to copy.
// Every kernel of the chain raises its waves' issue priority (s_setprio 3):
// in a served or batched fleet the chain runs on a stream lane of its own
// beside the BRKGA queue steps, its few latency-bound workgroups share CUs
// with their persistent ones, and at the default priority the CU's arbiter
// left them waiting (served mix at saturation 1320 -> 1510 solves/s).  Alone
// on the device the priority changes nothing.
// Buffers: keys[b] holds 2N rows (parents [0,N), children [N,2N)).
#include "launch.hpp"
#include "decoders.hpp"
#include "step_device.hpp"

namespace brkga {

namespace {

// Instance owning row b: instances are laid out with slot_begin = prefix
// of their populations.
__device__ __forceinline__ int find_inst(const InstDesc* __restrict__ d, int num, int b)
{
    int lo = 0, hi = num - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (d[mid].slot_begin <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ uint32_t order_key_asc(float f)
{
    if (f != f) return 0xFFFFFFFFu;
    if (f == 0.0f) f = 0.0f;
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ bool dominates2(float a0, float a1, float b0, float b1)
{
    // individual.hpp:99-114 for two objectives (minimisation)
    if (a0 > b0 || a1 > b1) return false;
    return a0 < b0 || a1 < b1;
}

template <int BLOCK>
__device__ void bitonic_u64(unsigned long long* sk, int n)
{
    for (int k = 2; k <= n; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < (n >> 1); i += BLOCK) {
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
}

}  // namespace

template <int KIND, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_nsga2_offspring(const InstDesc* __restrict__ descs, int num_inst, int init)
{
    __builtin_amdgcn_s_setprio(3);
    extern __shared__ __align__(16) char lds[];
    const int b = blockIdx.x;
    const InstDesc& d = descs[find_inst(descs, num_inst, b)];
    const int c = b - d.slot_begin;  // child (or, at init, parent) index
    const int N = d.pop;
    if (c >= N) return;
    // step counter -1 (or the init flag): build the first parents
    const int t = init ? -1 : *d.gen;
    if (t >= d.gen_limit) return;
    init = t < 0;
    const int cur = t & 1;
    const int row = init ? c : N + c;
    const int ld = d.ld, L = d.L;
    const int nq = (L + 3) >> 2;
    const int buf = init ? 0 : cur;
    float4* out = reinterpret_cast<float4*>(d.keys[buf] + (size_t)row * ld);
    const int nb = d.nb;
    const DecodeLayout lay = decode_layout(KIND, d.prob.n, L, nb, d.prob.aux);
    float4* kl = reinterpret_cast<float4*>(lds + lay.keys_off);

    if (init) {
        for (int q = threadIdx.x; q < nq; q += BLOCK) {
            const U4 r = rng_block(d.seed_lo, d.seed_hi, 0, 0, c, TAG_INIT, q);
            const float4 x = make_float4(unit_from_u32(r.x), unit_from_u32(r.y), unit_from_u32(r.z),
                                         unit_from_u32(r.w));
            out[q] = x;
            kl[q] = x;
        }
    } else {
        const uint32_t g = (uint32_t)t + 1u;
        const int* rank = d.rank[cur];
        const float* crowd = d.crowd[cur];
        const U4 pr = rng_block(d.seed_lo, d.seed_hi, 0, g, c, TAG_PARENTS, 0);
        // binary tournaments (solver.hpp:1042-1056)
        auto tour = [&](int a, int bb) {
            if (rank[a] < rank[bb]) return a;
            if (rank[bb] < rank[a]) return bb;
            if (crowd[a] > crowd[bb]) return a;
            return bb;
        };
        const int p1 = tour((int)(pr.x % (uint32_t)N), (int)(pr.y % (uint32_t)N));
        const int p2 = tour((int)(pr.z % (uint32_t)N), (int)(pr.w % (uint32_t)N));
        const float4* a4 = reinterpret_cast<const float4*>(d.keys[cur] + (size_t)p1 * ld);
        const float4* b4 = reinterpret_cast<const float4*>(d.keys[cur] + (size_t)p2 * ld);
        const float rho = d.rho, mut = d.mut_rate;
        for (int q = threadIdx.x; q < nq; q += BLOCK) {
            const U4 rc = rng_block(d.seed_lo, d.seed_hi, 0, g, c, TAG_COINS, q);
            const U4 rm = rng_block(d.seed_lo, d.seed_hi, 0, g, c, TAG_MUTCOIN, q);
            const U4 rk = rng_block(d.seed_lo, d.seed_hi, 0, g, c, TAG_MUTANT, q);
            const float4 pa = a4[q], pb = b4[q];
            float4 x;
            x.x = unit_from_u32(rc.x) < rho ? pa.x : pb.x;
            x.y = unit_from_u32(rc.y) < rho ? pa.y : pb.y;
            x.z = unit_from_u32(rc.z) < rho ? pa.z : pb.z;
            x.w = unit_from_u32(rc.w) < rho ? pa.w : pb.w;
            x.x = unit_from_u32(rm.x) < mut ? unit_from_u32(rk.x) : x.x;
            x.y = unit_from_u32(rm.y) < mut ? unit_from_u32(rk.y) : x.y;
            x.z = unit_from_u32(rm.z) < mut ? unit_from_u32(rk.z) : x.z;
            x.w = unit_from_u32(rm.w) < mut ? unit_from_u32(rk.w) : x.w;
            out[q] = x;
            kl[q] = x;
        }
    }
    if constexpr (KIND == KIND_HOST) {
        return;  // the caller evaluates the row on the host
    } else {
        __syncthreads();
        decode_block<KIND, BLOCK>(d.prob, lds, lay, nb, d.fit[buf] + (size_t)row * 2, nullptr);
    }
}

// LDS layout of the selection block for C = up to 2N individuals (xch: the
// exchange buffers of the register sorting network, for C <= 1024).
struct SelectLayout {
    int o0, o1, rk, cnt, lst, cr, fst, fs, sk, xch, bytes;
};

__host__ __device__ inline SelectLayout select_layout(int C)
{
    SelectLayout s;
    const int Cp = pow2_at_least(C < 2 ? 2 : C);
    int off = 0;
    s.o0 = off; off += C * 4;
    s.o1 = off; off += C * 4;
    s.rk = off; off += C * 4;
    s.cnt = off; off += C * 4;
    s.lst = off; off += (C + 1) * 4;
    s.cr = off; off += C * 4;
    s.fst = off; off += (C + 2) * 4;
    s.fs = off; off += (C + 2) * 4;
    off = round_up_int(off, 16);
    s.sk = off; off += Cp * 8;
    s.xch = off; off += Cp <= 1024 ? island_rank_bytes(Cp < 64 ? 64 : Cp) : 0;
    s.bytes = round_up_int(off, 16);
    return s;
}

constexpr int kSelectBlock = 1024;

// Ascending sort of the Cp (power of two) keys sk[0, Cp) of the selection
// block, left in sk.  Up to kSelectBlock keys: one key per thread through
// the register bitonic network of the island ranking (in-wave exchanges by
// DPP / swizzle / permute, one LDS exchange and barrier per step across
// waves: 10 barriers for 1024 keys); larger: the LDS network (a barrier per
// stage).  Ends with a barrier.
__device__ void select_sort(unsigned long long* sk, int Cp, unsigned long long* xch)
{
    constexpr int B = kSelectBlock;
    if (Cp > B) {
        bitonic_u64<B>(sk, Cp);
        return;
    }
    const int t = threadIdx.x;
    unsigned long long v = t < Cp ? sk[t] : ~0ull;
    if (Cp <= 64) v = block_bitonic_sort<64>(v, xch);
    else if (Cp <= 128) v = block_bitonic_sort<128>(v, xch);
    else if (Cp <= 256) v = block_bitonic_sort<256>(v, xch);
    else if (Cp <= 512) v = block_bitonic_sort<512>(v, xch);
    else v = block_bitonic_sort<1024>(v, xch);
    __syncthreads();  // every exchange read is done before sk is rewritten
    if (t < Cp) sk[t] = v;
    __syncthreads();
}
constexpr int kDomBlock = 256;
static_assert(kDomRows <= kSelectBlock, "one dominance row per selection thread");

// Dominated-by bitsets of the C rows the selection of every listed
// population ranks (C = 2N, or N at generation 0): bit j of word w of row q
// is set when row 32 w + j dominates row q.  One thread per (row, word);
// grid (row-word chunks of kDomBlock, populations).  Populations without
// bitsets (more than kDomRows rows) are skipped.
__global__ __launch_bounds__(kDomBlock) void k_nsga2_dominance(const InstDesc* __restrict__ descs, int mode)
{
    __builtin_amdgcn_s_setprio(3);
    __shared__ float2 o[kDomRows];
    const InstDesc& d = descs[blockIdx.y];
    if (!d.dom) return;
    const int N = d.pop;
    int cur, C;
    if (mode == 2) {
        // the N members of the current generation after a diversity injection
        if (d.div_interval <= 0 || *d.ls_flag == 0) return;
        cur = *d.gen & 1;
        C = N;
    } else {
        const int t = mode ? -1 : *d.gen;
        if (t >= d.gen_limit) return;
        const int first = t < 0;
        cur = first ? 0 : (t & 1);
        C = first ? N : 2 * N;
    }
    const int W = (C + 31) >> 5;
    if ((int)blockIdx.x * kDomBlock >= C * W) return;
    const float* fit = d.fit[cur];
    for (int i = threadIdx.x; i < C; i += kDomBlock) o[i] = make_float2(fit[2 * i], fit[2 * i + 1]);
    __syncthreads();
    const int idx = blockIdx.x * kDomBlock + threadIdx.x;
    if (idx >= C * W) return;
    const int q = idx / W, w = idx - q * W;
    const float2 b = o[q];
    const int j0 = w << 5;
    const int nj = C - j0 < 32 ? C - j0 : 32;
    unsigned bits = 0;
    for (int j = 0; j < nj; ++j) {
        const float2 a = o[j0 + j];
        bits |= (dominates2(a.x, a.y, b.x, b.y) ? 1u : 0u) << j;
    }
    d.dom[(size_t)q * W + w] = bits;
}

__global__ __launch_bounds__(kSelectBlock) void k_nsga2_select(const InstDesc* __restrict__ descs, int mode)
{
    __builtin_amdgcn_s_setprio(3);
    constexpr int B = kSelectBlock;
    extern __shared__ __align__(16) char lds[];
    const InstDesc& d = descs[blockIdx.x];
    const int N = d.pop;
    const int forced = mode == 1;  // generation 0 built by the host path: the counter stays
    int t, init, cur;
    if (mode == 2) {
        // rank and crowd the N members of the current generation again
        // (population.hpp:174-176); its counter and the gather flag stay
        if (d.div_interval <= 0 || *d.ls_flag == 0) return;
        t = *d.gen;
        init = 1;
        cur = t & 1;
    } else {
        t = mode ? -1 : *d.gen;
        if (t >= d.gen_limit) {
            if (threadIdx.x == 0) *d.done = 0;  // nothing for the gather
            return;
        }
        init = t < 0;
        cur = init ? 0 : (t & 1);
    }
    const int C = init ? N : 2 * N;
    const int Cp = pow2_at_least(C < 2 ? 2 : C);
    const SelectLayout L = select_layout(C);
    float* o0 = reinterpret_cast<float*>(lds + L.o0);
    float* o1 = reinterpret_cast<float*>(lds + L.o1);
    int* rk = reinterpret_cast<int*>(lds + L.rk);
    int* cnt = reinterpret_cast<int*>(lds + L.cnt);
    int* lst = reinterpret_cast<int*>(lds + L.lst);  // [0] = size, then members
    float* cr = reinterpret_cast<float*>(lds + L.cr);
    int* fst = reinterpret_cast<int*>(lds + L.fst);  // first sorted position of each front
    int* fs = reinterpret_cast<int*>(lds + L.fs);    // front sizes
    unsigned long long* sk = reinterpret_cast<unsigned long long*>(lds + L.sk);
    unsigned long long* xch = reinterpret_cast<unsigned long long*>(lds + L.xch);

    const float* fit = d.fit[cur];
    for (int i = threadIdx.x; i < C; i += B) {
        o0[i] = fit[2 * i];
        o1[i] = fit[2 * i + 1];
        rk[i] = -1;
    }
    __syncthreads();
    int r = 0;
    if (d.dom && C <= kDomRows) {
        // Domination counts and front peeling (population.hpp:334-440) from
        // the dominated-by bitsets: row q's count is the popcount of its
        // bitset, and a front takes away popcount(bitset & front members).
        // Thread q holds row q; wave k's ballot gives the front bits of rows
        // [64k, 64k + 64).
        __shared__ unsigned front[kSelectBlock / 32];
        const int W = (C + 31) >> 5;
        const int q = threadIdx.x;
        unsigned dq[kDomRows / 32];
        int c = 0;
#pragma unroll
        for (int w = 0; w < kDomRows / 32; ++w) {
            dq[w] = (q < C && w < W) ? d.dom[(size_t)q * W + w] : 0u;
            c += __popc(dq[w]);
        }
        if (q < C && c == 0) rk[q] = 0;
        __syncthreads();
        while (true) {
            const unsigned long long bal = __ballot(q < C && rk[q] == r);
            if ((q & (kWave - 1)) == 0) {
                front[2 * (q / kWave)] = (unsigned)bal;
                front[2 * (q / kWave) + 1] = (unsigned)(bal >> 32);
            }
            __syncthreads();
            int m = 0;
            for (int w = 0; w < W; ++w) m += __popc(front[w]);
            if (m == 0) break;
            if (q == 0) fs[r] = m;
            if (q < C && rk[q] < 0) {
                int dec = 0;
#pragma unroll
                for (int w = 0; w < kDomRows / 32; ++w) dec += __popc(dq[w] & front[w]);
                c -= dec;
                if (dec && c == 0) rk[q] = r + 1;
            }
            ++r;
            __syncthreads();
        }
    } else {
    // domination counts (population.hpp:409-440)
    for (int i = threadIdx.x; i < C; i += B) {
        const float b0 = o0[i], b1 = o1[i];
        int c = 0;
        for (int j = 0; j < C; ++j) c += dominates2(o0[j], o1[j], b0, b1) ? 1 : 0;
        cnt[i] = c;
        if (c == 0) rk[i] = 0;
    }
    __syncthreads();
    // front peeling (population.hpp:334-369)
    while (true) {
        if (threadIdx.x == 0) lst[0] = 0;
        __syncthreads();
        for (int i = threadIdx.x; i < C; i += B)
            if (rk[i] == r) lst[1 + atomicAdd(&lst[0], 1)] = i;
        __syncthreads();
        const int m = lst[0];
        if (m == 0) break;
        if (threadIdx.x == 0) fs[r] = m;
        for (int q = threadIdx.x; q < C; q += B) {
            if (rk[q] >= 0) continue;
            const float b0 = o0[q], b1 = o1[q];
            int dec = 0;
            for (int a = 0; a < m; ++a) {
                const int p = lst[1 + a];
                dec += dominates2(o0[p], o1[p], b0, b1) ? 1 : 0;
            }
            if (dec) {
                cnt[q] -= dec;
                if (cnt[q] == 0) rk[q] = r + 1;
            }
        }
        ++r;
        __syncthreads();
    }
    }
    const int nfront = r;
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int f = 0; f < nfront; ++f) {
            fst[f] = acc;
            acc += fs[f];
        }
    }
    // crowding distance (population.hpp:453-489); per objective the front
    // is ordered by (objective, index)
    for (int i = threadIdx.x; i < C; i += B) cr[i] = 0.0f;
    __syncthreads();
    for (int m = 0; m < 2; ++m) {
        const float* om = m ? o1 : o0;
        for (int i = threadIdx.x; i < Cp; i += B)
            sk[i] = i < C ? ((unsigned long long)rk[i] << 44) | ((unsigned long long)order_key_asc(om[i]) << 12) |
                                (unsigned long long)i
                          : ~0ull;
        __syncthreads();
        select_sort(sk, Cp, xch);
        for (int q = threadIdx.x; q < C; q += B) {
            const int e = (int)(sk[q] & 0xFFFull);
            const int re = rk[e];
            const int size = fs[re];
            if (size <= 2) {
                if (m == 1) cr[e] = INFINITY;
                continue;
            }
            // the front occupies sorted positions [start, start + size)
            const int start = fst[re];
            const bool first = q == start;
            const bool last = q == start + size - 1;
            const int lo = (int)(sk[start] & 0xFFFull);
            const int hi = (int)(sk[start + size - 1] & 0xFFFull);
            const float range = __fsub_rn(om[hi], om[lo]);
            // boundaries become infinite; interior members add the normalised
            // gap between their neighbours when the range is positive
            float acc = cr[e];
            if (first || last) {
                acc = INFINITY;
            } else if (range > 0.0f) {
                const float gap = __fsub_rn(om[(int)(sk[q + 1] & 0xFFFull)], om[(int)(sk[q - 1] & 0xFFFull)]);
                acc = __fadd_rn(acc, __fdiv_rn(gap, range));
            }
            cr[e] = acc;
        }
        __syncthreads();
    }
    // after the m = 1 pass, sk holds (rank, obj1, index) order
    const int nxt = cur ^ 1;
    if (init) {
        for (int i = threadIdx.x; i < C; i += B) {
            d.rank[cur][i] = rk[i];
            d.crowd[cur][i] = cr[i];
        }
        if (threadIdx.x == 0) {
            if (mode == 2) {
                *d.ls_flag = 0;  // every thread read it before the first barrier
            } else if (!forced) {
                // counter -1: generation 0 is complete
                *d.done = 0;
                *d.gen = 0;
            }
        }
        return;
    }
    // survivor order (population.hpp:491-517)
    int full = 0, cum = 0;
    for (int f = 0; f < nfront; ++f) {
        if (cum + fs[f] <= N) {
            cum += fs[f];
            full = f + 1;
        } else {
            break;
        }
    }
    int* order = d.perm[cur];
    // whole fronts: position = lower fronts + position in (obj1, index)
    // order for fronts of more than two, index order otherwise
    for (int q = threadIdx.x; q < C; q += B) {
        const int e = (int)(sk[q] & 0xFFFull);
        const int re = rk[e];
        if (re >= full) continue;
        const int base = fst[re];
        int pos;
        if (fs[re] > 2) {
            pos = q - base;
        } else if (fs[re] == 2) {
            const int o = (int)(sk[q == base ? q + 1 : q - 1] & 0xFFFull);
            pos = o < e ? 1 : 0;
        } else {
            pos = 0;
        }
        const int out = base + pos;
        order[out] = e;
        d.rank[nxt][out] = re;
        d.crowd[nxt][out] = cr[e];
    }
    __syncthreads();
    if (cum < N) {
        // overflowing front `full`: crowding descending, ties by index
        for (int i = threadIdx.x; i < Cp; i += B) {
            unsigned long long key = ~0ull;
            if (i < C && rk[i] == full) {
                const uint32_t ok = ~order_key_asc(cr[i]);
                key = ((unsigned long long)ok << 12) | (unsigned long long)i;
            }
            sk[i] = key;
        }
        __syncthreads();
        select_sort(sk, Cp, xch);
        for (int q = threadIdx.x; q < N - cum; q += B) {
            const int e = (int)(sk[q] & 0xFFFull);
            order[cum + q] = e;
            d.rank[nxt][cum + q] = full;
            d.crowd[nxt][cum + q] = cr[e];
        }
    }
    // every thread read *d.gen before the first barrier above
    if (threadIdx.x == 0) {
        *d.done = 1;
        *d.gen = t + 1;
    }
}

// ----------------------------------------------- diversity injection --

constexpr int kDivBlock = 1024;

// Whether two members count as one (population.hpp:139-146): the squared
// Euclidean distance of their FP32 objectives, from FP64 differences, is
// below eps2, the smallest double whose correctly rounded square root
// reaches 0.001.
__device__ __forceinline__ bool objectives_close(float2 a, float2 b, double eps2)
{
#pragma clang fp contract(off)
    const double d0 = (double)a.x - (double)b.x;
    const double d1 = (double)a.y - (double)b.y;
    return d0 * d0 + d1 * d1 < eps2;
}

// Injection after the gather of step t (t % div_interval == 0), one
// workgroup per population (Population::ensure_minimum_diversity,
// population.hpp:130-182).  Members are taken in order; a member is kept
// when it is not close to any member kept before it.  Candidates go 64 at a
// time: every wave tests the 64 against its share of the members already
// kept, then wave 0 settles the 64 among themselves in order (a candidate
// kept here rules out the later ones close to it).  When fewer than 70% are
// kept (10 U < 7 N, which is the reference's (double)U / N < 0.70), the
// kept rows move, in order, to rows [0, U) of the new generation (through
// its child rows [N, N + U), unused until the next step) and *ls_flag := U
// for the refill and the ranking that follow.
__global__ __launch_bounds__(kDivBlock) void k_nsga2_diversity(const InstDesc* __restrict__ descs)
{
    extern __shared__ __align__(16) char lds[];
    __shared__ unsigned kill[2];
    __shared__ int kept_count;
    const InstDesc& d = descs[blockIdx.x];
    if (d.div_interval <= 0) return;
    const int t1 = *d.gen;  // the generation step t just built: t1 = t + 1
    if (*d.done == 0 || t1 < 1 || (t1 - 1) % d.div_interval != 0) return;
    const int N = d.pop, b = t1 & 1;
    const int tid = threadIdx.x, wid = tid / kWave, lane = tid & (kWave - 1);
    constexpr int W = kDivBlock / kWave;
    float2* o = reinterpret_cast<float2*>(lds);
    int* kept = reinterpret_cast<int*>(lds + 8 * N);
    const float* fit = d.fit[b];
    for (int i = tid; i < N; i += kDivBlock) o[i] = make_float2(fit[2 * i], fit[2 * i + 1]);
    if (tid == 0) {
        kill[0] = kill[1] = 0u;
        kept_count = 0;
    }
    __syncthreads();
    const double eps2 = d.div_eps2;
    for (int c0 = 0; c0 < N; c0 += kWave) {
        const int i = c0 + lane;
        const bool valid = i < N;
        const float2 oi = valid ? o[i] : make_float2(0.0f, 0.0f);
        const int u = kept_count;
        bool close = false;
        if (valid)
            for (int k = wid; k < u && !close; k += W) close = objectives_close(oi, o[kept[k]], eps2);
        const unsigned long long bal = __ballot(close);
        if (lane == 0 && bal) {
            atomicOr(&kill[0], (unsigned)bal);
            atomicOr(&kill[1], (unsigned)(bal >> 32));
        }
        __syncthreads();
        if (wid == 0) {
            unsigned long long km = ((unsigned long long)kill[1] << 32) | kill[0];
            int nu = u;
            const int lim = N - c0 < kWave ? N - c0 : kWave;
            for (int l = 0; l < lim; ++l) {
                if ((km >> l) & 1ull) continue;
                if (lane == 0) kept[nu] = c0 + l;
                ++nu;
                const float2 ol = o[c0 + l];
                km |= __ballot(lane > l && valid && objectives_close(oi, ol, eps2));
            }
            if (lane == 0) {
                kept_count = nu;
                kill[0] = kill[1] = 0u;
            }
        }
        __syncthreads();
    }
    const int U = kept_count;
    if (10 * U >= 7 * N) return;  // at least 70% unique: the population stays
    const int nq = (d.L + 3) >> 2, ld4 = d.ld >> 2;
    float4* rows = reinterpret_cast<float4*>(d.keys[b]);
    for (int idx = tid; idx < U * nq; idx += kDivBlock) {
        const int r = idx / nq, q = idx - r * nq;
        rows[(size_t)(N + r) * ld4 + q] = rows[(size_t)kept[r] * ld4 + q];
    }
    __syncthreads();
    for (int idx = tid; idx < U * nq; idx += kDivBlock) {
        const int r = idx / nq, q = idx - r * nq;
        rows[(size_t)r * ld4 + q] = rows[(size_t)(N + r) * ld4 + q];
    }
    float* fo = d.fit[b];
    for (int r = tid; r < U; r += kDivBlock) {
        const float2 v = o[kept[r]];
        fo[2 * r] = v.x;
        fo[2 * r + 1] = v.y;
    }
    if (tid == 0) *d.ls_flag = U;
}

// New members of an injection: row r >= U of the current generation gets
// fresh keys from Philox stream (0, t + 1, r, DIVERSITY) (the reference's
// Individual::randomize, population.hpp:166-171) and is decoded.  Blocks
// as the offspring launch; rows of populations without an injection in
// progress are skipped.
template <int KIND, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_nsga2_refill(const InstDesc* __restrict__ descs, int num_inst)
{
    extern __shared__ __align__(16) char lds[];
    const int bidx = blockIdx.x;
    const InstDesc& d = descs[find_inst(descs, num_inst, bidx)];
    const int r = bidx - d.slot_begin;
    if (r >= d.pop || d.div_interval <= 0) return;
    const int U = *d.ls_flag;
    if (U == 0 || r < U) return;
    const int t1 = *d.gen, buf = t1 & 1;
    const int L = d.L, nq = (L + 3) >> 2, nb = d.nb;
    const DecodeLayout lay = decode_layout(KIND, d.prob.n, L, nb, d.prob.aux);
    float4* kl = reinterpret_cast<float4*>(lds + lay.keys_off);
    float4* out = reinterpret_cast<float4*>(d.keys[buf] + (size_t)r * d.ld);
    for (int q = threadIdx.x; q < nq; q += BLOCK) {
        const U4 w = rng_block(d.seed_lo, d.seed_hi, 0, (uint32_t)t1, (uint32_t)r, TAG_DIVERSITY, (uint32_t)q);
        const float4 x = make_float4(unit_from_u32(w.x), unit_from_u32(w.y), unit_from_u32(w.z), unit_from_u32(w.w));
        out[q] = x;
        kl[q] = x;
    }
    __syncthreads();
    decode_block<KIND, BLOCK>(d.prob, lds, lay, nb, d.fit[buf] + (size_t)r * 2, nullptr);
}

constexpr int kGatherRows = 4;  // rows per workgroup, one wave each

// Copies survivor i (rank order from the selection) of buffer t-1 into row
// i of buffer t, t being the counter the selection just advanced.
__global__ __launch_bounds__(kGatherRows * kWave) void k_nsga2_gather(const InstDesc* __restrict__ descs,
                                                                      int num_inst, int total_rows)
{
    __builtin_amdgcn_s_setprio(3);
    const int b = blockIdx.x * kGatherRows + threadIdx.x / kWave;
    if (b >= total_rows) return;
    const InstDesc& d = descs[find_inst(descs, num_inst, b)];
    const int i = b - d.slot_begin;
    if (i >= d.pop || *d.done == 0) return;
    const int t = *d.gen;
    const int cur = (t - 1) & 1, nxt = t & 1;
    const int src = d.perm[cur][i];
    const int nq = (d.L + 3) >> 2;
    const int lane = threadIdx.x & (kWave - 1);
    const float4* in = reinterpret_cast<const float4*>(d.keys[cur] + (size_t)src * d.ld);
    float4* out = reinterpret_cast<float4*>(d.keys[nxt] + (size_t)i * d.ld);
    for (int q = lane; q < nq; q += kWave) out[q] = in[q];
    if (lane < 2) d.fit[nxt][2 * i + lane] = d.fit[cur][2 * src + lane];
}

// ------------------------------------------------------------- launchers --

namespace {

template <int KIND, int BLOCK>
hipError_t off_one(const InstDesc* d, int num, int blocks, int lds, int init, hipStream_t s)
{
    if (lds > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_nsga2_offspring<KIND, BLOCK>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    hipLaunchKernelGGL((k_nsga2_offspring<KIND, BLOCK>), dim3(blocks), dim3(BLOCK), lds, s, d, num, init);
    return hipGetLastError();
}

template <int KIND>
hipError_t off_kind(int block, const InstDesc* d, int num, int blocks, int lds, int init, hipStream_t s)
{
    switch (block) {
    case 64: return off_one<KIND, 64>(d, num, blocks, lds, init, s);
    case 128: return off_one<KIND, 128>(d, num, blocks, lds, init, s);
    case 256: return off_one<KIND, 256>(d, num, blocks, lds, init, s);
    case 512: return off_one<KIND, 512>(d, num, blocks, lds, init, s);
    case 1024: return off_one<KIND, 1024>(d, num, blocks, lds, init, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_nsga2_offspring(int kind, int block, const InstDesc* d_descs, int num_inst, int blocks, int lds,
                                  int init, hipStream_t s)
{
    if (blocks <= 0) return hipSuccess;
    switch (kind) {
    case KIND_BITSP_COORDS: return off_kind<KIND_BITSP_COORDS>(block, d_descs, num_inst, blocks, lds, init, s);
    case KIND_BITSP_MATRIX: return off_kind<KIND_BITSP_MATRIX>(block, d_descs, num_inst, blocks, lds, init, s);
    case KIND_ZDT1: return off_kind<KIND_ZDT1>(block, d_descs, num_inst, blocks, lds, init, s);
    case KIND_MOVRP: return off_kind<KIND_MOVRP>(block, d_descs, num_inst, blocks, lds, init, s);
    case KIND_HOST: return off_kind<KIND_HOST>(block, d_descs, num_inst, blocks, lds, init, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_nsga2_select(const InstDesc* d_descs, int num_inst, int mode, int max_pop, hipStream_t s)
{
    if (num_inst <= 0) return hipSuccess;
    const int C = mode ? max_pop : 2 * max_pop;
    const int Cd = C < kDomRows ? C : kDomRows;  // larger populations count in the selection
    const int dom_blocks = (Cd * ((Cd + 31) / 32) + kDomBlock - 1) / kDomBlock;
    hipLaunchKernelGGL(k_nsga2_dominance, dim3(dom_blocks, num_inst), dim3(kDomBlock), 0, s, d_descs, mode);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int lds = select_layout(C).bytes;
    if (lds > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_nsga2_select),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    hipLaunchKernelGGL(k_nsga2_select, dim3(num_inst), dim3(kSelectBlock), lds, s, d_descs, mode);
    return hipGetLastError();
}

namespace {

template <int KIND, int BLOCK>
hipError_t refill_one(const InstDesc* d, int num, int rows, int lds, hipStream_t s)
{
    if (lds > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_nsga2_refill<KIND, BLOCK>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    hipLaunchKernelGGL((k_nsga2_refill<KIND, BLOCK>), dim3(rows), dim3(BLOCK), lds, s, d, num);
    return hipGetLastError();
}

template <int KIND>
hipError_t refill_kind(int block, const InstDesc* d, int num, int rows, int lds, hipStream_t s)
{
    switch (block) {
    case 64: return refill_one<KIND, 64>(d, num, rows, lds, s);
    case 128: return refill_one<KIND, 128>(d, num, rows, lds, s);
    case 256: return refill_one<KIND, 256>(d, num, rows, lds, s);
    case 512: return refill_one<KIND, 512>(d, num, rows, lds, s);
    case 1024: return refill_one<KIND, 1024>(d, num, rows, lds, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_nsga2_diversity(int kind, int block, const InstDesc* d_descs, int num_inst, int rows, int lds,
                                  int max_pop, hipStream_t s)
{
    if (num_inst <= 0 || rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_nsga2_diversity, dim3(num_inst), dim3(kDivBlock), 12 * max_pop, s, d_descs);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    switch (kind) {
    case KIND_BITSP_COORDS: e = refill_kind<KIND_BITSP_COORDS>(block, d_descs, num_inst, rows, lds, s); break;
    case KIND_BITSP_MATRIX: e = refill_kind<KIND_BITSP_MATRIX>(block, d_descs, num_inst, rows, lds, s); break;
    case KIND_ZDT1: e = refill_kind<KIND_ZDT1>(block, d_descs, num_inst, rows, lds, s); break;
    case KIND_MOVRP: e = refill_kind<KIND_MOVRP>(block, d_descs, num_inst, rows, lds, s); break;
    default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    return launch_nsga2_select(d_descs, num_inst, 2, max_pop, s);
}

hipError_t launch_nsga2_gather(const InstDesc* d_descs, int num_inst, int total_rows, hipStream_t s)
{
    if (total_rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_nsga2_gather, dim3((total_rows + kGatherRows - 1) / kGatherRows), dim3(kGatherRows * kWave),
                       0, s, d_descs, num_inst, total_rows);
    return hipGetLastError();
}

}  // namespace brkga
