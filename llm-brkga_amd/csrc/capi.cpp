// capi.cpp - the C ABI (include/brkga_hip.h): engine, problems, resident
// populations.  Host logic only; every kernel lives in kernels_*.hip.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "capi_internal.hpp"
#include "device_common.hpp"
#include "launch.hpp"

// Device staging of brkga_evaluate_host_noalloc: the keys, fitness and
// decoded rows of one call.  Each device the library runs on holds its own
// copy until the process ends.  Host code never uses this name as an
// address: engine creation resolves the device address once per device
// (hipGetSymbolAddress) into noalloc_scratch_ptr, and the entry point uses
// only that.
constexpr size_t kNoAllocScratchBytes = size_t(4) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_resolve_mu;
unsigned char* g_scratch_ptr[kMaxScratchDevices];  // device address of `scratch`, per device
std::mutex g_scratch_call_mu[kMaxScratchDevices];  // one call at a time per device

// Resolves the address of `scratch` on `device`, which must be current.
brkga_status resolve_noalloc_scratch(int device)
{
    if (device < 0 || device >= kMaxScratchDevices)
        return brkga::set_error(BRKGA_ERR_INVALID_ARG, "device index beyond the static scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_resolve_mu);
    if (g_scratch_ptr[device]) return BRKGA_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        return brkga::hip_error(e == hipSuccess ? hipErrorInvalidSymbol : e, "hipGetSymbolAddress(scratch)");
    }
    g_scratch_ptr[device] = static_cast<unsigned char*>(p);
    return BRKGA_OK;
}

size_t align256(size_t v)
{
    return (v + 255) & ~(size_t)255;
}
}  // namespace

namespace brkga {

namespace {
thread_local std::string g_last_error;

constexpr int kMaxGenes = 16000;   // fixed-point sums hold 2^14 terms
constexpr int kMaxMachines = 4096; // job-scheduling machine loads live in LDS
constexpr int kMaxIslandPop = 16384;
constexpr int kMaxNsgaPop = 2048;  // parents + children index in 12 bits

double max_abs(const float* a, size_t n)
{
    double m = 0.0;
    for (size_t i = 0; i < n; ++i) m = std::max(m, std::fabs((double)a[i]));
    return m;
}

double coords_diameter(const float* xy, int count)
{
    if (count <= 0) return 0.0;
    double x0 = xy[0], x1 = xy[0], y0 = xy[1], y1 = xy[1];
    for (int i = 1; i < count; ++i) {
        x0 = std::min(x0, (double)xy[2 * i]);
        x1 = std::max(x1, (double)xy[2 * i]);
        y0 = std::min(y0, (double)xy[2 * i + 1]);
        y1 = std::max(y1, (double)xy[2 * i + 1]);
    }
    const double dx = x1 - x0, dy = y1 - y0;
    return std::sqrt(dx * dx + dy * dy) * (1.0 + 1e-6);
}

bool all_finite(const float* a, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(a[i])) return false;
    return true;
}

void set_scales(ProblemDesc* d, int s0, int s1)
{
    d->scale0 = std::ldexp(1.0f, s0);
    d->inv0 = std::ldexp(1.0f, -s0);
    d->scale1 = std::ldexp(1.0f, s1);
    d->inv1 = std::ldexp(1.0f, -s1);
}

// Problem data go up asynchronously on the legacy stream (the host copy is
// staged by the runtime before the call returns, so the caller's buffer is
// free again); finish_problem records the problem's `ready` event after
// them, and problem_ready waits for it before any device use.  A
// synchronous copy per array cost about 17 us, most of it the wait, which
// made building many small problems (a burst of served requests) slow.
brkga_status upload(brkga_problem_s* pr, const void* host, size_t bytes, const void** dst)
{
    void* p = nullptr;
    brkga_status st = device_alloc(&p, bytes, "problem data");
    if (st != BRKGA_OK) return st;
    pr->owned.push_back(p);
    hipError_t e = hipMemcpyAsync(p, host, bytes, hipMemcpyHostToDevice, nullptr);
    if (e != hipSuccess) return hip_error(e, "hipMemcpyAsync(problem data)");
    pr->uploading = true;
    *dst = p;
    return BRKGA_OK;
}

void free_all(std::vector<void*>& v)
{
    for (void* p : v) (void)hipFree(p);
    v.clear();
}

brkga_status check_lds(const brkga_engine_s* eng, const ProblemDesc& p)
{
    const int bytes = decode_layout(p.kind, p.n, p.chrom_len, sort_buckets(p.n), p.aux).bytes;
    if ((size_t)bytes > eng->lds_limit)
        return set_error(BRKGA_ERR_INVALID_ARG, "chromosome too long: decode needs " + std::to_string(bytes) +
                                                    " bytes of LDS, the device allows " +
                                                    std::to_string(eng->lds_limit));
    return BRKGA_OK;
}

brkga_status new_problem(brkga_engine_t engine, int kind, int n, brkga_problem_s** out)
{
    if (!engine || !out) return set_error(BRKGA_ERR_INVALID_ARG, "null engine or output handle");
    if (n < 1 || n > kMaxGenes)
        return set_error(BRKGA_ERR_INVALID_ARG, "problem size must be in [1, " + std::to_string(kMaxGenes) + "]");
    auto* pr = new brkga_problem_s();
    pr->engine = engine;
    std::memset(&pr->desc, 0, sizeof(pr->desc));
    pr->desc.kind = kind;
    pr->desc.n = n;
    pr->desc.chrom_len = kind_chrom_len(kind, n);
    pr->desc.num_obj = kind_num_obj(kind);
    pr->desc.maximize = (kind == KIND_KNAP_THRESHOLD || kind == KIND_KNAP_GREEDY) ? 1 : 0;
    *out = pr;
    return BRKGA_OK;
}

// Checks that the decoder fits in LDS, then hands the problem out (or frees
// it on any error).
brkga_status finish_problem(brkga_problem_s* pr, brkga_status st, brkga_problem_t* out)
{
    if (st == BRKGA_OK) st = check_lds(pr->engine, pr->desc);
    if (st == BRKGA_OK && pr->uploading) {
        hipError_t e = hipEventCreateWithFlags(&pr->ready, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventRecord(pr->ready, nullptr);
        if (e != hipSuccess) st = hip_error(e, "problem upload event");
    }
    if (st != BRKGA_OK) {
        (void)hipStreamSynchronize(nullptr);  // no upload may still read or write what is freed
        if (pr->ready) (void)hipEventDestroy(pr->ready);
        free_all(pr->owned);
        delete pr;
        return st;
    }
    *out = pr;
    return BRKGA_OK;
}

}  // namespace

brkga_status problem_ready(brkga_problem_s* pr)
{
    if (!pr->uploading.load()) return BRKGA_OK;
    DeviceGuard g(pr->engine->device);
    const hipError_t e = hipEventSynchronize(pr->ready);
    if (e != hipSuccess) return hip_error(e, "problem upload");
    pr->uploading = false;
    return BRKGA_OK;
}

brkga_status set_error(brkga_status st, const std::string& msg)
{
    g_last_error = msg;
    return st;
}

brkga_status hip_error(hipError_t e, const char* what)
{
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        return set_error(BRKGA_ERR_OOM, std::string(what) + ": " + hipGetErrorString(e));
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice)
        return set_error(BRKGA_ERR_NO_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
    return set_error(BRKGA_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

brkga_status device_alloc(void** p, size_t bytes, const char* what)
{
    *p = nullptr;
    hipError_t e = hipMalloc(p, bytes < 16 ? 16 : bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
            return set_error(BRKGA_ERR_OOM, std::string("out of device memory allocating ") + what);
        return hip_error(e, what);
    }
    return BRKGA_OK;
}

// Shift S with |term| * 2^S <= 2^48 for every term below max_abs_term.
int fixed_shift(double max_abs_term)
{
    if (!(max_abs_term > 0.0) || !std::isfinite(max_abs_term)) return 30;
    int e = 0;
    std::frexp(max_abs_term, &e);
    return std::max(-60, std::min(60, 48 - e));
}

// Smallest double s whose correctly rounded square root is at least 0.001:
// sqrt(s) < 0.001 (the reference's test, population.hpp:139-146) holds
// exactly when s is below it, so the kernels compare squared distances.
static double diversity_eps2()
{
    double s = 0.001 * 0.001;
    while (std::sqrt(s) >= 0.001) s = std::nextafter(s, 0.0);
    while (std::sqrt(s) < 0.001) s = std::nextafter(s, 1.0);
    return s;
}

brkga_status fill_inst_desc(const brkga_problem_s* pr, const brkga_params* q, InstDesc* d, int* nsga2)
{
    if (!pr || !q || !d) return set_error(BRKGA_ERR_INVALID_ARG, "null argument");
    std::memset(d, 0, sizeof(*d));
    const ProblemDesc& p = pr->desc;
    const int mo = p.num_obj > 1;
    *nsga2 = mo;
    const int P = q->population_size;
    if (P <= 0) return set_error(BRKGA_ERR_INVALID_ARG, "population_size must be > 0");
    if (!(q->elite_prob >= 0.0 && q->elite_prob <= 1.0))
        return set_error(BRKGA_ERR_INVALID_ARG, "elite_prob must be in [0, 1]");
    d->prob = p;
    d->L = p.chrom_len;
    d->ld = round_up_int(p.chrom_len, 32);
    d->pop = P;
    d->seed_lo = (uint32_t)q->seed;
    d->seed_hi = (uint32_t)(q->seed >> 32);
    d->rho = (float)q->elite_prob;
    d->nb = sort_buckets(p.n);
    d->gen_limit = INT_MAX;
    if (mo) {
        if (q->num_islands > 1) return set_error(BRKGA_ERR_INVALID_ARG, "NSGA-II runs a single population");
        if (P > kMaxNsgaPop)
            return set_error(BRKGA_ERR_INVALID_ARG, "NSGA-II population_size must be <= " +
                                                        std::to_string(kMaxNsgaPop));
        if (q->diversity_interval < 0)
            return set_error(BRKGA_ERR_INVALID_ARG, "diversity_interval must be >= 0");
        if (q->diversity_interval > 0 && p.kind == KIND_HOST)
            return set_error(BRKGA_ERR_INVALID_ARG, "the diversity injection decodes its new members on the device; "
                                                    "host-evaluated problems run without it");
        d->islands = 1;
        d->Pi = P;
        d->mut_rate = (float)(5.0 / (double)p.chrom_len);
        d->div_interval = q->diversity_interval;
        static const double eps2 = diversity_eps2();
        d->div_eps2 = eps2;
        return BRKGA_OK;
    }
    const int I = q->num_islands < 1 ? 1 : q->num_islands;
    if (P % I != 0) return set_error(BRKGA_ERR_INVALID_ARG, "population_size must be a multiple of num_islands");
    const int Pi = P / I, Ei = q->elite_size / I, Mi = q->mutant_size / I;
    if (Pi > kMaxIslandPop)
        return set_error(BRKGA_ERR_INVALID_ARG, "island population must be <= " + std::to_string(kMaxIslandPop));
    if (Ei <= 0) return set_error(BRKGA_ERR_INVALID_ARG, "elite_size per island must be > 0");
    if (Mi < 0) return set_error(BRKGA_ERR_INVALID_ARG, "mutant_size must be >= 0");
    if (Ei + Mi > Pi) return set_error(BRKGA_ERR_INVALID_ARG, "elite_size + mutant_size exceeds population_size");
    if (Ei >= Pi && Pi - Ei - Mi > 0) return set_error(BRKGA_ERR_INVALID_ARG, "no non-elite parents");
    if (q->migration_interval < 0 || q->num_migrants < 0)
        return set_error(BRKGA_ERR_INVALID_ARG, "migration settings must be >= 0");
    if (I > 1 && q->num_migrants > Ei)
        return set_error(BRKGA_ERR_INVALID_ARG, "num_migrants must not exceed the elite per island");
    if (q->ls_interval < 0 || q->ls_individuals < 0 || q->ls_moves < 0)
        return set_error(BRKGA_ERR_INVALID_ARG, "local search settings must be >= 0");
    if (q->ls_interval > 0 && q->ls_individuals > 0 && p.kind != KIND_TSPJ)
        return set_error(BRKGA_ERR_INVALID_ARG, "the local search is defined for TSPJ problems only");
    if (q->ls_interval > 0 && q->ls_individuals > 0) {
        d->ls_interval = q->ls_interval;
        d->ls_k = std::min(q->ls_individuals, Ei);
        d->ls_moves = q->ls_moves;
    }
    d->islands = I;
    d->Pi = Pi;
    d->Ei = Ei;
    d->Mi = Mi;
    d->mig_interval = q->migration_interval;
    d->mig_k = I > 1 ? q->num_migrants : 0;
    return BRKGA_OK;
}

brkga_status alloc_inst_buffers(InstDesc* d, int nsga2, DeviceAllocator* a, std::vector<void*>* owned)
{
    const size_t rows = nsga2 ? (size_t)2 * d->pop : (size_t)d->pop;
    const int M = d->prob.num_obj;
    auto get = [&](void** p, size_t bytes, const char* what) -> brkga_status {
        brkga_status st = a ? a->alloc(p, bytes, what) : device_alloc(p, bytes, what);
        if (st == BRKGA_OK) owned->push_back(*p);
        return st;
    };
    brkga_status st;
    for (int b = 0; b < 2; ++b) {
        void* p;
        if ((st = get(&p, rows * d->ld * sizeof(float), "population keys")) != BRKGA_OK) return st;
        d->keys[b] = (float*)p;
        if ((st = get(&p, rows * M * sizeof(float), "fitness")) != BRKGA_OK) return st;
        d->fit[b] = (float*)p;
        if ((st = get(&p, rows * sizeof(int), "rank order")) != BRKGA_OK) return st;
        d->perm[b] = (int*)p;
        if (nsga2) {
            if ((st = get(&p, rows * sizeof(int), "front ranks")) != BRKGA_OK) return st;
            d->rank[b] = (int*)p;
            if ((st = get(&p, rows * sizeof(float), "crowding")) != BRKGA_OK) return st;
            d->crowd[b] = (float*)p;
        }
    }
    if (nsga2 && rows <= (size_t)kDomRows) {
        void* p;
        if ((st = get(&p, rows * ((rows + 31) / 32) * sizeof(unsigned), "dominance bitsets")) != BRKGA_OK) return st;
        d->dom = (unsigned*)p;
    }
    void* c;
    if ((st = get(&c, 4 * sizeof(int), "counters")) != BRKGA_OK) return st;
    d->gen = (int*)c;
    d->done = (int*)c + 1;
    d->ls_flag = (int*)c + 2;
    return BRKGA_OK;
}

namespace {

// Allocations of one population besides the ping-pong buffers.
brkga_status alloc_pop_extras(brkga_pop_s* pop, DeviceAllocator* a)
{
    auto get = [&](void** p, size_t bytes, const char* what) -> brkga_status {
        brkga_status r = a ? a->alloc(p, bytes, what) : device_alloc(p, bytes, what);
        if (r == BRKGA_OK) pop->owned.push_back(*p);
        return r;
    };
    brkga_status st;
    void* p = nullptr;
    if ((st = get(&p, sizeof(InstDesc), "descriptor")) != BRKGA_OK) return st;
    pop->d_desc = (InstDesc*)p;
    if ((st = get(&p, sizeof(int) * pop->host.L, "decode row")) != BRKGA_OK) return st;
    pop->d_decode_row = (int*)p;
    if ((st = get(&p, sizeof(float) * 2, "fitness row")) != BRKGA_OK) return st;
    pop->d_fit_row = (float*)p;
    return BRKGA_OK;
}

// Counts bytes (at the arena's 256-byte granularity) instead of allocating.
struct CountingAllocator : DeviceAllocator {
    size_t bytes = 0;
    uintptr_t next = 256;
    brkga_status alloc(void** p, size_t n, const char*) override
    {
        const size_t r = (n + 255) & ~(size_t)255;
        bytes += r;
        *p = reinterpret_cast<void*>(next);
        next += r;
        return BRKGA_OK;
    }
    void release(void*) override {}
};

}  // namespace

brkga_status population_bytes(const brkga_problem_s* pr, const brkga_params* q, size_t* bytes)
{
    InstDesc d;
    int nsga2 = 0;
    brkga_status st = fill_inst_desc(pr, q, &d, &nsga2);
    if (st != BRKGA_OK) return st;
    CountingAllocator c;
    std::vector<void*> owned;
    st = alloc_inst_buffers(&d, nsga2, &c, &owned);
    if (st != BRKGA_OK) return st;
    brkga_pop_s tmp;
    tmp.host = d;
    st = alloc_pop_extras(&tmp, &c);
    if (st != BRKGA_OK) return st;
    *bytes = c.bytes;
    return BRKGA_OK;
}

brkga_status create_population(brkga_problem_s* problem, const brkga_params* params, DeviceAllocator* a,
                               hipStream_t s, int gen0, brkga_pop_s** out, PopSetup* deferred)
{
    if (!problem || !params || !out) return set_error(BRKGA_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    const brkga_status rs = problem_ready(problem);
    if (rs != BRKGA_OK) return rs;
    auto* pop = new brkga_pop_s();
    pop->problem = problem;
    pop->allocator = a;
    pop->params = *params;
    pop->d_desc = nullptr;
    pop->capture_stream = nullptr;
    brkga_status st = fill_inst_desc(problem, params, &pop->host, &pop->nsga2);
    if (st == BRKGA_OK) st = alloc_inst_buffers(&pop->host, pop->nsga2, a, &pop->owned);
    if (st == BRKGA_OK) st = alloc_pop_extras(pop, a);
    if (st == BRKGA_OK) {
        const int n = problem->desc.n;
        pop->block = default_block(n);
        pop->host_t = 0;
        pop->step_lds = pop->nsga2 ? decode_layout(problem->desc.kind, n, pop->host.L, pop->host.nb, problem->desc.aux).bytes
                                   : step_lds_bytes(pop->host);
        if ((size_t)partition_lds_bytes(pop->host.Pi) > problem->engine->lds_limit ||
            (size_t)pop->step_lds > problem->engine->lds_limit)
            st = set_error(BRKGA_ERR_INVALID_ARG, "population or chromosome too large for the LDS kernels");
    }
    if (st == BRKGA_OK && deferred) {
        const InstDesc& d = pop->host;
        const size_t rows = pop->nsga2 ? (size_t)2 * d.pop : (size_t)d.pop;
        deferred->keys[0] = d.keys[0];
        deferred->keys[1] = d.keys[1];
        deferred->key_floats = rows * d.ld;
        deferred->gen = d.gen;
        deferred->gen_edge = gen0 < 0 ? -1 : 0;
        deferred->d_desc = pop->d_desc;
        deferred->desc = d;
    } else if (st == BRKGA_OK) {
        // clear the key buffers (their padding columns are never written by
        // the RNG paths, and copies must stay deterministic) and the
        // counters; pop->host outlives the descriptor copy
        const InstDesc& d = pop->host;
        const size_t rows = pop->nsga2 ? (size_t)2 * d.pop : (size_t)d.pop;
        hipError_t e = hipSuccess;
        for (int b = 0; b < 2 && e == hipSuccess; ++b)
            e = hipMemsetAsync(d.keys[b], 0, rows * d.ld * sizeof(float), s);
        if (e == hipSuccess) e = hipMemsetAsync(d.gen, 0, 4 * sizeof(int), s);
        if (e == hipSuccess && gen0 < 0) e = hipMemsetAsync(d.gen, 0xFF, sizeof(int), s);
        if (e == hipSuccess && gen0 < 0) e = hipMemsetAsync(d.gen + 3, 0xFF, sizeof(int), s);
        if (e == hipSuccess) e = hipMemcpyAsync(pop->d_desc, &pop->host, sizeof(InstDesc), hipMemcpyHostToDevice, s);
        if (e != hipSuccess) st = hip_error(e, "population upload");
    }
    if (st != BRKGA_OK) {
        (void)hipStreamSynchronize(s);
        destroy_population(pop);
        return st;
    }
    *out = pop;
    return BRKGA_OK;
}

void destroy_population(brkga_pop_s* pop)
{
    if (!pop) return;
    for (auto& kv : pop->graphs) (void)hipGraphExecDestroy(kv.second);
    if (pop->capture_stream) (void)hipStreamDestroy(pop->capture_stream);
    if (pop->allocator) {
        for (void* p : pop->owned) pop->allocator->release(p);
        pop->owned.clear();
    } else {
        free_all(pop->owned);
    }
    delete pop;
}

int pick_best(const InstDesc& d, int nsga2, const float* fit, const int* order_or_rank, const float* crowd)
{
    int best = -1;
    if (nsga2) {
        for (int i = 0; i < d.pop; ++i)
            if (order_or_rank[i] == 0 && (best < 0 || crowd[i] > crowd[best])) best = i;
        return best;
    }
    for (int i = 0; i < d.islands; ++i) {
        const int cand = i * d.Pi + order_or_rank[(size_t)i * d.Pi];
        if (best < 0) {
            best = cand;
            continue;
        }
        const float a = fit[cand], c = fit[best];
        if (d.prob.maximize ? a > c : a < c) best = cand;
    }
    return best;
}

}  // namespace brkga

using namespace brkga;

namespace {

hipStream_t as_stream(void* s)
{
    return reinterpret_cast<hipStream_t>(s);
}

const char* const kNoDecoder =
    "the problem is evaluated on the host (no device decoder): use brkga_population_host_begin / _end";

bool host_kind(const brkga_pop_s* pop)
{
    return pop->problem->desc.kind == KIND_HOST;
}

// Device-decoded generations of a host-evaluated population are refused.
brkga_status device_path_check(const brkga_pop_s* pop)
{
    if (!pop) return set_error(BRKGA_ERR_INVALID_ARG, "null population");
    if (host_kind(pop)) return set_error(BRKGA_ERR_INVALID_ARG, kNoDecoder);
    return BRKGA_OK;
}

// Whether the population's steps rank their parent islands themselves: the
// island fits the workgroup, and no local search needs each generation's
// rank order.  A run of such steps then needs no partition per generation,
// only one at its end (pop_close).
bool pop_ranks(const brkga_pop_s* pop)
{
    return !pop->nsga2 && step_can_rank(pop->host, pop->block, pop->problem->engine->lds_limit);
}

// Children per workgroup the population's ranking steps run with (1 when
// the multi-child form does not apply at the current workgroup size).
int pop_children(const brkga_pop_s* pop)
{
    if (pop->children <= 1) return 1;
    return step_multi_ok(pop->host, pop->block, pop->children, pop->problem->engine->lds_limit) ? pop->children : 1;
}

// Enqueues generation 0 (init) or one step whose parent generation has
// buffer parity `cur`.
brkga_status pop_step(brkga_pop_t pop, int init, int cur, hipStream_t s)
{
    const InstDesc& d = pop->host;
    hipError_t e;
    if (pop->nsga2) {
        e = launch_nsga2_offspring(d.prob.kind, pop->block, pop->d_desc, 1, d.pop, pop->step_lds, init, s);
        if (e == hipSuccess) e = launch_nsga2_select(pop->d_desc, 1, init, d.pop, s);
        if (e == hipSuccess && !init) e = launch_nsga2_gather(pop->d_desc, 1, d.pop, s);
        // the injection's kernels check on the device whether this step is
        // due, so graphs replay at any generation
        if (e == hipSuccess && !init && d.div_interval > 0)
            e = launch_nsga2_diversity(d.prob.kind, pop->block, pop->d_desc, 1, d.pop, pop->step_lds, d.pop, s);
    } else {
        const int rank = !init && pop_ranks(pop);
        e = launch_step_one(d, pop->block, cur, init, rank, s, nullptr, nullptr, rank ? pop_children(pop) : 1);
        if (e == hipSuccess && !rank) e = launch_partition_one(d, init ? 0 : cur ^ 1, init ? 0 : 1, s);
        // TSPJ local search: the kernels check on the device whether the
        // generation just built is due, so graphs replay at any generation
        if (e == hipSuccess && !init && d.ls_interval > 0) {
            e = launch_tspj_local_search(pop->d_desc, 1, d.islands * d.ls_k, pop->block,
                                         tspj_ls_lds_bytes(d.prob.n, d.nb), s);
            if (e == hipSuccess) e = launch_partition_one(d, -1, 2, s);
        }
    }
    if (e != hipSuccess) return hip_error(e, "kernel launch");
    return BRKGA_OK;
}

// Ends a run of generations whose last one has buffer parity `buf`: steps
// that ranked their parents leave the rank order of the newest generation
// and the canonical step counter to one partition launch.
brkga_status pop_close(brkga_pop_t pop, int buf, hipStream_t s)
{
    if (!pop_ranks(pop)) return BRKGA_OK;
    hipError_t e = launch_partition_one(pop->host, buf, 3, s);
    if (e != hipSuccess) return hip_error(e, "partition launch");
    return BRKGA_OK;
}

}  // namespace

brkga_status brkga::population_steps(brkga_pop_s* pop, int init, int steps, hipStream_t s)
{
    if (init) {
        BRKGA_HIP_TRY(hipMemsetAsync(pop->host.gen, 0, 4 * sizeof(int), s));
        pop->host_t = 0;
        const brkga_status st = pop_step(pop, 1, 0, s);
        if (st != BRKGA_OK) return st;
    }
    for (int i = 0; i < steps; ++i) {
        const brkga_status st = pop_step(pop, 0, pop->host_t & 1, s);
        if (st != BRKGA_OK) return st;
        ++pop->host_t;
    }
    return steps > 0 ? pop_close(pop, pop->host_t & 1, s) : BRKGA_OK;
}

namespace {

brkga_status read_gen(brkga_pop_t pop, hipStream_t s, int* gen)
{
    BRKGA_HIP_TRY(hipStreamSynchronize(s));
    BRKGA_HIP_TRY(hipMemcpy(gen, pop->host.gen, sizeof(int), hipMemcpyDeviceToHost));
    return BRKGA_OK;
}

}  // namespace

extern "C" {

const char* brkga_version(void)
{
    return "brkga-mi355x 0.1.0 (gfx950)";
}

const char* brkga_last_error(void)
{
    return g_last_error.c_str();
}

brkga_status brkga_device_count(int32_t* count)
{
    if (!count) return set_error(BRKGA_ERR_INVALID_ARG, "null count");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        c = 0;
    }
    *count = c;
    return BRKGA_OK;
}

brkga_status brkga_engine_create(int32_t device, brkga_engine_t* out)
{
    if (!out) return set_error(BRKGA_ERR_INVALID_ARG, "null output handle");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        (void)hipGetLastError();
        return set_error(BRKGA_ERR_NO_DEVICE, "no HIP device available");
    }
    if (device < 0 || device >= count) return set_error(BRKGA_ERR_INVALID_ARG, "device index out of range");
    DeviceGuard g(device);
    int lds = 0, lds_optin = 0, cus = 0;
    BRKGA_HIP_TRY(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    if (hipDeviceGetAttribute(&lds_optin, hipDeviceAttributeSharedMemPerBlockOptin, device) != hipSuccess) {
        (void)hipGetLastError();
        lds_optin = 0;
    }
    BRKGA_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    // set-up of the allocation-free entry point, once per device, with the
    // device current
    const brkga_status sst = resolve_noalloc_scratch(device);
    if (sst != BRKGA_OK) return sst;
    auto* eng = new brkga_engine_s();
    eng->device = device;
    eng->lds_limit = (size_t)std::max(lds, lds_optin);
    eng->cu_count = cus;
    *out = eng;
    return BRKGA_OK;
}

brkga_status brkga_engine_destroy(brkga_engine_t engine)
{
    delete engine;
    return BRKGA_OK;
}

brkga_status brkga_problem_create_tsp_coords(brkga_engine_t engine, const float* xy, int32_t n,
                                             brkga_problem_t* out)
{
    if (!xy) return set_error(BRKGA_ERR_INVALID_ARG, "null coordinates");
    if (n > 0 && !all_finite(xy, (size_t)2 * n)) return set_error(BRKGA_ERR_INVALID_ARG, "non-finite coordinates");
    brkga_problem_s* pr;
    brkga_status st = new_problem(engine, KIND_TSP_COORDS, n, &pr);
    if (st != BRKGA_OK) return st;
    DeviceGuard g(engine->device);
    pr->desc.dim = n;
    set_scales(&pr->desc, fixed_shift(coords_diameter(xy, n)), 30);
    st = upload(pr, xy, sizeof(float) * 2 * (size_t)n, (const void**)&pr->desc.a);
    return finish_problem(pr, st, out);
}

brkga_status brkga_problem_create_tsp_matrix(brkga_engine_t engine, const float* dist, int32_t n,
                                             brkga_problem_t* out)
{
    if (!dist) return set_error(BRKGA_ERR_INVALID_ARG, "null matrix");
    if (n > 0 && !all_finite(dist, (size_t)n * n)) return set_error(BRKGA_ERR_INVALID_ARG, "non-finite distances");
    brkga_problem_s* pr;
    brkga_status st = new_problem(engine, KIND_TSP_MATRIX, n, &pr);
    if (st != BRKGA_OK) return st;
    DeviceGuard g(engine->device);
    pr->desc.dim = n;
    set_scales(&pr->desc, fixed_shift(max_abs(dist, (size_t)n * n)), 30);
    st = upload(pr, dist, sizeof(float) * (size_t)n * n, (const void**)&pr->desc.a);
    return finish_problem(pr, st, out);
}

brkga_status brkga_problem_create_knapsack(brkga_engine_t engine, const int32_t* weights, const int32_t* values,
                                           int32_t n, int64_t capacity, int32_t greedy, brkga_problem_t* out)
{
    if (!weights || !values) return set_error(BRKGA_ERR_INVALID_ARG, "null weights or values");
    if (greedy) {
        if (capacity < 0) return set_error(BRKGA_ERR_INVALID_ARG, "greedy knapsack needs capacity >= 0");
        for (int32_t i = 0; i < n; ++i)
            if (weights[i] < 0) return set_error(BRKGA_ERR_INVALID_ARG, "greedy knapsack needs weights >= 0");
    }
    brkga_problem_s* pr;
    brkga_status st = new_problem(engine, greedy ? KIND_KNAP_GREEDY : KIND_KNAP_THRESHOLD, n, &pr);
    if (st != BRKGA_OK) return st;
    DeviceGuard g(engine->device);
    pr->desc.dim = n;
    pr->desc.cap = capacity;
    set_scales(&pr->desc, 0, 0);
    st = upload(pr, weights, sizeof(int32_t) * (size_t)n, (const void**)&pr->desc.w);
    if (st == BRKGA_OK) st = upload(pr, values, sizeof(int32_t) * (size_t)n, (const void**)&pr->desc.v);
    return finish_problem(pr, st, out);
}

brkga_status brkga_problem_create_tspj(brkga_engine_t engine, const float* travel, const float* jobs, int32_t dim,
                                       brkga_problem_t* out)
{
    if (!travel || !jobs) return set_error(BRKGA_ERR_INVALID_ARG, "null matrices");
    if (dim < 2) return set_error(BRKGA_ERR_INVALID_ARG, "TSPJ needs dim >= 2 (depot + one city)");
    const size_t cells = (size_t)dim * dim;
    if (!all_finite(travel, cells) || !all_finite(jobs, cells))
        return set_error(BRKGA_ERR_INVALID_ARG, "non-finite matrix entries");
    brkga_problem_s* pr;
    brkga_status st = new_problem(engine, KIND_TSPJ, dim - 1, &pr);
    if (st != BRKGA_OK) return st;
    DeviceGuard g(engine->device);
    pr->desc.dim = dim;
    set_scales(&pr->desc, fixed_shift(std::max(max_abs(travel, cells), max_abs(jobs, cells))), 30);
    st = upload(pr, travel, sizeof(float) * cells, (const void**)&pr->desc.a);
    if (st == BRKGA_OK) st = upload(pr, jobs, sizeof(float) * cells, (const void**)&pr->desc.b);
    return finish_problem(pr, st, out);
}

brkga_status brkga_problem_create_bitsp_coords(brkga_engine_t engine, const float* xy0, const float* xy1, int32_t n,
                                               brkga_problem_t* out)
{
    if (!xy0 || !xy1) return set_error(BRKGA_ERR_INVALID_ARG, "null coordinates");
    if (n > 0 && (!all_finite(xy0, (size_t)2 * (n + 1)) || !all_finite(xy1, (size_t)2 * (n + 1))))
        return set_error(BRKGA_ERR_INVALID_ARG, "non-finite coordinates");
    brkga_problem_s* pr;
    brkga_status st = new_problem(engine, KIND_BITSP_COORDS, n, &pr);
    if (st != BRKGA_OK) return st;
    DeviceGuard g(engine->device);
    pr->desc.dim = n + 1;
    set_scales(&pr->desc, fixed_shift(coords_diameter(xy0, n + 1)), fixed_shift(coords_diameter(xy1, n + 1)));
    st = upload(pr, xy0, sizeof(float) * 2 * (size_t)(n + 1), (const void**)&pr->desc.a);
    if (st == BRKGA_OK) st = upload(pr, xy1, sizeof(float) * 2 * (size_t)(n + 1), (const void**)&pr->desc.b);
    return finish_problem(pr, st, out);
}

brkga_status brkga_problem_create_bitsp_matrix(brkga_engine_t engine, const float* m0, const float* m1, int32_t dim,
                                               brkga_problem_t* out)
{
    if (!m0 || !m1) return set_error(BRKGA_ERR_INVALID_ARG, "null matrices");
    if (dim < 2) return set_error(BRKGA_ERR_INVALID_ARG, "bi-objective TSP needs dim >= 2 (depot + one city)");
    const size_t cells = (size_t)dim * dim;
    if (!all_finite(m0, cells) || !all_finite(m1, cells))
        return set_error(BRKGA_ERR_INVALID_ARG, "non-finite matrix entries");
    brkga_problem_s* pr;
    brkga_status st = new_problem(engine, KIND_BITSP_MATRIX, dim - 1, &pr);
    if (st != BRKGA_OK) return st;
    DeviceGuard g(engine->device);
    pr->desc.dim = dim;
    set_scales(&pr->desc, fixed_shift(max_abs(m0, cells)), fixed_shift(max_abs(m1, cells)));
    st = upload(pr, m0, sizeof(float) * cells, (const void**)&pr->desc.a);
    if (st == BRKGA_OK) st = upload(pr, m1, sizeof(float) * cells, (const void**)&pr->desc.b);
    return finish_problem(pr, st, out);
}

brkga_status brkga_problem_create_continuous(brkga_engine_t engine, int32_t function, int32_t n, float lower,
                                             float upper, brkga_problem_t* out)
{
    if (function < FN_SPHERE || function > FN_ROSENBROCK)
        return set_error(BRKGA_ERR_INVALID_ARG, "function must be 0 (sphere), 1 (rastrigin) or 2 (rosenbrock)");
    if (!std::isfinite(lower) || !std::isfinite(upper)) return set_error(BRKGA_ERR_INVALID_ARG, "non-finite bounds");
    brkga_problem_s* pr;
    brkga_status st = new_problem(engine, KIND_CONTINUOUS, n, &pr);
    if (st != BRKGA_OK) return st;
    pr->desc.dim = n;
    pr->desc.ip0 = function;
    pr->desc.f0 = lower;
    pr->desc.f1 = upper;
    const double m = std::max(std::fabs((double)lower), std::fabs((double)upper));
    const double bound = function == FN_SPHERE ? m * m
                         : function == FN_RASTRIGIN ? m * m + 10.0
                                                    : 100.0 * (m + m * m) * (m + m * m) + (1.0 + m) * (1.0 + m);
    set_scales(&pr->desc, fixed_shift(bound * (1.0 + 1e-6)), 30);
    return finish_problem(pr, BRKGA_OK, out);
}

brkga_status brkga_problem_create_job_scheduling(brkga_engine_t engine, const float* job_times, int32_t n,
                                                 int32_t machines, brkga_problem_t* out)
{
    if (!job_times) return set_error(BRKGA_ERR_INVALID_ARG, "null job times");
    if (machines < 1 || machines > kMaxMachines)
        return set_error(BRKGA_ERR_INVALID_ARG, "machines must be in [1, " + std::to_string(kMaxMachines) + "]");
    if (n > 0 && !all_finite(job_times, (size_t)n)) return set_error(BRKGA_ERR_INVALID_ARG, "non-finite job times");
    brkga_problem_s* pr;
    brkga_status st = new_problem(engine, KIND_JOBSHOP, n, &pr);
    if (st != BRKGA_OK) return st;
    DeviceGuard g(engine->device);
    pr->desc.dim = n;
    pr->desc.ip0 = machines;
    pr->desc.aux = 8 * machines;
    set_scales(&pr->desc, fixed_shift(max_abs(job_times, (size_t)n)), 30);
    st = upload(pr, job_times, sizeof(float) * (size_t)n, (const void**)&pr->desc.a);
    return finish_problem(pr, st, out);
}

brkga_status brkga_problem_create_zdt1(brkga_engine_t engine, int32_t n, brkga_problem_t* out)
{
    brkga_problem_s* pr;
    brkga_status st = new_problem(engine, KIND_ZDT1, n, &pr);
    if (st != BRKGA_OK) return st;
    pr->desc.dim = n;
    set_scales(&pr->desc, fixed_shift(1.0), 30);
    return finish_problem(pr, BRKGA_OK, out);
}

brkga_status brkga_problem_create_vrp_days(brkga_engine_t engine, const float* time_matrix, int32_t dim,
                                           int32_t max_stops_per_day, double max_drive_minutes,
                                           brkga_problem_t* out)
{
    if (!time_matrix) return set_error(BRKGA_ERR_INVALID_ARG, "null time matrix");
    if (dim < 2) return set_error(BRKGA_ERR_INVALID_ARG, "VRP needs dim >= 2 (depot + one stop)");
    const size_t cells = (size_t)dim * dim;
    if (!all_finite(time_matrix, cells)) return set_error(BRKGA_ERR_INVALID_ARG, "non-finite matrix entries");
    if (!std::isfinite(max_drive_minutes)) return set_error(BRKGA_ERR_INVALID_ARG, "non-finite drive limit");
    brkga_problem_s* pr;
    brkga_status st = new_problem(engine, KIND_VRP_DAYS, dim - 1, &pr);
    if (st != BRKGA_OK) return st;
    DeviceGuard g(engine->device);
    const int n = dim - 1;
    pr->desc.dim = dim;
    pr->desc.ip0 = max_stops_per_day;
    pr->desc.d0 = max_drive_minutes;
    pr->desc.aux = 8 * n + 16;
    const int per_day = std::max(0, std::min(max_stops_per_day, n));
    const double bound = std::max((per_day + 1.0) * max_abs(time_matrix, cells), 2000.0 * n);
    set_scales(&pr->desc, fixed_shift(bound), 30);
    st = upload(pr, time_matrix, sizeof(float) * cells, (const void**)&pr->desc.a);
    return finish_problem(pr, st, out);
}

brkga_status brkga_problem_create_movrp(brkga_engine_t engine, const float* xy, const int32_t* demands, int32_t dim,
                                        int64_t capacity, int32_t vehicles, brkga_problem_t* out)
{
    if (!xy || !demands) return set_error(BRKGA_ERR_INVALID_ARG, "null coordinates or demands");
    if (dim < 2) return set_error(BRKGA_ERR_INVALID_ARG, "VRP needs dim >= 2 (depot + one customer)");
    if (!all_finite(xy, (size_t)2 * dim)) return set_error(BRKGA_ERR_INVALID_ARG, "non-finite coordinates");
    if (vehicles < 0) return set_error(BRKGA_ERR_INVALID_ARG, "vehicles must be >= 0");
    int32_t max_dem = 0;
    for (int32_t i = 0; i < dim; ++i) {
        if (demands[i] < 0) return set_error(BRKGA_ERR_INVALID_ARG, "demands must be >= 0");
        max_dem = std::max(max_dem, demands[i]);
    }
    brkga_problem_s* pr;
    brkga_status st = new_problem(engine, KIND_MOVRP, dim - 1, &pr);
    if (st != BRKGA_OK) return st;
    DeviceGuard g(engine->device);
    const int n = dim - 1;
    pr->desc.dim = dim;
    pr->desc.cap = capacity;
    pr->desc.ip0 = vehicles;
    pr->desc.aux = 4 * round_up_int(n + 1, 2) + 16 * n + 16;
    const double bound = std::max(coords_diameter(xy, dim), std::max(100.0 * max_dem, 500.0 * n));
    set_scales(&pr->desc, fixed_shift(bound), fixed_shift(bound));
    st = upload(pr, xy, sizeof(float) * 2 * (size_t)dim, (const void**)&pr->desc.a);
    if (st == BRKGA_OK) st = upload(pr, demands, sizeof(int32_t) * (size_t)dim, (const void**)&pr->desc.w);
    return finish_problem(pr, st, out);
}

brkga_status brkga_problem_create_host(brkga_engine_t engine, int32_t chrom_len, int32_t num_obj, int32_t maximize,
                                       brkga_problem_t* out)
{
    if (num_obj != 1 && num_obj != 2)
        return set_error(BRKGA_ERR_INVALID_ARG, "host-evaluated problems have 1 or 2 objectives");
    if (num_obj == 2 && maximize)
        return set_error(BRKGA_ERR_INVALID_ARG, "NSGA-II minimises every objective (individual.hpp:99-114)");
    brkga_problem_s* pr;
    brkga_status st = new_problem(engine, KIND_HOST, chrom_len, &pr);
    if (st != BRKGA_OK) return st;
    pr->desc.num_obj = num_obj;
    pr->desc.maximize = maximize ? 1 : 0;
    pr->desc.dim = chrom_len;
    set_scales(&pr->desc, 30, 30);
    return finish_problem(pr, st, out);
}

brkga_status brkga_problem_info(brkga_problem_t problem, int32_t* kind, int32_t* chrom_len, int32_t* num_obj,
                                int32_t* maximize)
{
    if (!problem) return set_error(BRKGA_ERR_INVALID_ARG, "null problem");
    if (kind) *kind = problem->desc.kind;
    if (chrom_len) *chrom_len = problem->desc.chrom_len;
    if (num_obj) *num_obj = problem->desc.num_obj;
    if (maximize) *maximize = problem->desc.maximize;
    return BRKGA_OK;
}

brkga_status brkga_problem_destroy(brkga_problem_t problem)
{
    if (!problem) return BRKGA_OK;
    DeviceGuard g(problem->engine->device);
    (void)problem_ready(problem);
    if (problem->ready) (void)hipEventDestroy(problem->ready);
    free_all(problem->owned);
    delete problem;
    return BRKGA_OK;
}

static brkga_status evaluate_impl(brkga_problem_t problem, const float* d_keys, float* d_fitness, int* d_decoded,
                                  int32_t pop_size, int32_t chrom_len, void* stream)
{
    if (!problem || !d_keys) return set_error(BRKGA_ERR_INVALID_ARG, "null problem or keys");
    if (problem->desc.kind == KIND_HOST) return set_error(BRKGA_ERR_INVALID_ARG, kNoDecoder);
    if (pop_size < 0) return set_error(BRKGA_ERR_INVALID_ARG, "negative pop_size");
    if (chrom_len != problem->desc.chrom_len)
        return set_error(BRKGA_ERR_INVALID_ARG, "chrom_len " + std::to_string(chrom_len) + " does not match the problem (" +
                                                    std::to_string(problem->desc.chrom_len) + ")");
    if (pop_size == 0) return BRKGA_OK;
    DeviceGuard g(problem->engine->device);
    const brkga_status rs = problem_ready(problem);
    if (rs != BRKGA_OK) return rs;
    hipError_t e = launch_evaluate(problem->desc, d_keys, chrom_len, pop_size, d_fitness, d_decoded,
                                   default_block(problem->desc.n), as_stream(stream));
    if (e != hipSuccess) return hip_error(e, "evaluate launch");
    return BRKGA_OK;
}

brkga_status brkga_evaluate(brkga_problem_t problem, const float* d_keys, float* d_fitness, int32_t pop_size,
                            int32_t chrom_len, void* stream)
{
    if (!d_fitness) return set_error(BRKGA_ERR_INVALID_ARG, "null fitness buffer");
    return evaluate_impl(problem, d_keys, d_fitness, nullptr, pop_size, chrom_len, stream);
}

brkga_status brkga_decode(brkga_problem_t problem, const float* d_keys, int32_t* d_decoded, int32_t pop_size,
                          int32_t chrom_len, void* stream)
{
    if (!d_decoded) return set_error(BRKGA_ERR_INVALID_ARG, "null decoded buffer");
    return evaluate_impl(problem, d_keys, nullptr, d_decoded, pop_size, chrom_len, stream);
}

brkga_status brkga_evaluate_host(brkga_problem_t problem, const float* keys, int32_t count, float* fitness,
                                 int32_t* decoded)
{
    if (!problem || !keys || count < 0) return set_error(BRKGA_ERR_INVALID_ARG, "null problem / keys or negative count");
    if (problem->desc.kind == KIND_HOST) return set_error(BRKGA_ERR_INVALID_ARG, kNoDecoder);
    if (count == 0) return BRKGA_OK;
    DeviceGuard g(problem->engine->device);
    const ProblemDesc& p = problem->desc;
    const size_t nk = (size_t)count * p.chrom_len;
    void *dk = nullptr, *df = nullptr, *dd = nullptr;
    brkga_status st = problem_ready(problem);
    if (st != BRKGA_OK) return st;
    st = device_alloc(&dk, nk * sizeof(float), "keys");
    if (st == BRKGA_OK) st = device_alloc(&df, (size_t)count * p.num_obj * sizeof(float), "fitness");
    if (st == BRKGA_OK) st = device_alloc(&dd, nk * sizeof(int), "decoded");
    hipError_t e = hipSuccess;
    if (st == BRKGA_OK) e = hipMemcpy(dk, keys, nk * sizeof(float), hipMemcpyHostToDevice);
    if (st == BRKGA_OK && e == hipSuccess)
        e = launch_evaluate(p, (const float*)dk, p.chrom_len, count, (float*)df, (int*)dd, default_block(p.n), nullptr);
    if (st == BRKGA_OK && e == hipSuccess) e = hipDeviceSynchronize();
    if (st == BRKGA_OK && e == hipSuccess && fitness)
        e = hipMemcpy(fitness, df, (size_t)count * p.num_obj * sizeof(float), hipMemcpyDeviceToHost);
    if (st == BRKGA_OK && e == hipSuccess && decoded) e = hipMemcpy(decoded, dd, nk * sizeof(int), hipMemcpyDeviceToHost);
    if (st == BRKGA_OK && e != hipSuccess) st = hip_error(e, "evaluate_host");
    for (void* q : {dk, df, dd})
        if (q) (void)hipFree(q);
    return st;
}

brkga_status brkga_evaluate_host_noalloc(brkga_problem_t problem, const float* keys, int32_t count, float* fitness,
                                         int32_t* decoded)
{
    if (!problem || !keys || count < 0) return set_error(BRKGA_ERR_INVALID_ARG, "null problem / keys or negative count");
    if (problem->desc.kind == KIND_HOST) return set_error(BRKGA_ERR_INVALID_ARG, kNoDecoder);
    if (count == 0) return BRKGA_OK;
    const ProblemDesc& p = problem->desc;
    const int dev = problem->engine->device;
    const size_t nk = (size_t)count * p.chrom_len;
    const size_t off_fit = align256(nk * sizeof(float));
    const size_t off_dec = off_fit + align256((size_t)count * p.num_obj * sizeof(float));
    const size_t need = off_dec + (decoded ? nk * sizeof(int) : 0);
    if (need > kNoAllocScratchBytes)
        return set_error(BRKGA_ERR_INVALID_ARG, "evaluate_host_noalloc: " + std::to_string(count) + " rows of " +
                                                    std::to_string(p.chrom_len) + " keys need " + std::to_string(need) +
                                                    " bytes of device staging; the static scratch holds " +
                                                    std::to_string(kNoAllocScratchBytes));
    if (dev < 0 || dev >= kMaxScratchDevices || !g_scratch_ptr[dev])
        return set_error(BRKGA_ERR_DEVICE, "static scratch not resolved for this device");
    DeviceGuard g(dev);
    const brkga_status rs = problem_ready(problem);
    if (rs != BRKGA_OK) return rs;
    // every call on a device shares its scratch: calls from several host
    // threads run one after the other
    std::lock_guard<std::mutex> lock(g_scratch_call_mu[dev]);
    unsigned char* base = g_scratch_ptr[dev];
    float* dk = reinterpret_cast<float*>(base);
    float* df = reinterpret_cast<float*>(base + off_fit);
    int* dd = decoded ? reinterpret_cast<int*>(base + off_dec) : nullptr;
    // the decode writes every fitness and decoded entry of the rows it is
    // given, so nothing left by an earlier call is read
    hipError_t e = hipMemcpy(dk, keys, nk * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_evaluate(p, dk, p.chrom_len, count, df, dd, default_block(p.n), nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e == hipSuccess && fitness)
        e = hipMemcpy(fitness, df, (size_t)count * p.num_obj * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess && decoded) e = hipMemcpy(decoded, dd, nk * sizeof(int), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_error(e, "evaluate_host_noalloc");
    return BRKGA_OK;
}

brkga_status brkga_population_create(brkga_problem_t problem, const brkga_params* params, brkga_pop_t* out)
{
    if (!problem || !params || !out) return set_error(BRKGA_ERR_INVALID_ARG, "null argument");
    DeviceGuard g(problem->engine->device);
    brkga_status st = create_population(problem, params, nullptr, nullptr, 0, out);
    if (st != BRKGA_OK) return st;
    hipError_t e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) {
        destroy_population(*out);
        *out = nullptr;
        return hip_error(e, "hipStreamSynchronize");
    }
    return BRKGA_OK;
}

brkga_status brkga_population_destroy(brkga_pop_t pop)
{
    if (!pop) return BRKGA_OK;
    DeviceGuard g(pop->problem->engine->device);
    (void)hipDeviceSynchronize();
    destroy_population(pop);
    return BRKGA_OK;
}

brkga_status brkga_population_initialize(brkga_pop_t pop, void* stream)
{
    if (brkga_status hs = device_path_check(pop)) return hs;
    if (!pop) return set_error(BRKGA_ERR_INVALID_ARG, "null population");
    DeviceGuard g(pop->problem->engine->device);
    hipStream_t s = as_stream(stream);
    BRKGA_HIP_TRY(hipMemsetAsync(pop->host.gen, 0, 4 * sizeof(int), s));
    pop->host_t = 0;
    return pop_step(pop, 1, 0, s);
}

brkga_status brkga_evolve(brkga_pop_t pop, int32_t generations, void* stream)
{
    if (brkga_status hs = device_path_check(pop)) return hs;
    if (!pop) return set_error(BRKGA_ERR_INVALID_ARG, "null population");
    if (generations < 0) return set_error(BRKGA_ERR_INVALID_ARG, "negative generation count");
    DeviceGuard g(pop->problem->engine->device);
    hipStream_t s = as_stream(stream);
    for (int i = 0; i < generations; ++i) {
        brkga_status st = pop_step(pop, 0, pop->host_t & 1, s);
        if (st != BRKGA_OK) return st;
        ++pop->host_t;
    }
    return generations > 0 ? pop_close(pop, pop->host_t & 1, s) : BRKGA_OK;
}

brkga_status brkga_evolve_graph(brkga_pop_t pop, int32_t generations, void* stream)
{
    if (brkga_status hs = device_path_check(pop)) return hs;
    if (!pop) return set_error(BRKGA_ERR_INVALID_ARG, "null population");
    if (generations < 0) return set_error(BRKGA_ERR_INVALID_ARG, "negative generation count");
    if (generations == 0) return BRKGA_OK;
    std::lock_guard<std::mutex> lock(pop->mu);
    DeviceGuard g(pop->problem->engine->device);
    const std::pair<int, int> key(generations, pop->host_t & 1);
    auto it = pop->graphs.find(key);
    if (it == pop->graphs.end()) {
        if (!pop->capture_stream)
            BRKGA_HIP_TRY(hipStreamCreateWithFlags(&pop->capture_stream, hipStreamNonBlocking));
        hipGraph_t graph = nullptr;
        BRKGA_HIP_TRY(hipStreamBeginCapture(pop->capture_stream, hipStreamCaptureModeThreadLocal));
        brkga_status st = BRKGA_OK;
        for (int i = 0; i < generations && st == BRKGA_OK; ++i)
            st = pop_step(pop, 0, (pop->host_t + i) & 1, pop->capture_stream);
        if (st == BRKGA_OK) st = pop_close(pop, (pop->host_t + generations) & 1, pop->capture_stream);
        hipError_t e = hipStreamEndCapture(pop->capture_stream, &graph);
        if (st != BRKGA_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return st;
        }
        if (e != hipSuccess) return hip_error(e, "hipStreamEndCapture");
        hipGraphExec_t exec = nullptr;
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return hip_error(e, "hipGraphInstantiate");
        it = pop->graphs.emplace(key, exec).first;
    }
    BRKGA_HIP_TRY(hipGraphLaunch(it->second, as_stream(stream)));
    pop->host_t += generations;
    return BRKGA_OK;
}

brkga_status brkga_population_generation(brkga_pop_t pop, int32_t* generation)
{
    if (!pop || !generation) return set_error(BRKGA_ERR_INVALID_ARG, "null argument");
    DeviceGuard g(pop->problem->engine->device);
    int gen = 0;
    brkga_status st = read_gen(pop, nullptr, &gen);
    if (st != BRKGA_OK) return st;
    *generation = gen;
    return BRKGA_OK;
}

brkga_status brkga_population_read(brkga_pop_t pop, float* keys, float* fitness, int32_t* order, int32_t* rank,
                                   float* crowding, void* stream)
{
    if (!pop) return set_error(BRKGA_ERR_INVALID_ARG, "null population");
    DeviceGuard g(pop->problem->engine->device);
    hipStream_t s = as_stream(stream);
    int gen = 0;
    brkga_status st = read_gen(pop, s, &gen);
    if (st != BRKGA_OK) return st;
    const InstDesc& d = pop->host;
    const int b = gen & 1;
    const size_t P = (size_t)d.pop;
    if (keys)
        BRKGA_HIP_TRY(hipMemcpy2D(keys, sizeof(float) * d.L, d.keys[b], sizeof(float) * d.ld, sizeof(float) * d.L, P,
                                  hipMemcpyDeviceToHost));
    if (fitness)
        BRKGA_HIP_TRY(hipMemcpy(fitness, d.fit[b], sizeof(float) * P * d.prob.num_obj, hipMemcpyDeviceToHost));
    if (order) BRKGA_HIP_TRY(hipMemcpy(order, d.perm[b], sizeof(int) * P, hipMemcpyDeviceToHost));
    if (pop->nsga2) {
        if (rank) BRKGA_HIP_TRY(hipMemcpy(rank, d.rank[b], sizeof(int) * P, hipMemcpyDeviceToHost));
        if (crowding) BRKGA_HIP_TRY(hipMemcpy(crowding, d.crowd[b], sizeof(float) * P, hipMemcpyDeviceToHost));
    }
    return BRKGA_OK;
}

brkga_status brkga_population_write(brkga_pop_t pop, const float* keys, void* stream)
{
    if (brkga_status hs = device_path_check(pop)) return hs;
    if (!pop || !keys) return set_error(BRKGA_ERR_INVALID_ARG, "null argument");
    DeviceGuard g(pop->problem->engine->device);
    hipStream_t s = as_stream(stream);
    const InstDesc& d = pop->host;
    BRKGA_HIP_TRY(hipStreamSynchronize(s));
    BRKGA_HIP_TRY(hipMemcpy2D(d.keys[0], sizeof(float) * d.ld, keys, sizeof(float) * d.L, sizeof(float) * d.L,
                              (size_t)d.pop, hipMemcpyHostToDevice));
    BRKGA_HIP_TRY(hipMemsetAsync(d.gen, 0, 4 * sizeof(int), s));
    pop->host_t = 0;
    hipError_t e = launch_evaluate(d.prob, d.keys[0], d.ld, d.pop, d.fit[0], nullptr, pop->block, s);
    if (e != hipSuccess) return hip_error(e, "evaluate launch");
    e = pop->nsga2 ? launch_nsga2_select(pop->d_desc, 1, 1, d.pop, s) : launch_partition_one(d, 0, 0, s);
    if (e != hipSuccess) return hip_error(e, "partition launch");
    return BRKGA_OK;
}

brkga_status brkga_population_restore(brkga_pop_t pop, const float* keys, const float* fitness, const int32_t* rank,
                                      const float* crowding, int32_t generation, void* stream)
{
    if (!pop || !keys || !fitness) return set_error(BRKGA_ERR_INVALID_ARG, "null population, keys or fitness");
    if (generation < 0) return set_error(BRKGA_ERR_INVALID_ARG, "negative generation");
    if (pop->nsga2 && (!rank || !crowding))
        return set_error(BRKGA_ERR_INVALID_ARG, "an NSGA-II checkpoint needs its ranks and crowding distances");
    if (pop->host_open) return set_error(BRKGA_ERR_INVALID_ARG, "a host-evaluated generation is open");
    DeviceGuard g(pop->problem->engine->device);
    hipStream_t s = as_stream(stream);
    const InstDesc& d = pop->host;
    const int b = generation & 1;  // step t reads buffer t & 1
    const size_t P = (size_t)d.pop;
    const int M = d.prob.num_obj;
    BRKGA_HIP_TRY(hipStreamSynchronize(s));
    BRKGA_HIP_TRY(hipMemcpy2D(d.keys[b], sizeof(float) * d.ld, keys, sizeof(float) * d.L, sizeof(float) * d.L, P,
                              hipMemcpyHostToDevice));
    BRKGA_HIP_TRY(hipMemcpy(d.fit[b], fitness, sizeof(float) * P * M, hipMemcpyHostToDevice));
    if (pop->nsga2) {
        BRKGA_HIP_TRY(hipMemcpy(d.rank[b], rank, sizeof(int) * P, hipMemcpyHostToDevice));
        BRKGA_HIP_TRY(hipMemcpy(d.crowd[b], crowding, sizeof(float) * P, hipMemcpyHostToDevice));
    }
    // both step counter slots (the ranking steps read the slot of their
    // parity), the gather / arrival counter and the flag
    const int counters[4] = {generation, 0, 0, generation};
    BRKGA_HIP_TRY(hipMemcpy(d.gen, counters, sizeof(counters), hipMemcpyHostToDevice));
    pop->host_t = generation;
    if (!pop->nsga2) {
        // the rank order is the stable partition of the fitness (as after any
        // generation), so the next step reads the same parents
        hipError_t e = launch_partition_one(d, b, 0, s);
        if (e != hipSuccess) return hip_error(e, "partition launch");
    }
    BRKGA_HIP_TRY(hipStreamSynchronize(s));
    return BRKGA_OK;
}

brkga_status brkga_get_best(brkga_pop_t pop, float* fitness, float* keys, int32_t* decoded, void* stream)
{
    if (!pop) return set_error(BRKGA_ERR_INVALID_ARG, "null population");
    DeviceGuard g(pop->problem->engine->device);
    hipStream_t s = as_stream(stream);
    int gen = 0;
    brkga_status st = read_gen(pop, s, &gen);
    if (st != BRKGA_OK) return st;
    const InstDesc& d = pop->host;
    const int b = gen & 1;
    const int P = d.pop, M = d.prob.num_obj;
    std::vector<float> fit((size_t)P * M);
    BRKGA_HIP_TRY(hipMemcpy(fit.data(), d.fit[b], sizeof(float) * fit.size(), hipMemcpyDeviceToHost));
    std::vector<int> ord(P);
    std::vector<float> crowd;
    if (pop->nsga2) {
        crowd.resize(P);
        BRKGA_HIP_TRY(hipMemcpy(ord.data(), d.rank[b], sizeof(int) * P, hipMemcpyDeviceToHost));
        BRKGA_HIP_TRY(hipMemcpy(crowd.data(), d.crowd[b], sizeof(float) * P, hipMemcpyDeviceToHost));
    } else {
        BRKGA_HIP_TRY(hipMemcpy(ord.data(), d.perm[b], sizeof(int) * P, hipMemcpyDeviceToHost));
    }
    const int best = pick_best(d, pop->nsga2, fit.data(), ord.data(), crowd.data());
    if (best < 0) return set_error(BRKGA_ERR_DEVICE, "no best individual found");
    if (fitness)
        for (int m = 0; m < M; ++m) fitness[m] = fit[(size_t)best * M + m];
    const float* row = d.keys[b] + (size_t)best * d.ld;
    if (keys) BRKGA_HIP_TRY(hipMemcpy(keys, row, sizeof(float) * d.L, hipMemcpyDeviceToHost));
    if (decoded && host_kind(pop)) return set_error(BRKGA_ERR_INVALID_ARG, kNoDecoder);
    if (decoded) {
        hipError_t e = launch_evaluate(d.prob, row, d.ld, 1, pop->d_fit_row, pop->d_decode_row, pop->block, s);
        if (e != hipSuccess) return hip_error(e, "decode launch");
        BRKGA_HIP_TRY(hipStreamSynchronize(s));
        BRKGA_HIP_TRY(hipMemcpy(decoded, pop->d_decode_row, sizeof(int) * d.L, hipMemcpyDeviceToHost));
    }
    return BRKGA_OK;
}

brkga_status brkga_population_device_buffers(brkga_pop_t pop, float** d_keys, float** d_fitness, int32_t** d_order,
                                             int32_t* ld)
{
    if (!pop) return set_error(BRKGA_ERR_INVALID_ARG, "null population");
    DeviceGuard g(pop->problem->engine->device);
    int gen = 0;
    brkga_status st = read_gen(pop, nullptr, &gen);
    if (st != BRKGA_OK) return st;
    const int b = gen & 1;
    if (d_keys) *d_keys = pop->host.keys[b];
    if (d_fitness) *d_fitness = pop->host.fit[b];
    if (d_order) *d_order = pop->host.perm[b];
    if (ld) *ld = pop->host.ld;
    return BRKGA_OK;
}

brkga_status brkga_population_autotune(brkga_pop_t pop, int32_t reps, int32_t* block, void* stream)
{
    if (pop && host_kind(pop)) return set_error(BRKGA_ERR_INVALID_ARG, kNoDecoder);
    if (!pop || reps <= 0) return set_error(BRKGA_ERR_INVALID_ARG, "null population or no repetitions");
    std::lock_guard<std::mutex> lock(pop->mu);
    DeviceGuard g(pop->problem->engine->device);
    hipStream_t s = as_stream(stream);
    const InstDesc& d = pop->host;
    const int n = d.prob.n;
    hipEvent_t a, b;
    BRKGA_HIP_TRY(hipEventCreate(&a));
    hipError_t e = hipEventCreate(&b);
    if (e != hipSuccess) {
        (void)hipEventDestroy(a);
        return hip_error(e, "hipEventCreate");
    }
    const int cur = pop->host_t & 1;
    const int keep = pop->block;
    int best = keep;
    float best_ms = 0.0f;
    bool have = false;
    // the step (BRKGA) / offspring (NSGA-II) kernel rewrites the same rows
    // from unchanged inputs on every launch: timing it changes nothing.  A
    // BRKGA generation is timed as it would run at that workgroup size: the
    // ranking step alone, or the step and a partition of the rows it wrote
    // (mode 0: the counter stays).  A ranking step writes the next counter
    // slot; the closing partition below sets both slots back.
    const size_t lds_limit = pop->problem->engine->lds_limit;
    bool ranked = false;
    const int keep_ch = pop->children;
    int best_ch = keep_ch;
    auto launch = [&](int blk, int ch) {
        if (pop->nsga2)
            return launch_nsga2_offspring(d.prob.kind, blk, pop->d_desc, 1, d.pop, pop->step_lds, 0, s);
        if (step_can_rank(d, blk, lds_limit)) {
            ranked = true;
            return launch_step_one(d, blk, cur, 0, 1, s, nullptr, nullptr, ch);
        }
        hipError_t r = launch_step_one(d, blk, cur, 0, 0, s);
        return r == hipSuccess ? launch_partition_one(d, cur ^ 1, 0, s) : r;
    };
    for (int blk = 64; blk <= 1024 && e == hipSuccess; blk <<= 1) {
        if (blk * kSortItems < n) continue;
        for (int ch = 1; ch <= 4 && e == hipSuccess; ch <<= 1) {
            if (ch > 1 && (pop->nsga2 || !step_multi_ok(d, blk, ch, lds_limit))) continue;
            e = launch(blk, ch);
            if (e == hipSuccess) e = hipEventRecord(a, s);
            for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch(blk, ch);
            if (e == hipSuccess) e = hipEventRecord(b, s);
            if (e == hipSuccess) e = hipEventSynchronize(b);
            float ms = 0.0f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
            if (e == hipSuccess && (!have || ms < best_ms)) {
                have = true;
                best_ms = ms;
                best = blk;
                best_ch = ch;
            }
        }
    }
    if (e == hipSuccess && ranked) e = launch_partition_one(d, cur, 3, s);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (e != hipSuccess) return hip_error(e, "autotune launches");
    if (best != keep || best_ch != keep_ch) {
        pop->block = best;
        pop->children = best_ch;
        for (auto& kv : pop->graphs) (void)hipGraphExecDestroy(kv.second);
        pop->graphs.clear();
    }
    if (block) *block = best;
    return BRKGA_OK;
}

brkga_status brkga_population_set_block(brkga_pop_t pop, int32_t block)
{
    if (!pop) return set_error(BRKGA_ERR_INVALID_ARG, "null population");
    const int n = pop->host.prob.n;
    if (block != 64 && block != 128 && block != 256 && block != 512 && block != 1024)
        return set_error(BRKGA_ERR_INVALID_ARG, "block must be 64, 128, 256, 512 or 1024");
    if ((long long)block * kSortItems < n)
        return set_error(BRKGA_ERR_INVALID_ARG, "block too small: a workgroup decodes at most 16 keys per thread");
    if (block != pop->block) {
        pop->block = block;
        for (auto& kv : pop->graphs) (void)hipGraphExecDestroy(kv.second);
        pop->graphs.clear();
    }
    return BRKGA_OK;
}

brkga_status brkga_population_host_begin(brkga_pop_t pop, int32_t initialize, float* keys, float* fitness,
                                         int32_t* needs, void* stream)
{
    if (!pop || !keys) return set_error(BRKGA_ERR_INVALID_ARG, "null population or keys");
    if (!host_kind(pop)) return set_error(BRKGA_ERR_INVALID_ARG, "the problem has a device decoder: use brkga_evolve");
    if (pop->host_open) return set_error(BRKGA_ERR_INVALID_ARG, "a host-evaluated generation is already open");
    DeviceGuard g(pop->problem->engine->device);
    hipStream_t s = as_stream(stream);
    const InstDesc& d = pop->host;
    const int M = d.prob.num_obj;
    if (initialize) {
        BRKGA_HIP_TRY(hipMemsetAsync(d.gen, 0, 4 * sizeof(int), s));
        pop->host_t = 0;
    }
    const int cur = pop->host_t & 1;
    hipError_t e;
    const float* rows;
    const float* fit;
    if (pop->nsga2) {
        // generation 0, or the N children of step host_t (rows [N, 2N))
        e = launch_nsga2_offspring(KIND_HOST, pop->block, pop->d_desc, 1, d.pop, pop->step_lds, initialize, s);
        const int buf = initialize ? 0 : cur;
        const size_t first = initialize ? 0 : (size_t)d.pop;
        rows = d.keys[buf] + first * d.ld;
        fit = d.fit[buf] + first * M;
    } else {
        e = launch_host_children(d, cur, initialize, s);
        const int nxt = initialize ? 0 : cur ^ 1;
        rows = d.keys[nxt];
        fit = d.fit[nxt];
    }
    if (e != hipSuccess) return hip_error(e, "kernel launch");
    BRKGA_HIP_TRY(hipStreamSynchronize(s));
    BRKGA_HIP_TRY(hipMemcpy2D(keys, sizeof(float) * d.L, rows, sizeof(float) * d.ld, sizeof(float) * d.L,
                              (size_t)d.pop, hipMemcpyDeviceToHost));
    // rows the step copied (elite, migrants) keep their fitness; the caller
    // evaluates the others (population.hpp:626-653 skips evaluated members)
    const int t = pop->host_t;
    const bool migrates = !pop->nsga2 && !initialize && d.islands > 1 && d.mig_k > 0 && d.mig_interval > 0 &&
                          t > 0 && t % d.mig_interval == 0;
    std::vector<int32_t> need((size_t)d.pop, 1);
    if (!pop->nsga2 && !initialize)
        for (int i = 0; i < d.pop; ++i) {
            const int sl = i % d.Pi;
            if (sl < d.Ei || (migrates && sl >= d.Pi - d.mig_k)) need[i] = 0;
        }
    if (fitness) {
        if (pop->nsga2 || initialize) {
            std::fill(fitness, fitness + (size_t)d.pop * M, 0.0f);
        } else {
            BRKGA_HIP_TRY(hipMemcpy(fitness, fit, sizeof(float) * (size_t)d.pop * M, hipMemcpyDeviceToHost));
            for (int i = 0; i < d.pop; ++i)
                if (need[i])
                    for (int m = 0; m < M; ++m) fitness[(size_t)i * M + m] = 0.0f;
        }
    }
    if (needs) std::copy(need.begin(), need.end(), needs);
    pop->host_open = initialize ? 1 : 2;
    return BRKGA_OK;
}

brkga_status brkga_population_host_end(brkga_pop_t pop, const float* fitness, void* stream)
{
    if (!pop || !fitness) return set_error(BRKGA_ERR_INVALID_ARG, "null population or fitness");
    if (!pop->host_open) return set_error(BRKGA_ERR_INVALID_ARG, "no host-evaluated generation is open");
    DeviceGuard g(pop->problem->engine->device);
    hipStream_t s = as_stream(stream);
    const InstDesc& d = pop->host;
    const int M = d.prob.num_obj;
    const int init = pop->host_open == 1;
    const int cur = pop->host_t & 1;
    float* dst;
    if (pop->nsga2) dst = d.fit[init ? 0 : cur] + (init ? 0 : (size_t)d.pop * M);
    else dst = d.fit[init ? 0 : cur ^ 1];
    BRKGA_HIP_TRY(hipMemcpyAsync(dst, fitness, sizeof(float) * (size_t)d.pop * M, hipMemcpyHostToDevice, s));
    hipError_t e;
    if (pop->nsga2) {
        e = launch_nsga2_select(pop->d_desc, 1, init, d.pop, s);
        if (e == hipSuccess && !init) e = launch_nsga2_gather(pop->d_desc, 1, d.pop, s);
    } else {
        e = launch_partition_one(d, init ? 0 : cur ^ 1, init ? 0 : 1, s);
    }
    if (e != hipSuccess) return hip_error(e, "kernel launch");
    BRKGA_HIP_TRY(hipStreamSynchronize(s));
    if (!init) ++pop->host_t;
    pop->host_open = 0;
    return BRKGA_OK;
}

brkga_status brkga_population_set_children(brkga_pop_t pop, int32_t children)
{
    if (!pop) return set_error(BRKGA_ERR_INVALID_ARG, "null population");
    if (children != 1 && children != 2 && children != 4)
        return set_error(BRKGA_ERR_INVALID_ARG, "children per workgroup must be 1, 2 or 4");
    if (children != pop->children) {
        pop->children = children;
        for (auto& kv : pop->graphs) (void)hipGraphExecDestroy(kv.second);
        pop->graphs.clear();
    }
    return BRKGA_OK;
}

brkga_status brkga_population_step_shape(brkga_pop_t pop, int32_t* block, int32_t* children)
{
    if (!pop) return set_error(BRKGA_ERR_INVALID_ARG, "null population");
    if (block) *block = pop->block;
    if (children) *children = pop_ranks(pop) ? pop_children(pop) : 1;
    return BRKGA_OK;
}

brkga_status brkga_population_profile(brkga_pop_t pop, int32_t reps, double* step_us, double* close_us,
                                      void* stream)
{
    if (pop && host_kind(pop)) return set_error(BRKGA_ERR_INVALID_ARG, kNoDecoder);
    if (!pop || reps <= 0) return set_error(BRKGA_ERR_INVALID_ARG, "null population or no repetitions");
    DeviceGuard g(pop->problem->engine->device);
    hipStream_t s = as_stream(stream);
    const InstDesc& d = pop->host;
    // reps real generations, enqueued as brkga_evolve does.  BRKGA: the step
    // and partition kernels take start / stop events from their own
    // dispatch (no marker packets between kernels); steps that rank their
    // parents have no partition per generation (close_us = 0), one closes
    // the run.  NSGA-II: marker events around the offspring kernel and
    // around selection + gather.  One synchronisation at the end; the
    // population advances reps generations.
    const bool timed = !pop->nsga2 && d.ls_interval <= 0;
    const bool ranks = pop_ranks(pop);
    std::vector<hipEvent_t> ev(4 * (size_t)reps, nullptr);
    hipError_t e = hipSuccess;
    for (auto& x : ev)
        if (e == hipSuccess) e = hipEventCreate(&x);
    for (int i = 0; i < reps && e == hipSuccess; ++i) {
        hipEvent_t* q = &ev[4 * (size_t)i];
        const int cur = pop->host_t & 1;
        if (timed) {
            e = launch_step_one(d, pop->block, cur, 0, ranks, s, q[0], q[1], ranks ? pop_children(pop) : 1);
            if (e == hipSuccess && !ranks) e = launch_partition_one(d, cur ^ 1, 1, s, q[2], q[3]);
        } else {
            e = hipEventRecord(q[0], s);
            if (e == hipSuccess) {
                if (pop->nsga2)
                    e = launch_nsga2_offspring(d.prob.kind, pop->block, pop->d_desc, 1, d.pop, pop->step_lds, 0, s);
                else
                    e = launch_step_one(d, pop->block, cur, 0, 0, s);
            }
            if (e == hipSuccess) e = hipEventRecord(q[1], s);
            if (e == hipSuccess) e = hipEventRecord(q[2], s);
            if (e == hipSuccess) {
                if (pop->nsga2) {
                    e = launch_nsga2_select(pop->d_desc, 1, 0, d.pop, s);
                    if (e == hipSuccess) e = launch_nsga2_gather(pop->d_desc, 1, d.pop, s);
                    if (e == hipSuccess && d.div_interval > 0)
                        e = launch_nsga2_diversity(d.prob.kind, pop->block, pop->d_desc, 1, d.pop, pop->step_lds,
                                                   d.pop, s);
                } else {
                    e = launch_partition_one(d, cur ^ 1, 1, s);
                    if (e == hipSuccess)
                        e = launch_tspj_local_search(pop->d_desc, 1, d.islands * d.ls_k, pop->block,
                                                     tspj_ls_lds_bytes(d.prob.n, d.nb), s);
                    if (e == hipSuccess) e = launch_partition_one(d, -1, 2, s);
                }
            }
            if (e == hipSuccess) e = hipEventRecord(q[3], s);
        }
        ++pop->host_t;
    }
    if (e == hipSuccess && ranks) {
        const brkga_status st = pop_close(pop, pop->host_t & 1, s);
        if (st != BRKGA_OK) {
            for (auto& x : ev)
                if (x) (void)hipEventDestroy(x);
            return st;
        }
    }
    double sa = 0.0, sb = 0.0;
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    for (int i = 0; i < reps && e == hipSuccess; ++i) {
        const hipEvent_t* q = &ev[4 * (size_t)i];
        float a = 0.f, b = 0.f;
        e = hipEventElapsedTime(&a, q[0], q[1]);
        if (e == hipSuccess && !(timed && ranks)) e = hipEventElapsedTime(&b, q[2], q[3]);
        sa += a;
        sb += b;
    }
    for (auto& x : ev)
        if (x) (void)hipEventDestroy(x);
    if (e != hipSuccess) return hip_error(e, "profiled generations");
    if (step_us) *step_us = 1000.0 * sa / reps;
    if (close_us) *close_us = 1000.0 * sb / reps;
    return BRKGA_OK;
}

}  // extern "C"
