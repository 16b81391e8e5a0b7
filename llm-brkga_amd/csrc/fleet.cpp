// fleet.cpp - resident populations advanced one generation per tick.
#include "fleet.hpp"

#include <algorithm>
#include <cstdlib>
#include <map>
#include <numeric>
#include <tuple>

#include "launch.hpp"

namespace brkga {

namespace {

constexpr int kQueueStride = 32;  // ints between two buckets' counters
constexpr int kBlocks[] = {64, 128, 256, 512, 1024};

// Lane balancing charges a child n + kChildFixedWork: the build, the
// barriers and the reductions of a child cost about as much as ranking a
// thousand keys, so buckets of small instances take longer than their key
// count says.
constexpr double kChildFixedWork = 1024.0;

// Lane calibration: the probe's length (about 60 us of dependent ALU steps)
// and how many candidate streams are tried for the kLanes - 1 lanes.
constexpr int kProbeIters = 15000;
constexpr int kLaneCandidates = 8;
constexpr int kProbeRounds = 3;
constexpr int kChunks[] = {1, 2, 4, 8, 16};

// The autotuner keeps the fastest of this many timed rounds per launch
// shape, so one slow round (clocks ramping, a neighbouring kernel) does not
// decide the shape.
constexpr int kTuneRounds = 3;

// Default slots per work item (size class n, a power of two), from the
// tuner's picks on the served mix: small children want one item each (the
// first item of a workgroup needs no queue atomic, so short items cost
// only their staging), mid-size ones a few slots to share the staging of
// the island's rank order and coordinates, the largest one again.
int default_chunk(int n)
{
    if (n <= 256) return 1;
    if (n <= 512) return 2;
    if (n <= 2048) return 4;
    if (n <= 4096) return 2;
    return 1;
}

}  // namespace

int fleet_key(int n)
{
    int k = 64;
    while (k < n) k <<= 1;
    return k;
}

Fleet::~Fleet()
{
    DeviceGuard g(eng_->device);
    drop_graphs();
    if (capture_) (void)hipStreamDestroy(capture_);
    for (int k = 1; k < kLanes; ++k) {
        if (lanes_[k]) (void)hipStreamDestroy(lanes_[k]);
        if (join_[k]) (void)hipEventDestroy(join_[k]);
    }
    if (fork_) (void)hipEventDestroy(fork_);
    if (d_descs_) (void)hipFree(d_descs_);
    if (d_ls_descs_) (void)hipFree(d_ls_descs_);
    if (d_items_) (void)hipFree(d_items_);
    if (d_segs_) (void)hipFree(d_segs_);
    if (d_queues_) (void)hipFree(d_queues_);
}

void Fleet::drop_graphs()
{
    for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
    graphs_.clear();
}

brkga_status Fleet::reserve(void** p, size_t* cap, size_t bytes, const char* what)
{
    if (bytes <= *cap && *p) return BRKGA_OK;
    size_t want = std::max<size_t>(bytes, 2 * *cap);
    want = std::max<size_t>(want, 4096);
    if (*p) {
        // hipFree waits for the device; kernels of earlier ticks may still read it
        (void)hipFree(*p);
        *p = nullptr;
        *cap = 0;
    }
    brkga_status st = device_alloc(p, want, what);
    if (st != BRKGA_OK) return st;
    *cap = want;
    return BRKGA_OK;
}

brkga_status Fleet::set_members(const std::vector<brkga_pop_s*>& pops, hipStream_t s)
{
    DeviceGuard g(eng_->device);
    for (auto* p : pops)
        if (!p) return set_error(BRKGA_ERR_INVALID_ARG, "null population in fleet");
    // earlier uploads read descs_ / items_ asynchronously
    BRKGA_HIP_TRY(hipStreamSynchronize(s));
    // BRKGA members by (kind, size class, n descending), NSGA-II by (kind, n descending)
    std::vector<brkga_pop_s*> order(pops);
    std::stable_sort(order.begin(), order.end(), [](const brkga_pop_s* a, const brkga_pop_s* b) {
        if (a->nsga2 != b->nsga2) return a->nsga2 < b->nsga2;
        const int ka = a->host.prob.kind, kb = b->host.prob.kind;
        if (ka != kb) return ka < kb;
        const int na = a->host.prob.n, nb = b->host.prob.n;
        if (!a->nsga2 && fleet_key(na) != fleet_key(nb)) return fleet_key(na) > fleet_key(nb);
        return na > nb;
    });
    members_ = order;
    descs_.assign(members_.size(), InstDesc());
    buckets_.clear();
    groups_.clear();
    n_brkga_ = 0;
    total_islands_ = 0;
    max_Pi_ = 0;
    for (size_t i = 0; i < members_.size(); ++i) {
        const brkga_pop_s* p = members_[i];
        InstDesc d = p->host;
        d.slot_begin = 0;
        d.seg_begin = 0;
        const int n = d.prob.n;
        if (!p->nsga2) {
            d.seg_begin = total_islands_;
            total_islands_ += d.islands;
            max_Pi_ = std::max(max_Pi_, d.Pi);
            ++n_brkga_;
            const int key = fleet_key(n);
            if (buckets_.empty() || buckets_.back().kind != d.prob.kind || buckets_.back().key != key) {
                FleetBucket b{};
                b.kind = d.prob.kind;
                b.key = key;
                b.first = (int)i;
                b.n_min = n;
                b.n_max = n;
                std::lock_guard<std::mutex> lock(eng_->tuned_mu);
                auto it = eng_->tuned.find({b.kind, key});
                if (it != eng_->tuned.end()) {
                    b.block = it->second.block;
                    b.chunk = it->second.chunk;
                    b.xys = it->second.xys;
                    b.child_us = it->second.child_us;
                    b.tuned_children = it->second.children;
                } else {
                    b.block = default_block(key);
                    b.chunk = default_chunk(key);
                    b.xys = b.kind == KIND_TSP_COORDS && key <= 2048;
                }
                buckets_.push_back(b);
            }
            FleetBucket& b = buckets_.back();
            ++b.count;
            b.children += d.pop;
            b.work += (double)d.pop * (n + kChildFixedWork);
            b.n_min = std::min(b.n_min, n);
            b.n_max = std::max(b.n_max, n);
            b.max_Pi = std::max(b.max_Pi, d.Pi);
            b.lds = std::max(b.lds, step_layout(d.prob.kind, n, d.L, d.nb, d.Pi, b.xys, d.prob.aux).bytes);
        } else {
            if (groups_.empty() || groups_.back().kind != d.prob.kind) {
                FleetGroup gr{};
                gr.kind = d.prob.kind;
                gr.first = (int)i;
                groups_.push_back(gr);
            }
            FleetGroup& gr = groups_.back();
            d.slot_begin = gr.rows;
            gr.rows += d.pop;
            ++gr.count;
            gr.block = std::max(gr.block, p->block);
            gr.lds = std::max(gr.lds, p->step_lds);
            gr.max_pop = std::max(gr.max_pop, d.pop);
            gr.diversity = gr.diversity || d.div_interval > 0;
        }
        descs_[i] = d;
    }
    ls_descs_.clear();
    ls_items_ = ls_block_ = ls_lds_ = 0;
    for (int i = 0; i < n_brkga_; ++i) {
        const InstDesc& d = descs_[i];
        if (d.ls_interval <= 0 || d.ls_k <= 0) continue;
        ls_descs_.push_back(d);
        ls_items_ = std::max(ls_items_, d.islands * d.ls_k);
        ls_block_ = std::max(ls_block_, default_block(d.prob.n));
        ls_lds_ = std::max(ls_lds_, tspj_ls_lds_bytes(d.prob.n, d.nb));
    }
    for (auto& b : buckets_) {
        // the decoders hold up to 16 keys per thread
        while (b.block * kSortItems < b.n_max) b.block <<= 1;
        if ((size_t)b.lds > eng_->lds_limit)
            return set_error(BRKGA_ERR_INVALID_ARG, "fleet member too large for the LDS kernels");
    }
    form_launches();
    brkga_status st = ensure_lanes(s);
    if (st != BRKGA_OK) return st;
    assign_lanes();
    drop_graphs();
    st = reserve(reinterpret_cast<void**>(&d_descs_), &descs_cap_, sizeof(InstDesc) * descs_.size(),
                              "fleet descriptors");
    if (st != BRKGA_OK) return st;
    st = reserve(reinterpret_cast<void**>(&d_queues_), &queues_cap_, sizeof(int) * kQueueStride * 2 * buckets_.size(),
                 "fleet queues");
    if (st != BRKGA_OK) return st;
    if (!descs_.empty())
        BRKGA_HIP_TRY(hipMemcpyAsync(d_descs_, descs_.data(), sizeof(InstDesc) * descs_.size(),
                                     hipMemcpyHostToDevice, s));
    if (!ls_descs_.empty()) {
        st = reserve(reinterpret_cast<void**>(&d_ls_descs_), &ls_descs_cap_, sizeof(InstDesc) * ls_descs_.size(),
                     "fleet local-search descriptors");
        if (st != BRKGA_OK) return st;
        BRKGA_HIP_TRY(hipMemcpyAsync(d_ls_descs_, ls_descs_.data(), sizeof(InstDesc) * ls_descs_.size(),
                                     hipMemcpyHostToDevice, s));
    }
    BRKGA_HIP_TRY(hipMemsetAsync(d_queues_, 0, queues_cap_, s));
    return build_items(s);
}

brkga_status Fleet::build_items(hipStream_t s)
{
    // the previous upload from items_ must be done before it is rewritten
    BRKGA_HIP_TRY(hipStreamSynchronize(s));
    std::vector<int4>& items = items_;
    items.clear();
    for (auto& L : launches_) {
        L.item_begin = (int)items.size();
        L.n_max = 0;
        L.lds = 0;
        for (int bi : L.buckets) {
            FleetBucket& b = buckets_[bi];
            b.item_begin = (int)items.size();
            for (int m = b.first; m < b.first + b.count; ++m) {
                const InstDesc& d = descs_[m];
                for (int isl = 0; isl < d.islands; ++isl)
                    for (int s0 = 0; s0 < d.Pi; s0 += b.chunk)
                        items.push_back(
                            make_int4(m, isl * d.Pi + s0, isl * d.Pi + std::min(s0 + b.chunk, d.Pi), 0));
            }
            b.item_count = (int)items.size() - b.item_begin;
            const int per_cu = step_queue_blocks_per_cu(b.kind, b.block, b.xys, b.n_max, b.lds);
            if (per_cu <= 0) return set_error(BRKGA_ERR_DEVICE, "queue step kernel does not fit on a CU");
            b.grid = std::min(b.item_count, per_cu * eng_->cu_count);
            L.n_max = std::max(L.n_max, b.n_max);
            L.lds = std::max(L.lds, b.lds);
        }
        L.item_count = (int)items.size() - L.item_begin;
        const int per_cu = step_queue_blocks_per_cu(L.kind, L.block, L.xys, L.n_max, L.lds);
        if (per_cu <= 0) return set_error(BRKGA_ERR_DEVICE, "queue step kernel does not fit on a CU");
        L.grid = std::min(L.item_count, per_cu * eng_->cu_count);
    }
    brkga_status st =
        reserve(reinterpret_cast<void**>(&d_items_), &items_cap_, sizeof(int4) * items.size(), "fleet work items");
    if (st != BRKGA_OK) return st;
    if (!items.empty())
        BRKGA_HIP_TRY(hipMemcpyAsync(d_items_, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice, s));
    // the islands each lane ranks: those of its launches' members
    segs_.clear();
    for (int l = 0; l < kLanes; ++l) {
        lane_seg_begin_[l] = (int)segs_.size();
        for (const auto& L : launches_) {
            if (L.lane != l) continue;
            for (int bi : L.buckets) {
                const FleetBucket& b = buckets_[bi];
                for (int m = b.first; m < b.first + b.count; ++m)
                    for (int isl = 0; isl < descs_[m].islands; ++isl) segs_.push_back(make_int2(m, isl));
            }
        }
        lane_seg_count_[l] = (int)segs_.size() - lane_seg_begin_[l];
    }
    st = reserve(reinterpret_cast<void**>(&d_segs_), &segs_cap_, sizeof(int2) * segs_.size(), "fleet island list");
    if (st != BRKGA_OK) return st;
    if (!segs_.empty())
        BRKGA_HIP_TRY(hipMemcpyAsync(d_segs_, segs_.data(), sizeof(int2) * segs_.size(), hipMemcpyHostToDevice, s));
    return BRKGA_OK;
}

// Buckets whose shapes select the same kernel share a launch; buckets stay
// in their order (largest n first) inside it.
void Fleet::form_launches()
{
    launches_.clear();
    std::map<std::tuple<int, int, bool, int>, int> index;
    for (int i = 0; i < (int)buckets_.size(); ++i) {
        const FleetBucket& b = buckets_[i];
        const auto key = std::make_tuple(b.kind, b.block, b.xys, queue_items_for(b.kind, b.block, b.n_max));
        auto it = index.find(key);
        if (it == index.end()) {
            FleetLaunch L{};
            L.kind = b.kind;
            L.block = b.block;
            L.xys = b.xys;
            launches_.push_back(L);
            it = index.emplace(key, (int)launches_.size() - 1).first;
        }
        launches_[it->second].buckets.push_back(i);
    }
}

// Largest-work-first onto the least-loaded lane, one launch at a time; with
// NSGA-II members the last lane is theirs.  The work of a bucket is its
// measured time (child slots x tuned time per child) when every bucket has
// been tuned, else population x (n + kChildFixedWork).
void Fleet::assign_lanes()
{
    const int brkga_lanes = groups_.empty() || nlanes_ < 2 ? nlanes_ : nlanes_ - 1;
    nsga_lane_ = groups_.empty() || buckets_.empty() || nlanes_ < 2 ? 0 : nlanes_ - 1;
    // A bucket's measured time per child is trusted when the measurement ran
    // over at least half its current children: measured on a few members
    // (a size class the server first saw with one request) a launch is
    // latency-bound and its time per child is many times the rate of a full
    // bucket.  Other buckets are costed by the work model, scaled to
    // microseconds by the trusted buckets' measured rate.
    auto trusted = [&](const FleetBucket& b) { return b.child_us > 0.0 && 2 * b.tuned_children >= b.children; };
    double us = 0.0, work = 0.0;
    for (const auto& b : buckets_)
        if (trusted(b)) {
            us += b.children * b.child_us;
            work += b.work;
        }
    const double scale = work > 0.0 ? us / work : 1.0;
    for (auto& L : launches_) {
        L.cost = 0.0;
        for (int bi : L.buckets) {
            const FleetBucket& b = buckets_[bi];
            L.cost += trusted(b) ? b.children * b.child_us : b.work * scale;
        }
    }
    std::vector<int> idx(launches_.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return launches_[a].cost > launches_[b].cost; });
    double load[kLanes] = {};
    for (int i : idx) {
        int best = 0;
        for (int k = 1; k < brkga_lanes; ++k)
            if (load[k] < load[best]) best = k;
        launches_[i].lane = best;
        for (int bi : launches_[i].buckets) buckets_[bi].lane = best;
        load[best] += launches_[i].cost;
    }
}

// Lanes wanted by the environment: BRKGA_FLEET_LANES = 1 .. kLanes fixes
// the count (no probing); anything else, or unset, leaves it to the probes.
static int fixed_lanes()
{
    const char* v = std::getenv("BRKGA_FLEET_LANES");
    if (!v || !*v) return 0;
    char* end = nullptr;
    const long n = std::strtol(v, &end, 10);
    return (end && *end == 0 && n >= 1 && n <= kLanes) ? (int)n : 0;
}

brkga_status Fleet::ensure_lanes(hipStream_t s)
{
    if (calibrated_) return BRKGA_OK;
    // events first (kept if the calibration below fails; it runs again on
    // the next call)
    if (!fork_) BRKGA_HIP_TRY(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
    for (int k = 1; k < kLanes; ++k)
        if (!join_[k]) BRKGA_HIP_TRY(hipEventCreateWithFlags(&join_[k], hipEventDisableTiming));
    const int fixed = fixed_lanes();
    if (fixed > 0) {
        int found = 1;
        hipError_t e = hipSuccess;
        for (; found < fixed && e == hipSuccess; ++found) {
            if (lanes_[found]) continue;
            e = hipStreamCreateWithFlags(&lanes_[found], hipStreamNonBlocking);
        }
        if (e != hipSuccess) return hip_error(e, "stream lanes");
        nlanes_ = fixed;
        calibrated_ = true;
        return BRKGA_OK;
    }
    for (int k = 1; k < kLanes; ++k)
        if (lanes_[k]) {
            (void)hipStreamDestroy(lanes_[k]);
            lanes_[k] = nullptr;
        }
    float* probe = nullptr;
    hipEvent_t e0 = nullptr, ea = nullptr, eb = nullptr;
    std::vector<hipStream_t> rejected;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&probe), 2 * 64 * sizeof(float));
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&ea);
    if (e == hipSuccess) e = hipEventCreate(&eb);
    // start-to-end time of probes launched together on a and (unless null) b
    auto timed = [&](hipStream_t a, hipStream_t b, float* ms) {
        hipError_t r = hipEventRecord(e0, s);
        if (r == hipSuccess && a != s) r = hipStreamWaitEvent(a, e0, 0);
        if (r == hipSuccess && b) r = hipStreamWaitEvent(b, e0, 0);
        if (r == hipSuccess) r = launch_lane_probe(kProbeIters, probe, a);
        if (r == hipSuccess && b) r = launch_lane_probe(kProbeIters, probe + 64, b);
        if (r == hipSuccess) r = hipEventRecord(ea, a);
        if (r == hipSuccess && b) r = hipEventRecord(eb, b);
        if (r == hipSuccess) r = hipEventSynchronize(ea);
        if (r == hipSuccess && b) r = hipEventSynchronize(eb);
        float ta = 0.0f, tb = 0.0f;
        if (r == hipSuccess) r = hipEventElapsedTime(&ta, e0, ea);
        if (r == hipSuccess && b) r = hipEventElapsedTime(&tb, e0, eb);
        *ms = std::max(ta, tb);
        return r;
    };
    // medians of kProbeRounds timings, so one disturbed probe (other work
    // on the device while the fleet is created) decides nothing alone
    auto median_timed = [&](hipStream_t a, hipStream_t b, float* ms) {
        float t[kProbeRounds];
        hipError_t r = hipSuccess;
        for (int i = 0; i < kProbeRounds && r == hipSuccess; ++i) r = timed(a, b, &t[i]);
        std::sort(t, t + kProbeRounds);
        *ms = t[kProbeRounds / 2];
        return r;
    };
    float alone = 0.0f;
    if (e == hipSuccess) e = timed(s, nullptr, &alone);  // warm-up
    if (e == hipSuccess) e = median_timed(s, nullptr, &alone);
    int found = 1;
    for (int c = 0; c < kLaneCandidates && found < kLanes && e == hipSuccess; ++c) {
        hipStream_t cand = nullptr;
        e = hipStreamCreateWithFlags(&cand, hipStreamNonBlocking);
        if (e != hipSuccess) break;
        bool concurrent = true;
        for (int k = 0; k < found && concurrent && e == hipSuccess; ++k) {
            float ms = 0.0f;
            e = median_timed(k == 0 ? s : lanes_[k], cand, &ms);
            concurrent = ms < 1.5f * alone;
        }
        if (concurrent && e == hipSuccess) lanes_[found++] = cand;
        else rejected.push_back(cand);
    }
    for (hipStream_t r : rejected) (void)hipStreamDestroy(r);
    if (e0) (void)hipEventDestroy(e0);
    if (ea) (void)hipEventDestroy(ea);
    if (eb) (void)hipEventDestroy(eb);
    if (probe) (void)hipFree(probe);
    if (e != hipSuccess) {
        for (int k = 1; k < found; ++k) {
            (void)hipStreamDestroy(lanes_[k]);
            lanes_[k] = nullptr;
        }
        nlanes_ = 1;
        return hip_error(e, "stream lane calibration");
    }
    nlanes_ = found;
    calibrated_ = true;
    return BRKGA_OK;
}

brkga_status Fleet::prepare_lanes(hipStream_t s)
{
    DeviceGuard g(eng_->device);
    return ensure_lanes(s);
}

brkga_status Fleet::set_counters(int gen, hipStream_t s)
{
    DeviceGuard g(eng_->device);
    if (descs_.empty()) return BRKGA_OK;
    hipError_t e = launch_set_counters(d_descs_, (int)descs_.size(), gen, s);
    if (e != hipSuccess) return hip_error(e, "set counters");
    return BRKGA_OK;
}

brkga_status Fleet::ticks(int count, hipStream_t s)
{
    if (count <= 0) return BRKGA_OK;
    DeviceGuard g(eng_->device);
    brkga_status st = ensure_lanes(s);
    if (st != BRKGA_OK) return st;
    if (!ls_descs_.empty()) {
        for (int k = 0; k < count && st == BRKGA_OK; ++k) st = joined_tick(s);
        return st;
    }
    bool used[kLanes] = {};
    for (const auto& L : launches_) used[L.lane] = true;
    if (!groups_.empty()) used[nsga_lane_] = true;
    auto lane = [&](int k) { return k == 0 ? s : lanes_[k]; };
    hipError_t e = hipSuccess;
    bool forked = false;
    for (int k = 1; k < kLanes && e == hipSuccess; ++k)
        if (used[k]) {
            if (!forked) e = hipEventRecord(fork_, s);
            forked = true;
            if (e == hipSuccess) e = hipStreamWaitEvent(lanes_[k], fork_, 0);
        }
    // Each lane launches its smallest-n (latency-bound) launches first, so
    // the kernels still running at the end of its tick are the ones with the
    // most parallel work; then it ranks its members' islands.
    std::vector<int> order(launches_.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return launches_[a].n_max < launches_[b].n_max; });
    for (int k = 0; k < count && e == hipSuccess; ++k) {
        for (size_t r = 0; r < order.size() && e == hipSuccess; ++r) {
            const int j = order[r];
            const FleetLaunch& L = launches_[j];
            e = launch_step_queue(L.kind, L.block, L.xys, L.n_max, d_descs_, d_items_ + L.item_begin, L.item_count,
                                  d_queues_ + kQueueStride * (buckets_.size() + j), L.grid, L.lds, lane(L.lane));
        }
        for (int l = 0; l < kLanes && e == hipSuccess; ++l)
            if (lane_seg_count_[l] > 0)
                e = launch_partition_list(d_descs_, d_segs_ + lane_seg_begin_[l], lane_seg_count_[l], max_Pi_, -1, 1,
                                          lane(l));
        for (size_t i = 0; i < groups_.size() && e == hipSuccess; ++i) {
            const FleetGroup& gr = groups_[i];
            const InstDesc* d = d_descs_ + gr.first;
            hipStream_t ls = lane(nsga_lane_);
            e = launch_nsga2_offspring(gr.kind, gr.block, d, gr.count, gr.rows, gr.lds, 0, ls);
            if (e == hipSuccess) e = launch_nsga2_select(d, gr.count, 0, gr.max_pop, ls);
            if (e == hipSuccess) e = launch_nsga2_gather(d, gr.count, gr.rows, ls);
            if (e == hipSuccess && gr.diversity)
                e = launch_nsga2_diversity(gr.kind, gr.block, d, gr.count, gr.rows, gr.lds, gr.max_pop, ls);
        }
    }
    // every lane that was forked joins s, also after a failed launch, so
    // that waiting for s waits for everything the call queued
    for (int k = 1; k < kLanes && forked; ++k)
        if (used[k]) {
            hipError_t j = hipEventRecord(join_[k], lanes_[k]);
            if (j == hipSuccess) j = hipStreamWaitEvent(s, join_[k], 0);
            if (e == hipSuccess) e = j;
        }
    if (e != hipSuccess) return hip_error(e, "fleet ticks");
    return BRKGA_OK;
}

brkga_status Fleet::joined_tick(hipStream_t s)
{
    bool used[kLanes] = {};
    for (const auto& L : launches_) used[L.lane] = true;
    if (!groups_.empty()) used[nsga_lane_] = true;
    auto lane = [&](int k) { return k == 0 ? s : lanes_[k]; };
    hipError_t e = hipSuccess;
    bool forked = false;
    for (int k = 1; k < kLanes && e == hipSuccess; ++k)
        if (used[k]) {
            if (!forked) e = hipEventRecord(fork_, s);
            forked = true;
            if (e == hipSuccess) e = hipStreamWaitEvent(lanes_[k], fork_, 0);
        }
    // Each lane launches its smallest-n (latency-bound) launches first, so
    // the kernels still running at the end of the tick are the ones with the
    // most parallel work.
    std::vector<int> order(launches_.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return launches_[a].n_max < launches_[b].n_max; });
    for (size_t r = 0; r < order.size() && e == hipSuccess; ++r) {
        const int j = order[r];
        const FleetLaunch& L = launches_[j];
        e = launch_step_queue(L.kind, L.block, L.xys, L.n_max, d_descs_, d_items_ + L.item_begin, L.item_count,
                              d_queues_ + kQueueStride * (buckets_.size() + j), L.grid, L.lds, lane(L.lane));
    }
    for (size_t i = 0; i < groups_.size() && e == hipSuccess; ++i) {
        const FleetGroup& gr = groups_[i];
        const InstDesc* d = d_descs_ + gr.first;
        hipStream_t ls = lane(nsga_lane_);
        e = launch_nsga2_offspring(gr.kind, gr.block, d, gr.count, gr.rows, gr.lds, 0, ls);
        if (e == hipSuccess) e = launch_nsga2_select(d, gr.count, 0, gr.max_pop, ls);
        if (e == hipSuccess) e = launch_nsga2_gather(d, gr.count, gr.rows, ls);
        if (e == hipSuccess && gr.diversity)
            e = launch_nsga2_diversity(gr.kind, gr.block, d, gr.count, gr.rows, gr.lds, gr.max_pop, ls);
    }
    // join the BRKGA lanes before the partition, the NSGA-II lane at the end
    for (int k = 1; k < kLanes && e == hipSuccess; ++k)
        if (used[k] && !(k == nsga_lane_ && !groups_.empty())) {
            e = hipEventRecord(join_[k], lanes_[k]);
            if (e == hipSuccess) e = hipStreamWaitEvent(s, join_[k], 0);
        }
    if (e == hipSuccess && n_brkga_ > 0)
        e = launch_partition_many(d_descs_, n_brkga_, total_islands_, max_Pi_, -1, 1, s);
    if (e == hipSuccess && !ls_descs_.empty()) {
        // members whose new generation is due run their local search, then
        // are sorted again (both kernels skip everyone else)
        e = launch_tspj_local_search(d_ls_descs_, (int)ls_descs_.size(), ls_items_, ls_block_, ls_lds_, s);
        if (e == hipSuccess) e = launch_partition_many(d_descs_, n_brkga_, total_islands_, max_Pi_, -1, 2, s);
    }
    if (e == hipSuccess && !groups_.empty() && nsga_lane_ != 0) {
        e = hipEventRecord(join_[nsga_lane_], lanes_[nsga_lane_]);
        if (e == hipSuccess) e = hipStreamWaitEvent(s, join_[nsga_lane_], 0);
    }
    if (e != hipSuccess) {
        // after a failed launch every forked lane still joins s, so that
        // waiting for s waits for everything the tick queued
        for (int k = 1; k < kLanes && forked; ++k)
            if (used[k] && hipEventRecord(join_[k], lanes_[k]) == hipSuccess)
                (void)hipStreamWaitEvent(s, join_[k], 0);
        return hip_error(e, "fleet tick");
    }
    return BRKGA_OK;
}

brkga_status Fleet::tick_graph(int ticks, hipStream_t s)
{
    if (ticks <= 0) return BRKGA_OK;
    DeviceGuard g(eng_->device);
    auto it = graphs_.find(ticks);
    if (it == graphs_.end()) {
        if (!capture_) BRKGA_HIP_TRY(hipStreamCreateWithFlags(&capture_, hipStreamNonBlocking));
        brkga_status ls = ensure_lanes(s);  // no stream / event creation while capturing
        if (ls != BRKGA_OK) return ls;
        hipGraph_t graph = nullptr;
        BRKGA_HIP_TRY(hipStreamBeginCapture(capture_, hipStreamCaptureModeThreadLocal));
        brkga_status st = this->ticks(ticks, capture_);
        hipError_t e = hipStreamEndCapture(capture_, &graph);
        if (st != BRKGA_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return st;
        }
        if (e != hipSuccess) return hip_error(e, "hipStreamEndCapture");
        hipGraphExec_t exec = nullptr;
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return hip_error(e, "hipGraphInstantiate");
        it = graphs_.emplace(ticks, exec).first;
    }
    BRKGA_HIP_TRY(hipGraphLaunch(it->second, s));
    return BRKGA_OK;
}

// LDS of bucket b's launch with coordinates staged or not.
static int bucket_lds(const std::vector<InstDesc>& descs, const FleetBucket& b, bool xys)
{
    int lds = 0;
    for (int m = b.first; m < b.first + b.count; ++m) {
        const InstDesc& d = descs[m];
        lds = std::max(lds, step_layout(d.prob.kind, d.prob.n, d.L, d.nb, d.Pi, xys, d.prob.aux).bytes);
    }
    return lds;
}

brkga_status Fleet::autotune(int reps, hipStream_t s)
{
    return tune(reps, false, s);
}

brkga_status Fleet::autotune_new(int reps, hipStream_t s)
{
    return tune(reps, true, s);
}

brkga_status Fleet::tune(int reps, bool only_new, hipStream_t s)
{
    DeviceGuard g(eng_->device);
    std::vector<bool> todo(buckets_.size(), true);
    bool any = false;
    {
        std::lock_guard<std::mutex> lock(eng_->tuned_mu);
        for (size_t bi = 0; bi < buckets_.size(); ++bi) {
            if (only_new && eng_->tuned.count({buckets_[bi].kind, buckets_[bi].key})) todo[bi] = false;
            any = any || todo[bi];
        }
    }
    if (!any) return BRKGA_OK;
    if (reps < 1) reps = 1;
    hipEvent_t a = nullptr, b = nullptr;
    BRKGA_HIP_TRY(hipEventCreate(&a));
    hipError_t e = hipEventCreate(&b);
    if (e != hipSuccess) {
        (void)hipEventDestroy(a);
        return hip_error(e, "hipEventCreate");
    }
    brkga_status st = BRKGA_OK;
    for (size_t bi = 0; bi < buckets_.size() && st == BRKGA_OK; ++bi) {
        if (!todo[bi]) continue;
        float best_ms = 0.0f;
        Shape best{buckets_[bi].block, buckets_[bi].chunk, buckets_[bi].xys};
        bool have = false;
        const int stagings = buckets_[bi].kind == KIND_TSP_COORDS ? 2 : 1;
        for (int xi = 0; xi < stagings && st == BRKGA_OK; ++xi) {
            const bool xys = buckets_[bi].kind == KIND_TSP_COORDS && xi == 0;
            const int lds = bucket_lds(descs_, buckets_[bi], xys);
            if ((size_t)lds > eng_->lds_limit) continue;
            for (int blk : kBlocks) {
                if (blk * kSortItems < buckets_[bi].n_max) continue;
                if (step_queue_blocks_per_cu(buckets_[bi].kind, blk, xys, buckets_[bi].n_max, lds) <= 0) continue;
                for (int ch : kChunks) {
                    if (ch > buckets_[bi].max_Pi) continue;
                    buckets_[bi].block = blk;
                    buckets_[bi].chunk = ch;
                    buckets_[bi].xys = xys;
                    buckets_[bi].lds = lds;
                    if ((st = build_items(s)) != BRKGA_OK) break;
                    const FleetBucket& bk = buckets_[bi];
                    int* q = d_queues_ + kQueueStride * bi;
                    e = launch_step_queue(bk.kind, bk.block, bk.xys, bk.n_max, d_descs_, d_items_ + bk.item_begin,
                                          bk.item_count, q, bk.grid, bk.lds, s);
                    // the fastest of kTuneRounds timed rounds of `reps` launches
                    float ms = 0.0f;
                    for (int round = 0; round < kTuneRounds && e == hipSuccess; ++round) {
                        e = hipEventRecord(a, s);
                        for (int r = 0; r < reps && e == hipSuccess; ++r)
                            e = launch_step_queue(bk.kind, bk.block, bk.xys, bk.n_max, d_descs_,
                                                  d_items_ + bk.item_begin, bk.item_count, q, bk.grid, bk.lds, s);
                        if (e == hipSuccess) e = hipEventRecord(b, s);
                        if (e == hipSuccess) e = hipEventSynchronize(b);
                        float r_ms = 0.0f;
                        if (e == hipSuccess) e = hipEventElapsedTime(&r_ms, a, b);
                        if (e == hipSuccess && (round == 0 || r_ms < ms)) ms = r_ms;
                    }
                    if (e != hipSuccess) {
                        st = hip_error(e, "autotune");
                        break;
                    }
                    if (!have || ms < best_ms) {
                        have = true;
                        best_ms = ms;
                        best = Shape{blk, ch, xys, 1000.0 * ms / reps / std::max(1, buckets_[bi].children),
                                     buckets_[bi].children};
                    }
                }
                if (st != BRKGA_OK) break;
            }
        }
        buckets_[bi].block = best.block;
        buckets_[bi].chunk = best.chunk;
        buckets_[bi].xys = best.xys;
        buckets_[bi].lds = bucket_lds(descs_, buckets_[bi], best.xys);
        buckets_[bi].child_us = best.child_us;
        buckets_[bi].tuned_children = best.children;
        std::lock_guard<std::mutex> lock(eng_->tuned_mu);
        eng_->tuned[{buckets_[bi].kind, buckets_[bi].key}] = best;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (st != BRKGA_OK) return st;
    form_launches();
    assign_lanes();
    drop_graphs();
    return build_items(s);
}

}  // namespace brkga
