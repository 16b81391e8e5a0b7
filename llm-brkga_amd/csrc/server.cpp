// server.cpp - request loop with continuous batching (brkga_server_*).
//
// One worker thread owns a Fleet of resident requests.  Each round it
//   1. admits pending requests first-come first-served while each one's
//      population (one contiguous slab of the HBM arena) fits the budget and
//      the resident cap; admitted populations start at step -1, so the next
//      tick builds their generation 0; requests that reached their deadline
//      in the previous round leave the fleet,
//   2. enqueues `ticks_per_round` ticks of every resident request on its
//      stream (members at their deadline are skipped by the kernels),
//   3. while those run, collects the results of the requests that left:
//      best individual to host on a second stream, slab back to the arena,
//   4. waits for the ticks and moves the requests now at their deadline to
//      the leaving set.
// Arena slabs are never returned to the driver while the server runs, and
// problems handed over with a request are destroyed only after the round's
// ticks are done, so no round waits on a device-wide free.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "capi_internal.hpp"
#include "fleet.hpp"
#include "launch.hpp"

using namespace brkga;

namespace {

constexpr size_t kSlabAlign = 256;

// First-fit allocator over one device allocation; free ranges coalesce.
class Arena {
public:
    brkga_status reserve(size_t bytes)
    {
        void* p = nullptr;
        brkga_status st = device_alloc(&p, bytes, "server arena");
        if (st != BRKGA_OK) return st;
        base_ = static_cast<char*>(p);
        size_ = bytes;
        free_[0] = bytes;
        return BRKGA_OK;
    }
    ~Arena()
    {
        if (base_) (void)hipFree(base_);
    }
    char* take(size_t bytes)
    {
        bytes = (bytes + kSlabAlign - 1) & ~(kSlabAlign - 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const size_t off = it->first, len = it->second;
            free_.erase(it);
            if (len > bytes) free_[off + bytes] = len - bytes;
            used_ += bytes;
            peak_ = std::max(peak_, used_);
            live_[off] = bytes;
            return base_ + off;
        }
        return nullptr;
    }
    void give(char* p)
    {
        const size_t off = (size_t)(p - base_);
        auto lv = live_.find(off);
        if (lv == live_.end()) return;
        size_t len = lv->second;
        live_.erase(lv);
        used_ -= len;
        size_t start = off;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && next->first == off + len) {
            len += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                start = prev->first;
                len += prev->second;
                free_.erase(prev);
            }
        }
        free_[start] = len;
    }
    size_t size() const { return size_; }
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }

private:
    char* base_ = nullptr;
    size_t size_ = 0, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;  // offset -> length
    std::unordered_map<size_t, size_t> live_;
};

// Carves one population's buffers out of a single arena slab.
struct SlabAllocator : DeviceAllocator {
    char* base = nullptr;
    size_t size = 0, used = 0;
    brkga_status alloc(void** p, size_t n, const char* what) override
    {
        const size_t r = (n + kSlabAlign - 1) & ~(kSlabAlign - 1);
        if (used + r > size)
            return set_error(BRKGA_ERR_OOM, std::string("server slab exhausted (") + what + ")");
        *p = base + used;
        used += r;
        return BRKGA_OK;
    }
    void release(void*) override {}
};

struct Request {
    uint64_t id = 0;
    brkga_problem_s* problem = nullptr;
    bool own_problem = false;
    brkga_params params{};
    int generations = 0;
    size_t bytes = 0;
    char* slab = nullptr;
    SlabAllocator alloc;
    brkga_pop_s* pop = nullptr;
    int ticks = 0;
    double gen_us = 0.0;  // device time per generation in its last lone round
    brkga_request_info info{};
    std::vector<float> fitness;
    std::vector<int32_t> decoded;
};

double now_us(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct brkga_server_s {
    brkga_engine_s* engine = nullptr;
    brkga_server_config cfg{};
    Arena arena;
    std::unique_ptr<Fleet> fleet;
    hipStream_t stream = nullptr;     // ticks and population set-up
    hipStream_t collect_stream = nullptr;  // results of leaving requests
    hipStream_t solo_streams[4] = {};      // one per request of a solo round
    hipEvent_t setup_done = nullptr;       // population set-up on `stream` before solo steps
    char* staging = nullptr;               // pinned host staging of results
    size_t staging_bytes = 0;
    brkga::PopSetup* setup_host = nullptr;  // pinned: set-ups of a round's admissions
    brkga::PopSetup* setup_dev = nullptr;   // their device copy (launch_setup_pops)
    int round_ticks = 0;                   // ticks of the last fleet round (grows while quiet)
    size_t setup_cap = 0;
    std::chrono::steady_clock::time_point t0;
    std::thread worker;

    std::mutex mu;
    std::condition_variable cv_work;  // worker: new requests or stop
    std::condition_variable cv_done;  // waiters: a request finished
    bool stopping = false;
    uint64_t next_id = 1;
    std::deque<std::unique_ptr<Request>> pending;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> finished;
    std::unordered_map<uint64_t, int> in_flight;  // ids admitted and not finished
    std::unordered_set<uint64_t> cancelled;       // in-flight ids to retire at the next round
    brkga_server_stats stats{};
    std::string worker_error;

    // worker-only state
    std::vector<std::unique_ptr<Request>> resident;
    std::vector<std::unique_ptr<Request>> leaving;  // at their deadline, results not collected yet
    std::vector<brkga_problem_s*> doomed;           // owned problems to destroy after the round
    bool fleet_stale = false;                       // resident set changed outside the ticks

    void run();
    bool solo_round(double* waited);
    bool admit();
    bool retire_cancelled();
    void collect_leaving();
    void finish(std::unique_ptr<Request> r, brkga_status st);
};

// Completes a request: releases its population and slab; an owned problem
// is queued for destruction after the round.
void brkga_server_s::finish(std::unique_ptr<Request> r, brkga_status st)
{
    r->info.status = st;
    r->info.done_us = now_us(t0);
    if (r->pop) {
        destroy_population(r->pop);
        r->pop = nullptr;
    }
    if (r->slab) {
        arena.give(r->slab);
        r->slab = nullptr;
    }
    if (r->own_problem && r->problem) {
        doomed.push_back(r->problem);
        r->problem = nullptr;
    }
    std::lock_guard<std::mutex> lock(mu);
    in_flight.erase(r->id);
    ++stats.completed;
    stats.bytes_in_use = arena.used();
    finished[r->id] = std::move(r);
    cv_done.notify_all();
}

// Retires the resident requests cancelled since the last round: no launch
// that uses their populations is queued any more (every round ends by
// waiting for its steps).  True when the resident set changed.
bool brkga_server_s::retire_cancelled()
{
    std::vector<uint64_t> ids;
    {
        std::lock_guard<std::mutex> lock(mu);
        if (cancelled.empty()) return false;
        ids.assign(cancelled.begin(), cancelled.end());
        cancelled.clear();
    }
    bool changed = false;
    std::vector<std::unique_ptr<Request>> keep;
    for (auto& r : resident) {
        if (std::find(ids.begin(), ids.end(), r->id) != ids.end()) {
            r->info.generations = r->ticks > 0 ? r->ticks - 1 : 0;
            finish(std::move(r), BRKGA_ERR_CANCELLED);
            changed = true;
        } else {
            keep.push_back(std::move(r));
        }
    }
    resident.swap(keep);
    return changed;
}

// Admits pending requests in arrival order while they fit; true when the
// resident set changed.  The new populations' device set-up (cleared key
// buffers, counters, descriptors) is one upload and one launch on `stream`
// for all of them: issued one by one, its six runtime calls per request
// held a burst of 200 requests about 8 ms in admission.
bool brkga_server_s::admit()
{
    bool changed = false;
    std::vector<PopSetup> setups;
    std::vector<Request*> fresh;
    for (;;) {
        std::unique_ptr<Request> r;
        {
            std::lock_guard<std::mutex> lock(mu);
            if (pending.empty()) break;
            if (cfg.max_resident > 0 && (int)resident.size() >= cfg.max_resident) break;
            Request* head = pending.front().get();
            char* slab = arena.take(head->bytes);
            if (!slab) break;  // first-come first-served: the head waits for memory
            r = std::move(pending.front());
            pending.pop_front();
            r->slab = slab;
            in_flight[r->id] = 1;
            ++stats.admitted;
            stats.bytes_in_use = arena.used();
            stats.peak_bytes = arena.peak();
        }
        r->alloc.base = r->slab;
        r->alloc.size = r->bytes;
        r->info.admit_us = now_us(t0);
        PopSetup su;
        brkga_status st = create_population(r->problem, &r->params, &r->alloc, stream, -1, &r->pop, &su);
        if (st != BRKGA_OK) {
            r->pop = nullptr;
            finish(std::move(r), st);
            continue;
        }
        r->pop->host.gen_limit = r->generations;
        su.desc.gen_limit = r->generations;
        setups.push_back(su);
        fresh.push_back(r.get());
        resident.push_back(std::move(r));
        changed = true;
    }
    if (setups.empty()) return changed;
    brkga_status st = BRKGA_OK;
    hipError_t e = hipSuccess;
    if (setups.size() > setup_cap) {
        // the previous lists are idle: every round ends by waiting for `stream`
        if (setup_host) (void)hipHostFree(setup_host);
        if (setup_dev) (void)hipFree(setup_dev);
        setup_host = nullptr;
        setup_dev = nullptr;
        setup_cap = 0;
        const size_t cap = std::max<size_t>(setups.size(), 256);
        e = hipHostMalloc(reinterpret_cast<void**>(&setup_host), cap * sizeof(PopSetup), hipHostMallocDefault);
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&setup_dev), cap * sizeof(PopSetup));
        if (e == hipSuccess) setup_cap = cap;
        else st = hip_error(e, "population set-up lists");
    }
    if (st == BRKGA_OK) {
        std::copy(setups.begin(), setups.end(), setup_host);
        e = hipMemcpyAsync(setup_dev, setup_host, setups.size() * sizeof(PopSetup), hipMemcpyHostToDevice, stream);
        if (e == hipSuccess) e = launch_setup_pops(setup_dev, (int)setups.size(), stream);
        if (e != hipSuccess) st = hip_error(e, "population set-up");
    }
    if (st != BRKGA_OK) {
        // the new requests end with the error; their populations were never set up
        (void)hipStreamSynchronize(stream);
        std::vector<std::unique_ptr<Request>> keep, failed;
        for (auto& r : resident) {
            if (std::find(fresh.begin(), fresh.end(), r.get()) != fresh.end()) failed.push_back(std::move(r));
            else keep.push_back(std::move(r));
        }
        resident.swap(keep);
        for (auto& r : failed) finish(std::move(r), st);
    }
    return changed;
}

// Results of the requests that left the fleet this round: their buffers
// are final (the previous round's ticks were waited for), so they are read
// on the collect stream while the current ticks run.  All copies go through
// one pinned staging area: two synchronisations per round, whatever the
// number of requests.
void brkga_server_s::collect_leaving()
{
    if (leaving.empty()) return;
    // layout per request: fitness [P][M], order or rank [P], crowding [P]
    std::vector<size_t> off(leaving.size() + 1, 0);
    for (size_t i = 0; i < leaving.size(); ++i) {
        const InstDesc& d = leaving[i]->pop->host;
        off[i + 1] = off[i] + (size_t)d.pop * (d.prob.num_obj + 2) * 4 + (size_t)d.L * 4 + 16;
        off[i + 1] = (off[i + 1] + 15) & ~(size_t)15;
    }
    brkga_status st = BRKGA_OK;
    if (off.back() > staging_bytes) {
        if (staging) (void)hipHostFree(staging);
        staging = nullptr;
        staging_bytes = 0;
        hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&staging), off.back(), hipHostMallocDefault);
        if (e != hipSuccess) st = hip_error(e, "hipHostMalloc(result staging)");
        else staging_bytes = off.back();
    }
    // a request that reached its deadline holds generation `generations`
    hipError_t e = hipSuccess;
    for (size_t i = 0; i < leaving.size() && st == BRKGA_OK && e == hipSuccess; ++i) {
        const brkga_pop_s* p = leaving[i]->pop;
        const InstDesc& d = p->host;
        const int b = leaving[i]->generations & 1;
        const size_t P = (size_t)d.pop, M = (size_t)d.prob.num_obj;
        char* base = staging + off[i];
        e = hipMemcpyAsync(base, d.fit[b], P * M * 4, hipMemcpyDeviceToHost, collect_stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(base + P * M * 4, p->nsga2 ? (const void*)d.rank[b] : (const void*)d.perm[b], P * 4,
                               hipMemcpyDeviceToHost, collect_stream);
        if (e == hipSuccess && p->nsga2)
            e = hipMemcpyAsync(base + P * (M + 1) * 4, d.crowd[b], P * 4, hipMemcpyDeviceToHost, collect_stream);
    }
    if (e == hipSuccess && st == BRKGA_OK) e = hipStreamSynchronize(collect_stream);
    std::vector<int> best(leaving.size(), -1);
    for (size_t i = 0; i < leaving.size() && e == hipSuccess && st == BRKGA_OK; ++i) {
        const brkga_pop_s* p = leaving[i]->pop;
        const InstDesc& d = p->host;
        const int b = leaving[i]->generations & 1;
        const size_t P = (size_t)d.pop, M = (size_t)d.prob.num_obj;
        char* base = staging + off[i];
        const float* fit = reinterpret_cast<const float*>(base);
        best[i] = pick_best(d, p->nsga2, fit, reinterpret_cast<const int*>(base + P * M * 4),
                            reinterpret_cast<const float*>(base + P * (M + 1) * 4));
        if (best[i] < 0) continue;
        leaving[i]->fitness.assign(fit + (size_t)best[i] * M, fit + (size_t)best[i] * M + M);
        // decode the best row into the population's scratch row, then stage it
        e = launch_evaluate(d.prob, d.keys[b] + (size_t)best[i] * d.ld, d.ld, 1, p->d_fit_row, p->d_decode_row,
                            p->block, collect_stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(base + P * (M + 2) * 4, p->d_decode_row, (size_t)d.L * 4, hipMemcpyDeviceToHost,
                               collect_stream);
    }
    if (e == hipSuccess && st == BRKGA_OK) e = hipStreamSynchronize(collect_stream);
    if (e != hipSuccess && st == BRKGA_OK) st = hip_error(e, "collecting results");
    for (size_t i = 0; i < leaving.size(); ++i) {
        auto& r = leaving[i];
        brkga_status rs = st;
        if (rs == BRKGA_OK && best[i] < 0) rs = set_error(BRKGA_ERR_DEVICE, "no best individual found");
        if (rs == BRKGA_OK) {
            const InstDesc& d = r->pop->host;
            const int32_t* dec =
                reinterpret_cast<const int32_t*>(staging + off[i] + (size_t)d.pop * (d.prob.num_obj + 2) * 4);
            r->decoded.assign(dec, dec + d.L);
        }
        r->info.generations = r->generations;
        r->info.device_bytes = r->bytes;
        finish(std::move(r), rs);
    }
    leaving.clear();
}

// Generations a lone request runs per round through the single-population
// launches (bounds how long a request arriving meanwhile waits to be
// admitted: 64 generations of the largest requests take about a
// millisecond).
constexpr int kSoloChunk = 64;
// A request alone on the device runs longer rounds, about kLoneRoundUs of
// device time each (sized from its previous round; the first runs
// kLoneFirst generations): few synchronisations for a short request, and a
// request arriving meanwhile waits at most about that long to be admitted.
constexpr int kLoneFirst = 8;
constexpr int kLoneMax = 256;
constexpr double kLoneRoundUs = 1500.0;
// Most requests a solo round advances side by side, each on a stream of its
// own; more resident requests share the batched ticks.
constexpr int kSoloMax = 4;
// A burst of submissions is admitted together: while at least two requests
// are pending and the newest arrived less than kBurstGapUs ago, the worker
// waits for the next one (at most kBurstMaxUs in all) before a round, so
// the burst's requests start their generations side by side rather than a
// few rounds apart (the last of them then finishes that many rounds
// sooner).  A lone pending request is admitted at once.
constexpr double kBurstGapUs = 200.0;
constexpr double kBurstMaxUs = 20000.0;
// Rounds with nothing pending and nothing admitted grow (doubling from
// ticks_per_round up to kQuietRoundMax ticks, and not past the earliest
// resident deadline beyond ticks_per_round): each round joins the stream
// lanes and waits for the device once, which costs several percent at 4
// ticks; a request arriving meanwhile waits for the round in progress.
constexpr int kQuietRoundMax = 16;

// One round for a few requests (at most kSoloMax) that have the device to
// themselves with nothing pending: a chunk of each one's generations
// through the single-population launches (one step launch per generation),
// each on a stream of its own so they overlap on the device.  A few
// requests finish sooner this way than through the batched ticks, whose
// per-tick fork, partition and join cost more than a few members' steps.
// Finished requests are collected in the same round.  A population's device
// state after these steps is what the ticks would have left (solo
// equivalence), so requests return to the ticks when more join.  False
// when the round is not such a round.
bool brkga_server_s::solo_round(double* waited_out)
{
    if (resident.empty() || (int)resident.size() > kSoloMax || !leaving.empty()) return false;
    {
        std::lock_guard<std::mutex> lock(mu);
        if (!pending.empty()) return false;
    }
    brkga_status st = BRKGA_OK;
    // populations admitted this round were set up on `stream`
    hipError_t e = hipEventRecord(setup_done, stream);
    std::vector<int> advanced(resident.size(), 0);
    for (size_t i = 0; i < resident.size() && st == BRKGA_OK; ++i) {
        Request* r = resident[i].get();
        hipStream_t ss = solo_streams[i];
        if (e == hipSuccess) e = hipStreamWaitEvent(ss, setup_done, 0);
        if (e != hipSuccess) {
            st = hip_error(e, "server solo streams");
            break;
        }
        const int init = r->ticks == 0;
        // tick 1 builds generation 0, tick g + 1 completes step g
        if (!init) r->pop->host_t = r->ticks - 1;
        const int done = init ? 0 : r->ticks - 1;
        int chunk = kSoloChunk;
        if (resident.size() == 1)
            chunk = r->gen_us > 0.0 ? (int)std::min<double>(kLoneMax, std::max(16.0, kLoneRoundUs / r->gen_us))
                                    : kLoneFirst;
        const int steps = std::min(chunk, r->generations - done);
        st = population_steps(r->pop, init, steps, ss);
        advanced[i] = steps + init;
    }
    const double h0 = now_us(t0);
    for (size_t i = 0; i < resident.size(); ++i) {
        const hipError_t se = hipStreamSynchronize(solo_streams[i]);
        if (se != hipSuccess && st == BRKGA_OK) st = hip_error(se, "server solo steps");
    }
    *waited_out = now_us(t0) - h0;
    if (st == BRKGA_OK && resident.size() == 1 && advanced[0] > 0)
        resident[0]->gen_us = *waited_out / advanced[0];
    if (st == BRKGA_OK) {
        std::vector<std::unique_ptr<Request>> keep;
        int most = 0;
        for (size_t i = 0; i < resident.size(); ++i) {
            auto& r = resident[i];
            r->ticks += advanced[i];
            most = std::max(most, advanced[i]);
            if (r->ticks >= r->generations + 1) leaving.push_back(std::move(r));
            else keep.push_back(std::move(r));
        }
        resident.swap(keep);
        {
            std::lock_guard<std::mutex> lock(mu);
            stats.ticks += most;
        }
        collect_leaving();
        return true;
    }
    {
        std::lock_guard<std::mutex> lock(mu);
        worker_error = brkga_last_error();
    }
    std::vector<std::unique_ptr<Request>> rs;
    rs.swap(resident);
    for (auto& r : rs) finish(std::move(r), st);
    return true;
}

void brkga_server_s::run()
{
    DeviceGuard g(engine->device);
    for (;;) {
        {
            std::unique_lock<std::mutex> lock(mu);
            cv_work.wait(lock, [&] {
                return stopping || !pending.empty() || !resident.empty() || !leaving.empty() || !cancelled.empty();
            });
            if (stopping && pending.empty() && resident.empty() && leaving.empty()) break;
            if (pending.empty() && resident.empty() && leaving.empty()) {
                cancelled.clear();  // ids that finished before their cancel was seen
                continue;
            }
            const double b0 = now_us(t0);
            while (!stopping && pending.size() >= 2) {
                const double now = now_us(t0);
                const double gap = now - pending.back()->info.submit_us;
                if (gap >= kBurstGapUs || now - b0 >= kBurstMaxUs) break;
                cv_work.wait_for(lock, std::chrono::microseconds((long long)(kBurstGapUs - gap) + 1));
            }
            ++stats.rounds;
        }
        const double h0 = now_us(t0);
        double waited = 0.0;
        brkga_status st = BRKGA_OK;
        if (retire_cancelled()) fleet_stale = true;
        const bool admitted = admit();
        if (solo_round(&waited)) {
            // the fleet's member list is rebuilt when the resident set
            // next changes under the ticks
            fleet_stale = true;
            for (auto* p : doomed) (void)brkga_problem_destroy(p);
            doomed.clear();
            std::lock_guard<std::mutex> lock(mu);
            stats.wait_us += waited;
            stats.host_us += now_us(t0) - h0 - waited;
            continue;
        }
        if (admitted || !leaving.empty() || fleet_stale) {
            fleet_stale = false;
            std::vector<brkga_pop_s*> pops;
            for (auto& r : resident) pops.push_back(r->pop);
            st = fleet->set_members(pops, stream);
            // step launches alone leave every population as it is
            if (st == BRKGA_OK && admitted && cfg.autotune_reps > 0)
                st = fleet->autotune_new(cfg.autotune_reps, stream);
        }
        const bool ticking = st == BRKGA_OK && !resident.empty();
        bool quiet = !admitted;
        {
            std::lock_guard<std::mutex> lock(mu);
            quiet = quiet && pending.empty();
        }
        const int base = cfg.ticks_per_round;
        int ticks = base;
        if (quiet) {
            round_ticks = std::min(std::max(2 * round_ticks, base), std::max(base, kQuietRoundMax));
            int rem = round_ticks;
            for (const auto& r : resident) rem = std::min(rem, r->generations + 1 - r->ticks);
            ticks = std::max(base, rem);
        } else {
            round_ticks = base;
        }
        if (ticking) st = fleet->ticks(ticks, stream);
        collect_leaving();
        // problems of the requests that left are referenced by no queued
        // launch; hipFree waits for the device, i.e. for this round's ticks
        for (auto* p : doomed) (void)brkga_problem_destroy(p);
        doomed.clear();
        if (ticking && st == BRKGA_OK) {
            const double w0 = now_us(t0);
            hipError_t e = hipStreamSynchronize(stream);
            waited = now_us(t0) - w0;
            if (e != hipSuccess) st = hip_error(e, "server tick");
        }
        if (ticking && st == BRKGA_OK) {
            {
                std::lock_guard<std::mutex> lock(mu);
                stats.ticks += ticks;
                stats.lanes = fleet->lanes();
            }
            // tick 1 builds generation 0, tick g + 1 completes step g
            std::vector<std::unique_ptr<Request>> keep;
            for (auto& r : resident) {
                r->ticks += ticks;
                if (r->ticks >= r->generations + 1) leaving.push_back(std::move(r));
                else keep.push_back(std::move(r));
            }
            resident.swap(keep);
        }
        for (auto* p : doomed) (void)brkga_problem_destroy(p);
        doomed.clear();
        {
            std::lock_guard<std::mutex> lock(mu);
            stats.wait_us += waited;
            stats.host_us += now_us(t0) - h0 - waited;
        }
        if (st != BRKGA_OK) {
            // a device failure ends every resident request with that status
            {
                std::lock_guard<std::mutex> lock(mu);
                worker_error = brkga_last_error();
            }
            (void)hipStreamSynchronize(stream);
            std::vector<std::unique_ptr<Request>> rs;
            rs.swap(resident);
            for (auto& r : leaving) rs.push_back(std::move(r));
            leaving.clear();
            for (auto& r : rs) finish(std::move(r), st);
            (void)fleet->set_members({}, stream);
            for (auto* p : doomed) (void)brkga_problem_destroy(p);
            doomed.clear();
        }
    }
}

extern "C" {

brkga_status brkga_server_start(brkga_engine_t engine, const brkga_server_config* config, brkga_server_t* out)
{
    if (!engine || !out) return set_error(BRKGA_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    DeviceGuard g(engine->device);
    std::unique_ptr<brkga_server_s> s(new brkga_server_s());
    s->engine = engine;
    if (config) s->cfg = *config;
    if (s->cfg.hbm_budget_bytes == 0) s->cfg.hbm_budget_bytes = (uint64_t)8 << 30;
    if (s->cfg.ticks_per_round <= 0) s->cfg.ticks_per_round = 4;
    if (s->cfg.max_resident < 0) return set_error(BRKGA_ERR_INVALID_ARG, "max_resident must be >= 0");
    brkga_status st = s->arena.reserve(s->cfg.hbm_budget_bytes);
    if (st != BRKGA_OK) return st;
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->collect_stream, hipStreamNonBlocking);
    for (int i = 0; i < kSoloMax && e == hipSuccess; ++i)
        e = hipStreamCreateWithFlags(&s->solo_streams[i], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->setup_done, hipEventDisableTiming);
    if (e != hipSuccess) return hip_error(e, "hipStreamCreate");
    s->fleet.reset(new Fleet(engine));
    // the lanes are calibrated here, not when the first requests arrive
    // (the server's own streams exist first, as when the ticks calibrate)
    (void)s->fleet->prepare_lanes(s->stream);
    s->stats.budget_bytes = s->cfg.hbm_budget_bytes;
    s->t0 = std::chrono::steady_clock::now();
    brkga_server_s* raw = s.get();
    s->worker = std::thread([raw] { raw->run(); });
    *out = s.release();
    return BRKGA_OK;
}

brkga_status brkga_server_submit(brkga_server_t server, brkga_problem_t problem, const brkga_params* params,
                                 int32_t generations, int32_t own_problem, uint64_t* request_id)
{
    if (!server || !problem || !params || !request_id)
        return set_error(BRKGA_ERR_INVALID_ARG, "null argument");
    if (problem->engine != server->engine) return set_error(BRKGA_ERR_INVALID_ARG, "problem of another engine");
    if (generations < 0) return set_error(BRKGA_ERR_INVALID_ARG, "negative generation count");
    if (problem->desc.kind == KIND_HOST)
        return set_error(BRKGA_ERR_INVALID_ARG, "host-evaluated problems run through brkga_population_host_*");
    std::unique_ptr<Request> r(new Request());
    r->problem = problem;
    r->own_problem = own_problem != 0;
    r->params = *params;
    r->generations = generations;
    brkga_status st = population_bytes(problem, params, &r->bytes);
    std::lock_guard<std::mutex> lock(server->mu);
    if (server->stopping) return set_error(BRKGA_ERR_INVALID_ARG, "server is stopping");
    if (st != BRKGA_OK) {
        ++server->stats.rejected;
        return st;
    }
    if (r->bytes > server->arena.size()) {
        ++server->stats.rejected;
        return set_error(BRKGA_ERR_OOM, "request needs " + std::to_string(r->bytes) +
                                            " bytes, more than the server budget");
    }
    r->id = server->next_id++;
    r->info.submit_us = now_us(server->t0);
    *request_id = r->id;
    ++server->stats.submitted;
    server->pending.push_back(std::move(r));
    server->cv_work.notify_one();
    return BRKGA_OK;
}

brkga_status brkga_server_wait(brkga_server_t server, uint64_t request_id, int32_t timeout_ms, float* fitness,
                               int32_t* decoded, brkga_request_info* info)
{
    if (!server) return set_error(BRKGA_ERR_INVALID_ARG, "null server");
    std::unique_lock<std::mutex> lock(server->mu);
    auto known = [&] {
        if (server->finished.count(request_id) || server->in_flight.count(request_id)) return true;
        for (auto& p : server->pending)
            if (p->id == request_id) return true;
        return false;
    };
    if (!known()) return set_error(BRKGA_ERR_NOT_FOUND, "unknown or already collected request");
    auto ready = [&] { return server->finished.count(request_id) > 0; };
    if (!ready()) {
        if (timeout_ms == 0) return BRKGA_ERR_BUSY;
        if (timeout_ms < 0) {
            server->cv_done.wait(lock, ready);
        } else if (!server->cv_done.wait_for(lock, std::chrono::milliseconds(timeout_ms), ready)) {
            return BRKGA_ERR_BUSY;
        }
    }
    std::unique_ptr<Request> r = std::move(server->finished[request_id]);
    server->finished.erase(request_id);
    if (info) *info = r->info;
    if (r->info.status == BRKGA_ERR_CANCELLED) return set_error(BRKGA_ERR_CANCELLED, "request cancelled");
    if (r->info.status != BRKGA_OK)
        return set_error((brkga_status)r->info.status, "request failed: " + server->worker_error);
    if (fitness) std::copy(r->fitness.begin(), r->fitness.end(), fitness);
    if (decoded) std::copy(r->decoded.begin(), r->decoded.end(), decoded);
    return BRKGA_OK;
}

brkga_status brkga_server_cancel(brkga_server_t server, uint64_t request_id)
{
    if (!server) return set_error(BRKGA_ERR_INVALID_ARG, "null server");
    std::unique_ptr<Request> dropped;
    {
        std::lock_guard<std::mutex> lock(server->mu);
        if (server->finished.count(request_id)) return BRKGA_OK;
        if (server->in_flight.count(request_id)) {
            server->cancelled.insert(request_id);
            server->cv_work.notify_one();
            return BRKGA_OK;
        }
        for (auto it = server->pending.begin(); it != server->pending.end(); ++it)
            if ((*it)->id == request_id) {
                dropped = std::move(*it);
                server->pending.erase(it);
                break;
            }
        if (!dropped) return set_error(BRKGA_ERR_NOT_FOUND, "unknown or already collected request");
        dropped->info.status = BRKGA_ERR_CANCELLED;
        dropped->info.done_us = now_us(server->t0);
        ++server->stats.completed;
    }
    // never admitted: no device memory of the request, and no launch uses
    // its problem
    if (dropped->own_problem && dropped->problem) {
        (void)brkga_problem_destroy(dropped->problem);
        dropped->problem = nullptr;
    }
    std::lock_guard<std::mutex> lock(server->mu);
    server->finished[request_id] = std::move(dropped);
    server->cv_done.notify_all();
    return BRKGA_OK;
}

brkga_status brkga_server_stats_get(brkga_server_t server, brkga_server_stats* stats)
{
    if (!server || !stats) return set_error(BRKGA_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lock(server->mu);
    *stats = server->stats;
    stats->pending = (int32_t)server->pending.size();
    stats->resident = (int32_t)server->in_flight.size();
    return BRKGA_OK;
}

brkga_status brkga_server_stop(brkga_server_t server)
{
    if (!server) return BRKGA_OK;
    {
        std::lock_guard<std::mutex> lock(server->mu);
        server->stopping = true;
        server->cv_work.notify_all();
    }
    if (server->worker.joinable()) server->worker.join();
    {
        DeviceGuard g(server->engine->device);
        (void)hipStreamSynchronize(server->stream);
        server->fleet.reset();
        (void)hipStreamDestroy(server->stream);
        (void)hipStreamDestroy(server->collect_stream);
        for (hipStream_t& ss : server->solo_streams)
            if (ss) (void)hipStreamDestroy(ss);
        if (server->setup_done) (void)hipEventDestroy(server->setup_done);
        if (server->staging) (void)hipHostFree(server->staging);
        if (server->setup_host) (void)hipHostFree(server->setup_host);
        if (server->setup_dev) (void)hipFree(server->setup_dev);
    }
    delete server;
    return BRKGA_OK;
}

}  // extern "C"
