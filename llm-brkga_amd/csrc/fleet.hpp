// fleet.hpp - a set of resident populations advanced together, one
// generation of every member per tick: the engine under the batched and the
// served forms.
//
// BRKGA members are grouped into buckets by (problem kind, n rounded up to
// a power of two); each bucket has its own workgroup size and work-item
// size (slots per item), which the autotuner picks by timing.  Buckets that
// end up with the same queue-step kernel share one persistent launch per
// tick (their work items side by side, largest n first).  One partition launch covers every
// BRKGA member; NSGA-II members run the offspring / selection / gather
// launches per problem kind.  Members may sit at different generations:
// each kernel reads the member's step counter (-1 = build generation 0) and
// leaves members at their generation limit untouched, so a member's result
// does not depend on which other members share its ticks.
#pragma once

#include <map>
#include <utility>
#include <vector>

#include "capi_internal.hpp"

namespace brkga {

struct FleetBucket {
    int kind;
    int key;           // smallest power of two >= n (at least 64) shared by the members
    int first, count;  // members [first, first + count) in the descriptor array
    int n_min, n_max, max_Pi;
    int lds;           // largest step LDS footprint among the members
    int block, chunk;  // launch shape; results do not depend on it
    bool xys;          // coordinates staged in LDS (coordinate kind only)
    int item_begin, item_count, grid;
    double work;       // lane-balancing cost (see assign_lanes)
    int children;      // child slots of the members (sum of populations)
    double child_us;   // measured step time per child slot, 0 if not tuned
    int tuned_children;  // child slots that measurement ran over
    int lane;          // stream lane of the bucket's launch (0 = caller's stream)
};

// One queue-step launch per tick: the buckets whose shapes select the same
// kernel instantiation, their work items contiguous.
struct FleetLaunch {
    int kind, block;
    bool xys;
    int n_max, lds;
    int item_begin, item_count, grid;
    double cost;  // sum of the buckets' lane-balancing costs
    int lane;
    std::vector<int> buckets;  // indices into the bucket list, largest n first
};

struct FleetGroup {  // NSGA-II members of one problem kind
    int kind, first, count, rows, block, lds, max_pop;
    bool diversity;  // some member runs the diversity injection
};

// Streams a tick spreads its independent launches over (the caller's stream
// and kLanes - 1 forked ones).
constexpr int kLanes = 4;

class Fleet {
public:
    explicit Fleet(brkga_engine_s* engine) : eng_(engine) {}
    ~Fleet();
    Fleet(const Fleet&) = delete;
    Fleet& operator=(const Fleet&) = delete;

    // Replaces the member list (populations are borrowed).  Uploads the
    // descriptors and work items on stream s; the host copies are reused
    // only by the next call, which the caller orders after s has drained.
    brkga_status set_members(const std::vector<brkga_pop_s*>& pops, hipStream_t s);
    // Sets every member's step counter to `gen` (-1: the next tick builds
    // generation 0).
    brkga_status set_counters(int gen, hipStream_t s);
    // One generation of every member below its limit.
    brkga_status tick(hipStream_t s) { return ticks(1, s); }
    // `count` generations of every member below its limit.  Each stream lane
    // steps and ranks its own members tick after tick; the lanes join s once,
    // at the end (every tick joins them when a member runs the TSPJ local
    // search, which ranks every member again after it).
    brkga_status ticks(int count, hipStream_t s);
    // `ticks` ticks replayed from a captured graph (cached per count).
    brkga_status tick_graph(int ticks, hipStream_t s);
    // Times every admissible (workgroup size, slots per item) pair of each
    // bucket on the members' current state and keeps the fastest.  Runs
    // step launches only, which rewrite the same next-generation buffers
    // from unchanged inputs, so the populations are not advanced.
    brkga_status autotune(int reps, hipStream_t s);
    // Same, for the buckets whose (kind, size class) the engine has no
    // tuned shape for yet.
    brkga_status autotune_new(int reps, hipStream_t s);

    // Calibrates the stream lanes now rather than in the first ticks (about
    // 10 ms of probe launches): a server does it when it starts, before any
    // request waits for it.  A failure leaves the calibration to the ticks.
    brkga_status prepare_lanes(hipStream_t s);

    const std::vector<FleetBucket>& buckets() const { return buckets_; }
    // Streams the ticks spread their launches over (1 until calibrated).
    int lanes() const { return nlanes_; }
    const std::vector<brkga_pop_s*>& members() const { return members_; }
    size_t size() const { return members_.size(); }

private:
    brkga_status build_items(hipStream_t s);
    // One tick whose lanes join before one partition of every member.
    brkga_status joined_tick(hipStream_t s);
    void form_launches();
    brkga_status tune(int reps, bool only_new, hipStream_t s);
    brkga_status reserve(void** p, size_t* cap, size_t bytes, const char* what);
    // Creates the stream lanes once, keeping only streams that run
    // concurrently with s and with each other (medians of timed probes),
    // or BRKGA_FLEET_LANES of them when that variable is 1 .. kLanes.  A
    // failed calibration leaves one lane and runs again on the next call.
    brkga_status ensure_lanes(hipStream_t s);
    void assign_lanes();
    void drop_graphs();

    brkga_engine_s* eng_;
    std::vector<brkga_pop_s*> members_;  // BRKGA members (bucket order), then NSGA-II groups
    std::vector<InstDesc> descs_;
    std::vector<int4> items_;  // host copy; kept until the next upload has been ordered after this one
    std::vector<FleetBucket> buckets_;
    std::vector<FleetLaunch> launches_;
    std::vector<FleetGroup> groups_;
    int n_brkga_ = 0, total_islands_ = 0, max_Pi_ = 0;
    using Shape = brkga_engine_s::Shape;

    InstDesc* d_descs_ = nullptr;
    size_t descs_cap_ = 0;
    // members with the TSPJ local search (descriptor copies)
    std::vector<InstDesc> ls_descs_;
    InstDesc* d_ls_descs_ = nullptr;
    size_t ls_descs_cap_ = 0;
    int ls_items_ = 0, ls_block_ = 0, ls_lds_ = 0;
    int4* d_items_ = nullptr;
    size_t items_cap_ = 0;
    // {member, island} of every BRKGA island, grouped by the lane that steps
    // it (lane l: [lane_seg_begin_[l], + lane_seg_count_[l]))
    std::vector<int2> segs_;
    int2* d_segs_ = nullptr;
    size_t segs_cap_ = 0;
    int lane_seg_begin_[kLanes] = {}, lane_seg_count_[kLanes] = {};
    int* d_queues_ = nullptr;  // two counters per bucket (tuning launches), then per launch; 128 bytes apart
    size_t queues_cap_ = 0;
    hipStream_t capture_ = nullptr;
    std::map<int, hipGraphExec_t> graphs_;
    hipStream_t lanes_[kLanes] = {};  // [0] unused: lane 0 is the caller's stream
    int nlanes_ = 1;                  // lanes in use: the caller's stream and lanes_[1, nlanes_)
    bool calibrated_ = false;
    hipEvent_t fork_ = nullptr;
    hipEvent_t join_[kLanes] = {};
    int nsga_lane_ = 0;
};

// Power-of-two size class of n used to bucket instances (at least 64).
int fleet_key(int n);

}  // namespace brkga
