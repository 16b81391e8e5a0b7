// batch.cpp - C ABI of the batched form (brkga_batch_*).
#include <memory>
#include <vector>

#include "capi_internal.hpp"
#include "fleet.hpp"

using namespace brkga;

struct brkga_batch_s {
    brkga_engine_s* engine;
    std::vector<brkga_pop_s*> pops;  // in creation order, owned
    std::unique_ptr<Fleet> fleet;
};

namespace {

hipStream_t as_stream(void* s)
{
    return reinterpret_cast<hipStream_t>(s);
}

bool valid_mode(int mode)
{
    return mode == BRKGA_BATCH_PLAIN || mode == BRKGA_BATCH_LAUNCH || mode == BRKGA_BATCH_GRAPH;
}

}  // namespace

extern "C" {

brkga_status brkga_batch_create(brkga_engine_t engine, int32_t count, const brkga_problem_t* problems,
                                const brkga_params* params, brkga_batch_t* out)
{
    if (!engine || !problems || !params || !out || count <= 0)
        return set_error(BRKGA_ERR_INVALID_ARG, "batch needs an engine and count > 0 problems / params");
    *out = nullptr;
    DeviceGuard g(engine->device);
    auto* b = new brkga_batch_s();
    b->engine = engine;
    brkga_status st = BRKGA_OK;
    for (int i = 0; i < count && st == BRKGA_OK; ++i) {
        if (!problems[i] || problems[i]->engine != engine) {
            st = set_error(BRKGA_ERR_INVALID_ARG, "batch member " + std::to_string(i) + ": problem of another engine");
            break;
        }
        if (problems[i]->desc.kind == KIND_HOST) {
            st = set_error(BRKGA_ERR_INVALID_ARG, "batch member " + std::to_string(i) +
                                                      ": host-evaluated problems run through brkga_population_host_*");
            break;
        }
        brkga_pop_s* p = nullptr;
        st = create_population(problems[i], &params[i], nullptr, nullptr, 0, &p);
        if (st == BRKGA_OK) b->pops.push_back(p);
    }
    if (st == BRKGA_OK) {
        hipError_t e = hipStreamSynchronize(nullptr);
        if (e != hipSuccess) st = hip_error(e, "hipStreamSynchronize");
    }
    if (st == BRKGA_OK) {
        b->fleet.reset(new Fleet(engine));
        st = b->fleet->set_members(b->pops, nullptr);
    }
    if (st != BRKGA_OK) {
        (void)hipDeviceSynchronize();
        b->fleet.reset();
        for (auto* p : b->pops) destroy_population(p);
        delete b;
        return st;
    }
    *out = b;
    return BRKGA_OK;
}

brkga_status brkga_batch_destroy(brkga_batch_t batch)
{
    if (!batch) return BRKGA_OK;
    DeviceGuard g(batch->engine->device);
    (void)hipDeviceSynchronize();
    batch->fleet.reset();
    for (auto* p : batch->pops) destroy_population(p);
    delete batch;
    return BRKGA_OK;
}

brkga_status brkga_batch_initialize(brkga_batch_t batch, int32_t mode, void* stream)
{
    if (!batch || !valid_mode(mode)) return set_error(BRKGA_ERR_INVALID_ARG, "null batch or unknown mode");
    DeviceGuard g(batch->engine->device);
    hipStream_t s = as_stream(stream);
    if (mode == BRKGA_BATCH_PLAIN) {
        for (auto* p : batch->pops) {
            brkga_status st = brkga_population_initialize(p, stream);
            if (st != BRKGA_OK) return st;
        }
        return BRKGA_OK;
    }
    // every member at -1; one tick builds generation 0 everywhere
    brkga_status st = batch->fleet->set_counters(-1, s);
    if (st == BRKGA_OK) st = batch->fleet->tick(s);
    if (st != BRKGA_OK) return st;
    for (auto* p : batch->pops) p->host_t = 0;
    return BRKGA_OK;
}

brkga_status brkga_batch_evolve(brkga_batch_t batch, int32_t generations, int32_t mode, void* stream)
{
    if (!batch || !valid_mode(mode)) return set_error(BRKGA_ERR_INVALID_ARG, "null batch or unknown mode");
    if (generations < 0) return set_error(BRKGA_ERR_INVALID_ARG, "negative generation count");
    if (generations == 0) return BRKGA_OK;
    DeviceGuard g(batch->engine->device);
    hipStream_t s = as_stream(stream);
    brkga_status st = BRKGA_OK;
    if (mode == BRKGA_BATCH_PLAIN) {
        for (auto* p : batch->pops)
            if ((st = brkga_evolve(p, generations, stream)) != BRKGA_OK) return st;
        return BRKGA_OK;
    }
    if (mode == BRKGA_BATCH_LAUNCH) {
        st = batch->fleet->ticks(generations, s);
    } else {
        st = batch->fleet->tick_graph(generations, s);
    }
    if (st != BRKGA_OK) return st;
    for (auto* p : batch->pops) p->host_t += generations;
    return BRKGA_OK;
}

brkga_status brkga_batch_autotune(brkga_batch_t batch, int32_t reps, void* stream)
{
    if (!batch) return set_error(BRKGA_ERR_INVALID_ARG, "null batch");
    return batch->fleet->autotune(reps, as_stream(stream));
}

brkga_status brkga_batch_population(brkga_batch_t batch, int32_t index, brkga_pop_t* out)
{
    if (!batch || !out || index < 0 || index >= (int)batch->pops.size())
        return set_error(BRKGA_ERR_INVALID_ARG, "null batch or member index out of range");
    *out = batch->pops[index];
    return BRKGA_OK;
}

brkga_status brkga_batch_num_buckets(brkga_batch_t batch, int32_t* count)
{
    if (!batch || !count) return set_error(BRKGA_ERR_INVALID_ARG, "null argument");
    *count = (int32_t)batch->fleet->buckets().size();
    return BRKGA_OK;
}

brkga_status brkga_batch_lanes(brkga_batch_t batch, int32_t* lanes)
{
    if (!batch || !lanes) return set_error(BRKGA_ERR_INVALID_ARG, "null argument");
    *lanes = batch->fleet->lanes();
    return BRKGA_OK;
}

brkga_status brkga_batch_bucket_info(brkga_batch_t batch, int32_t bucket, int32_t* kind, int32_t* n_min,
                                     int32_t* n_max, int32_t* members, int32_t* block, int32_t* chunk)
{
    if (!batch) return set_error(BRKGA_ERR_INVALID_ARG, "null batch");
    const auto& bs = batch->fleet->buckets();
    if (bucket < 0 || bucket >= (int)bs.size()) return set_error(BRKGA_ERR_INVALID_ARG, "bucket out of range");
    const FleetBucket& b = bs[bucket];
    if (kind) *kind = b.kind;
    if (n_min) *n_min = b.n_min;
    if (n_max) *n_max = b.n_max;
    if (members) *members = b.count;
    if (block) *block = b.block;
    if (chunk) *chunk = b.chunk;
    return BRKGA_OK;
}

}  // extern "C"
