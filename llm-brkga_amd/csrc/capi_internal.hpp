// capi_internal.hpp - host-side objects behind the C ABI handles.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/brkga_hip.h"
#include "engine_types.hpp"

namespace brkga {

brkga_status set_error(brkga_status st, const std::string& msg);
brkga_status hip_error(hipError_t e, const char* what);

#define BRKGA_HIP_TRY(expr)                                          \
    do {                                                             \
        hipError_t brkga_e_ = (expr);                                \
        if (brkga_e_ != hipSuccess) return ::brkga::hip_error(brkga_e_, #expr); \
    } while (0)

// Device allocation that reports OOM distinctly from other failures.
brkga_status device_alloc(void** p, size_t bytes, const char* what);

int fixed_shift(double max_abs_term);

// Makes `dev` current for the scope, restoring the previous device.
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        (void)hipSetDevice(dev);
    }
    ~DeviceGuard()
    {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

}  // namespace brkga

struct brkga_engine_s {
    int device;
    size_t lds_limit;  // bytes of LDS one workgroup may use
    int cu_count;
    // launch shapes the fleet autotuner picked, per (problem kind, size
    // class); shared by every batch and server of the engine
    struct Shape {
        int block, chunk;
        bool xys;
        double child_us;  // measured step time per child slot (0: not measured)
        int children;     // child slots the measurement ran over
    };
    std::map<std::pair<int, int>, Shape> tuned;
    std::mutex tuned_mu;
};

struct brkga_problem_s {
    brkga_engine_s* engine;
    brkga::ProblemDesc desc;  // device pointers
    std::vector<void*> owned;
    // The data uploads are asynchronous (the legacy stream, then `ready`):
    // every use of the data on the device first calls problem_ready.
    hipEvent_t ready = nullptr;
    std::atomic<bool> uploading{false};
};

namespace brkga {

// Source of a population's device buffers: plain hipMalloc / hipFree, or a
// sub-range of a pre-reserved arena (served populations), whose release
// does not wait for the device.
struct DeviceAllocator {
    virtual ~DeviceAllocator() = default;
    virtual brkga_status alloc(void** p, size_t bytes, const char* what) = 0;
    virtual void release(void* p) = 0;
};

}  // namespace brkga

struct brkga_pop_s {
    brkga_problem_s* problem;
    brkga::DeviceAllocator* allocator;  // nullptr: hipMalloc / hipFree
    brkga_params params;
    brkga::InstDesc host;     // descriptor mirror
    brkga::InstDesc* d_desc;  // descriptor in device memory
    std::vector<void*> owned;
    int nsga2;
    int block;                // workgroup size of the step kernel (threads per child)
    int children = 1;         // children per workgroup of the ranking step (1, 2, 4)
    int step_lds;             // LDS bytes of one step block
    int host_t;               // generation steps enqueued since initialize
    int* d_decode_row;        // scratch for get_best decode
    float* d_fit_row;
    int host_open = 0;        // host-evaluated generation begun and not yet ended: 1 initial, 2 step
    hipStream_t capture_stream;  // created on the first graph capture
    std::map<std::pair<int, int>, hipGraphExec_t> graphs;  // (steps, start parity)
    std::mutex mu;
};

namespace brkga {

// Waits for the problem's data uploads (once; later calls return at once).
brkga_status problem_ready(brkga_problem_s* pr);

// Validates params against a problem and fills `d` (pointers excluded).
brkga_status fill_inst_desc(const brkga_problem_s* pr, const brkga_params* q, InstDesc* d, int* nsga2);
// Allocates the ping-pong buffers and counters described by d (from `a`,
// or hipMalloc when null); create_population clears them.
brkga_status alloc_inst_buffers(InstDesc* d, int nsga2, DeviceAllocator* a, std::vector<void*>* owned);

// Device bytes create_population takes from an allocator for (pr, q), each
// buffer rounded up to 256 bytes.
brkga_status population_bytes(const brkga_problem_s* pr, const brkga_params* q, size_t* bytes);

// Population over `pr` with buffers from `a` (nullptr: hipMalloc).  The
// step counter starts at 0, or at -1 when gen0 < 0 (the first fleet tick
// then builds generation 0).  Uploads run on stream s; the caller orders
// later work after them.  With `deferred`, nothing is enqueued: the
// device set-up (cleared key buffers, counters, descriptor) is described
// there instead, for the caller to batch with others (launch_setup_pops).
brkga_status create_population(brkga_problem_s* pr, const brkga_params* q, DeviceAllocator* a, hipStream_t s,
                               int gen0, brkga_pop_s** out, PopSetup* deferred = nullptr);
// Frees a population's buffers through its allocator (no device-wide wait
// for arena-backed populations).
void destroy_population(brkga_pop_s* pop);

// Enqueues `steps` generation steps of one population through the
// single-population launches (as brkga_evolve does), after building its
// generation 0 first when init != 0; pop->host_t must hold the population's
// current generation.  Steps ignore gen_limit: the caller runs exactly the
// steps it wants.
brkga_status population_steps(brkga_pop_s* pop, int init, int steps, hipStream_t s);

// Slot of the best individual (Population::get_best): BRKGA, the better of
// the islands' rank-0 slots (order_or_rank = rank order); NSGA-II, the
// front-0 member with the largest crowding distance (order_or_rank = front
// ranks), first index on ties.
int pick_best(const InstDesc& d, int nsga2, const float* fit, const int* order_or_rank, const float* crowd);

}  // namespace brkga
