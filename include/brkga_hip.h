/*
 * brkga_hip.h - C ABI of the MI355X (gfx950) BRKGA / NSGA-II engine.
 *
 * The engine runs the whole generation step of the reference's BRKGA solver
 * on the GPU: key generation, decode + fitness, partition by fitness, elite
 * copy, biased crossover, mutants and island ring migration, plus NSGA-II
 * ranking / crowding / survivor selection.  Populations stay resident in
 * device memory; only results cross this boundary.
 *
 * Conventions
 *   - Every function returns a brkga_status and never throws.  On failure
 *     brkga_last_error() (thread-local) describes the cause.
 *   - `stream` arguments are hipStream_t values passed as void*; NULL means
 *     the null stream of the engine's device.  Functions that return host
 *     data synchronise the stream they use.
 *   - Pointers named d_* are device pointers owned by the caller; all other
 *     pointers are host memory owned by the caller.  Problem data is copied
 *     into engine-owned device memory at create time.
 *   - Handles are destroyed in reverse order of creation (populations,
 *     batches and servers before their problems, problems before the engine).
 *   - One engine per host thread, or external synchronisation; the server
 *     functions are thread-safe.
 *
 * Reference interfaces replaced (paths relative to the reference's brkga/):
 *   core/config.hpp       BRKGAConfig<T> (plug-in base, evaluate_population_gpu)
 *   core/population.hpp   Population<T>
 *   core/solver.hpp       Solver<T> generation loops and island model
 *   core/gpu_population.cuh  GPUPopulation<T> (device-resident population API)
 *   configs/<problem>_config.hpp  TSPConfig, KnapsackConfig, TSPJConfig, MultiTSPConfig
 */
#ifndef BRKGA_HIP_H
#define BRKGA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    BRKGA_OK = 0,
    BRKGA_ERR_INVALID_ARG = 1, /* bad sizes / parameters (config.hpp:159-179 is_valid) */
    BRKGA_ERR_OOM = 2,         /* device allocation failed or over budget */
    BRKGA_ERR_DEVICE = 3,      /* HIP runtime error */
    BRKGA_ERR_BUSY = 4,        /* result not ready yet */
    BRKGA_ERR_DEADLINE = 5,    /* request retired without completing */
    BRKGA_ERR_NOT_FOUND = 6,   /* unknown request id */
    BRKGA_ERR_NO_DEVICE = 7,   /* no usable GPU */
    BRKGA_ERR_CANCELLED = 8    /* request cancelled before its deadline (brkga_server_cancel) */
} brkga_status;

/* Problem kinds (decoders).  BRKGA runs the single-objective kinds,
 * NSGA-II the bi-objective ones. */
typedef enum {
    BRKGA_TSP_COORDS = 0,       /* configs/tsp_config.hpp (Euclidean, coordinates) */
    BRKGA_TSP_MATRIX = 1,       /* configs/tsp_config.hpp:38-61 (n x n matrix) */
    BRKGA_KNAPSACK_THRESHOLD = 2, /* configs/knapsack_config.hpp:43-75 */
    BRKGA_KNAPSACK_GREEDY = 3,  /* greedy fill in key order (no reference counterpart) */
    BRKGA_TSPJ = 4,             /* configs/tspj_config.hpp:42-65, 555-642 */
    BRKGA_BITSP_COORDS = 5,     /* configs/multi_tsp_config.hpp, two coordinate sets */
    BRKGA_BITSP_MATRIX = 6,     /* configs/multi_tsp_config.hpp:25-52 (two matrices) */
    BRKGA_CONTINUOUS = 7,       /* configs/continuous_opt_gpu_config.hpp:28-49, 116-160 */
    BRKGA_JOB_SCHEDULING = 8,   /* configs/job_scheduling_gpu_config.hpp:28-52, 138-158 */
    BRKGA_ZDT1 = 9,             /* examples/zdt1.cu:14-63 (two objectives) */
    BRKGA_VRP_DAYS = 10,        /* configs/vrp_blazer_config.hpp:52-296 */
    BRKGA_MOVRP = 11,           /* configs/multi_objective_vrp_config.hpp:57-300 (two objectives) */
    BRKGA_HOST = 12             /* no device decoder: evaluated by the caller (brkga_problem_create_host) */
} brkga_problem_kind;

/* Continuous test functions (BRKGA_CONTINUOUS). */
typedef enum { BRKGA_SPHERE = 0, BRKGA_RASTRIGIN = 1, BRKGA_ROSENBROCK = 2 } brkga_continuous_fn;

typedef struct brkga_engine_s* brkga_engine_t;
typedef struct brkga_problem_s* brkga_problem_t;
typedef struct brkga_pop_s* brkga_pop_t;

/* GA parameters: BRKGAConfig<T> fields (core/config.hpp:22-26) plus the
 * island / migration / seed settings the reference hard-codes
 * (core/solver.hpp:164-167). */
typedef struct {
    int32_t population_size;    /* total over all islands */
    int32_t elite_size;         /* total over all islands (BRKGA) */
    int32_t mutant_size;        /* total over all islands (BRKGA) */
    double elite_prob;          /* rho: P(gene from elite parent); NSGA-II crossover coin */
    int32_t num_islands;        /* >= 1; population_size must divide evenly */
    int32_t migration_interval; /* generations between ring migrations (0 = never) */
    int32_t num_migrants;       /* top-k rows sent to the next island, k <= elite per island */
    uint64_t seed;              /* Philox key */
    /* TSPJ local search (tspj_config.hpp:77-135, solver.hpp:458-462,562-647):
     * after every ls_interval-th generation, ls_moves random 2-opt moves on
     * the decoded city tour of the first ls_individuals (at most the elite)
     * of every island, kept when the makespan strictly improves.  0 = off
     * (the reference runs it every 2 generations on up to 100 elites with 5
     * moves).  TSPJ problems only. */
    int32_t ls_interval;
    int32_t ls_individuals;
    int32_t ls_moves;
    /* NSGA-II diversity injection (Population::ensure_minimum_diversity,
     * core/population.hpp:130-182, which Solver runs every 25 generations,
     * core/solver.hpp:812-814): after every generation step t with
     * t % diversity_interval == 0, members closer than 0.001 in objective
     * space to an earlier kept member count as duplicates; when fewer than
     * 70% are unique, the unique ones stay (in order) and the rest are
     * replaced by random individuals, and the population is ranked and
     * crowded again.  0 = off; 25 = the reference's schedule.  Ignored by
     * single-objective problems, as in the reference. */
    int32_t diversity_interval;
} brkga_params;

/* ------------------------------------------------------------ engine ---- */

/* Version string of the engine library. */
const char* brkga_version(void);

/* Thread-local description of the last failure on this thread. */
const char* brkga_last_error(void);

/* Number of visible HIP devices (0 when none). */
brkga_status brkga_device_count(int32_t* count);

/* Binds an engine to one device.  Replaces Solver::initialize_gpu_system /
 * detect_gpus (core/solver.hpp:1307-1394). */
brkga_status brkga_engine_create(int32_t device, brkga_engine_t* out);
brkga_status brkga_engine_destroy(brkga_engine_t engine);

/* ----------------------------------------------------------- problems --- */

/* Every constructor copies the caller's arrays to the device asynchronously
 * on the legacy stream; the arrays may be reused or freed as soon as the
 * call returns, and the problem's first device use waits for the copy. */

/* Euclidean TSP over n cities from FP32 coordinates xy[n][2]; fitness is
 * the closed-tour length, minimised.  Same decoder and edge formula as
 * TSPConfig (configs/tsp_config.hpp:180-225, 359-371), without the n x n
 * matrix. */
brkga_status brkga_problem_create_tsp_coords(brkga_engine_t engine, const float* xy, int32_t n,
                                             brkga_problem_t* out);

/* TSP over an explicit distance matrix dist[n][n] (TSPConfig ctor,
 * configs/tsp_config.hpp:38-61). */
brkga_status brkga_problem_create_tsp_matrix(brkga_engine_t engine, const float* dist, int32_t n,
                                             brkga_problem_t* out);

/* 0/1 knapsack with integer weights/values, maximised.  greedy = 0: the
 * reference's threshold decoder (key > 0.5, penalty 1000 per unit over
 * capacity; configs/knapsack_config.hpp:43-75).  greedy = 1: pack items in
 * ascending key order when they fit; exact integer value; weights and
 * capacity must be >= 0. */
brkga_status brkga_problem_create_knapsack(brkga_engine_t engine, const int32_t* weights, const int32_t* values,
                                           int32_t n, int64_t capacity, int32_t greedy, brkga_problem_t* out);

/* TSPJ from travel[dim][dim] and jobs[dim][dim] (row/column 0 = depot,
 * dim = n + 1); chromosome = n city keys then n job keys; makespan
 * minimised (TSPJConfig, configs/tspj_config.hpp:42-65, 555-642). */
brkga_status brkga_problem_create_tspj(brkga_engine_t engine, const float* travel, const float* jobs,
                                       int32_t dim, brkga_problem_t* out);

/* Bi-objective TSP (MultiTSPConfig, configs/multi_tsp_config.hpp:25-103):
 * depot 0 + n cities; objective m is the depot-to-depot tour length in
 * metric m.  _coords: two coordinate sets xy0/xy1 of n+1 points each;
 * _matrix: two matrices of dimension dim = n + 1. */
brkga_status brkga_problem_create_bitsp_coords(brkga_engine_t engine, const float* xy0, const float* xy1,
                                               int32_t n, brkga_problem_t* out);
brkga_status brkga_problem_create_bitsp_matrix(brkga_engine_t engine, const float* m0, const float* m1,
                                               int32_t dim, brkga_problem_t* out);

/* Continuous minimisation over n variables x_i = key_i * (upper - lower) +
 * lower (FP32): sphere sum x^2, Rastrigin 10 n + sum (x^2 - 10 cos(2 pi x)),
 * Rosenbrock sum 100 (x_{i+1} - x_i^2)^2 + (1 - x_i)^2
 * (ContinuousOptGPUConfig, configs/continuous_opt_gpu_config.hpp:28-49,
 * 116-160; the reference evaluates Rosenbrock on the host only).  Decoded
 * row: the FP32 bits of each x_i. */
brkga_status brkga_problem_create_continuous(brkga_engine_t engine, int32_t function, int32_t n, float lower,
                                             float upper, brkga_problem_t* out);

/* Job scheduling: job i (time job_times[i]) runs on machine int(key_i *
 * machines) (clamped to [0, machines)); fitness = largest machine load,
 * minimised (JobSchedulingGPUConfig, configs/job_scheduling_gpu_config.hpp:
 * 28-52, 138-158).  1 <= machines <= 4096.  Decoded row: machine per job. */
brkga_status brkga_problem_create_job_scheduling(brkga_engine_t engine, const float* job_times, int32_t n,
                                                 int32_t machines, brkga_problem_t* out);

/* ZDT1 over n variables (the keys): objectives f1 = x_1 and f2 = g (1 -
 * sqrt(f1 / g)), g = 1 + 9 sum(x_2..x_n) / (n - 1), both minimised; runs
 * on NSGA-II (ZDT1Config, examples/zdt1.cu:14-63).  Decoded row: the FP32
 * bits of each x_i. */
brkga_status brkga_problem_create_zdt1(brkga_engine_t engine, int32_t n, brkga_problem_t* out);

/* Multi-day routes (VrpBlazerConfig, configs/vrp_blazer_config.hpp:
 * 52-296): time_matrix[dim][dim] with location 0 the depot and n = dim - 1
 * stops.  Stops in key order fill days of at most max_stops_per_day stops
 * whose drive time (legs plus return to the depot) stays within
 * max_drive_minutes; fitness = total drive minutes + 1000 per entry of the
 * reference's unvisited list (which holds every skipped stop twice),
 * minimised.  Decoded row: day (1-based) of each stop, 0 if skipped. */
brkga_status brkga_problem_create_vrp_days(brkga_engine_t engine, const float* time_matrix, int32_t dim,
                                           int32_t max_stops_per_day, double max_drive_minutes,
                                           brkga_problem_t* out);

/* Bi-objective capacitated VRP (MultiObjectiveVRPConfig,
 * configs/multi_objective_vrp_config.hpp:57-300): xy[dim][2] and
 * demands[dim] with location 0 the depot and n = dim - 1 customers.
 * Customers in biased key order split greedily into routes by capacity;
 * objectives: total distance (+100 per unit of overload) and the longest
 * route, each +500 per route beyond `vehicles`; runs on NSGA-II.  Decoded
 * row: route (1-based) of each customer. */
brkga_status brkga_problem_create_movrp(brkga_engine_t engine, const float* xy, const int32_t* demands, int32_t dim,
                                        int64_t capacity, int32_t vehicles, brkga_problem_t* out);

/* A problem without a device decoder, evaluated by the caller on the host:
 * the reference's BRKGAConfig with a host fitness_function /
 * objective_functions (core/config.hpp:33-39), which it evaluates with
 * Population::evaluate_all_cpu (core/population.hpp:626-653).  Its
 * populations run through brkga_population_host_begin / _end (one call pair
 * per generation); keys are built and partitioned (BRKGA) or ranked and
 * selected (NSGA-II, num_obj = 2) on the device.  num_obj 1 or 2; maximize
 * for single-objective problems only.  brkga_evaluate / brkga_decode /
 * brkga_evolve and the batched and served forms refuse such problems. */
brkga_status brkga_problem_create_host(brkga_engine_t engine, int32_t chrom_len, int32_t num_obj, int32_t maximize,
                                       brkga_problem_t* out);

/* Kind, chromosome length, objectives and direction of a problem. */
brkga_status brkga_problem_info(brkga_problem_t problem, int32_t* kind, int32_t* chrom_len, int32_t* num_obj,
                                int32_t* maximize);
brkga_status brkga_problem_destroy(brkga_problem_t problem);

/* Decode + fitness of pop_size caller-owned device rows d_keys[pop][chrom_len];
 * writes d_fitness[pop][num_obj] (the layout of population.hpp:609-624).
 * Replaces BRKGAConfig::evaluate_population_gpu (core/config.hpp:50-53)
 * and the configs' fitness kernels; asynchronous on `stream`. */
brkga_status brkga_evaluate(brkga_problem_t problem, const float* d_keys, float* d_fitness, int32_t pop_size,
                            int32_t chrom_len, void* stream);

/* Decoded solutions d_decoded[pop][chrom_len] (int32): TSP / bi-TSP city
 * order (0-based), knapsack 0/1 selection per item, TSPJ 1-based city ids
 * then 1-based job ids.  Replaces BRKGAConfig::decoder (core/config.hpp:34). */
brkga_status brkga_decode(brkga_problem_t problem, const float* d_keys, int32_t* d_decoded, int32_t pop_size,
                          int32_t chrom_len, void* stream);

/* Host-buffer form of brkga_evaluate + brkga_decode for a few rows: copies
 * keys[count][chrom_len] to the device, decodes and evaluates them there and
 * copies fitness[count][num_obj] / decoded[count][chrom_len] back (either
 * may be NULL).  Backs the host-side fitness_function / decoder callbacks of
 * the configs (core/config.hpp:33-34), which the reference evaluates with
 * its CPU code.  Synchronous. */
brkga_status brkga_evaluate_host(brkga_problem_t problem, const float* keys, int32_t count, float* fitness,
                                 int32_t* decoded);

/* brkga_evaluate_host without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (tight
 * per-individual evaluation loops such as a config's fitness_function
 * callback, harnesses that forbid hipMalloc).  The device staging of keys,
 * fitness and decoded rows lives in a static 4 MiB array per device,
 * resolved when an engine is created on that device; a call needs
 * align256(4 count chrom_len) + align256(4 count num_obj) (+ 4 count
 * chrom_len when decoded is not NULL) bytes and is rejected with
 * BRKGA_ERR_INVALID_ARG above 4 MiB.  Every call on a device shares that
 * array: calls from several host threads are serialised by the library.
 * Results are identical to brkga_evaluate_host.  Synchronous. */
brkga_status brkga_evaluate_host_noalloc(brkga_problem_t problem, const float* keys, int32_t count, float* fitness,
                                         int32_t* decoded);

/* -------------------------------------------------------- populations --- */

/* Device-resident population: ping-pong key buffers [P][ld], fitness,
 * rank order per island (and NSGA-II rank / crowding).  Replaces
 * Population<T> (core/population.hpp:72-100) and GPUPopulation<T>
 * (core/gpu_population.cuh:52-120).  Validation follows
 * BRKGAConfig::is_valid (core/config.hpp:159-179). */
brkga_status brkga_population_create(brkga_problem_t problem, const brkga_params* params, brkga_pop_t* out);
brkga_status brkga_population_destroy(brkga_pop_t pop);

/* Generation 0: Philox keys, evaluation and partition
 * (Population::initialize, core/population.hpp:107-124). */
brkga_status brkga_population_initialize(brkga_pop_t pop, void* stream);

/* `generations` generation steps (Solver::evolve_generation,
 * core/solver.hpp:213-231 / 378-476).  BRKGA islands no wider than the step
 * kernel's workgroup, without local search: one step launch per generation
 * (each workgroup ranks its parent island), then one partition that leaves
 * the newest generation's rank order for readers.  Otherwise each step is
 * followed by its partition (and any local search). */
brkga_status brkga_evolve(brkga_pop_t pop, int32_t generations, void* stream);

/* Same steps replayed from a captured HIP graph of `generations` steps
 * (the graph is built on first use and cached per length). */
brkga_status brkga_evolve_graph(brkga_pop_t pop, int32_t generations, void* stream);

/* One generation of a host-evaluated population, in two calls.
 * _begin builds generation 0 (initialize != 0) or the next generation's
 * keys on the device and copies to the host: keys[rows][chrom_len];
 * fitness[rows][num_obj] (may be NULL) holds the fitness of rows that are
 * copies (BRKGA elite and migrant slots, as evaluate_all_cpu skips
 * evaluated members), 0 elsewhere; needs[rows] (may be NULL) is 1 for every
 * row the caller must evaluate.  rows = population_size: the population
 * (BRKGA, generation 0) or the N children (NSGA-II steps).
 * _end takes fitness[rows][num_obj] for every row (copied rows as _begin
 * gave them) and completes the generation: partition by fitness (BRKGA) or
 * NSGA-II ranking, crowding and survivor selection.  Both synchronise
 * `stream`. */
brkga_status brkga_population_host_begin(brkga_pop_t pop, int32_t initialize, float* keys, float* fitness,
                                         int32_t* needs, void* stream);
brkga_status brkga_population_host_end(brkga_pop_t pop, const float* fitness, void* stream);

/* Completed generation steps since initialize. */
brkga_status brkga_population_generation(brkga_pop_t pop, int32_t* generation);

/* Copies the current population to host (GPUPopulation::sync_to_host):
 * keys[P][chrom_len], fitness[P][num_obj], order[P] (per island, the local
 * slot of each rank), rank[P] and crowding[P] (NSGA-II only).  Any pointer
 * may be NULL.  Synchronises `stream`. */
brkga_status brkga_population_read(brkga_pop_t pop, float* keys, float* fitness, int32_t* order, int32_t* rank,
                                   float* crowding, void* stream);

/* Replaces the keys (GPUPopulation::sync_from_host) and re-evaluates /
 * re-partitions them as a fresh generation 0. */
brkga_status brkga_population_write(brkga_pop_t pop, const float* keys, void* stream);

/* Resumes a population from a checkpoint taken with brkga_population_read
 * after `generation` completed steps: keys[P][chrom_len], fitness[P][num_obj]
 * and, for NSGA-II, rank[P] and crowding[P] (they depend on the parents and
 * children the last selection saw, so they are part of the state).  The rank
 * order of a BRKGA population is recomputed from the fitness.  Steps after
 * the restore are bit-identical to those of the run the checkpoint was taken
 * from (the generation counter drives the Philox streams, migration, local
 * search and diversity schedules).  The population must have been created
 * with the same problem and parameters.  No reference counterpart (the
 * reference has no checkpoint). */
brkga_status brkga_population_restore(brkga_pop_t pop, const float* keys, const float* fitness, const int32_t* rank,
                                      const float* crowding, int32_t generation, void* stream);

/* Best individual over all islands (Population::get_best,
 * Solver::get_best_individual): fitness[num_obj], keys[chrom_len] and the
 * decoded solution[chrom_len]; any pointer may be NULL.  NSGA-II: the
 * rank-0 member with the largest crowding distance
 * (population.hpp update_best_from_front). */
brkga_status brkga_get_best(brkga_pop_t pop, float* fitness, float* keys, int32_t* decoded, void* stream);

/* Raw device buffers of the current generation (row stride *ld floats). */
brkga_status brkga_population_device_buffers(brkga_pop_t pop, float** d_keys, float** d_fitness,
                                             int32_t** d_order, int32_t* ld);

/* Picks the workgroup size of the population's generation kernel by timing
 * `reps` back-to-back launches of each admissible size (64 ... 1024, at
 * least n / 16) on the current generation; the launches rewrite the same
 * next-generation rows, so the population does not change.  Results never
 * depend on the size.  *block (may be NULL) receives the choice. */
brkga_status brkga_population_autotune(brkga_pop_t pop, int32_t reps, int32_t* block, void* stream);

/* Sets the workgroup size of the population's generation kernel (64, 128,
 * 256, 512 or 1024, at least n / 16) without timing anything; the
 * counterpart of BRKGAConfig::threads_per_block (core/config.hpp:30, 80-82).
 * Results never depend on it. */
brkga_status brkga_population_set_block(brkga_pop_t pop, int32_t block);

/* Children per workgroup of the BRKGA ranking step (1, 2 or 4): with more
 * than one, a workgroup builds that many children of one island and ranks
 * the island once for all of them (Euclidean TSP whose islands fit the
 * workgroup size and whose keys fit 4 per thread; otherwise 1 is used).
 * Results never depend on it; brkga_population_autotune picks it together
 * with the workgroup size. */
brkga_status brkga_population_set_children(brkga_pop_t pop, int32_t children);

/* Launch shape the generation kernel runs with: threads per child (*block)
 * and children per workgroup (*children, 1 unless the multi-child step
 * applies).  Either pointer may be NULL. */
brkga_status brkga_population_step_shape(brkga_pop_t pop, int32_t* block, int32_t* children);

/* Mean device time of the generation kernel (BRKGA step / NSGA-II
 * offspring) and of the kernels that close a generation (partition and any
 * local search / NSGA-II selection + gather), in microseconds, over `reps`
 * real generations enqueued on `stream` as brkga_evolve does.  BRKGA without
 * local search: each kernel's own start / stop time (events attached to its
 * dispatch); otherwise HIP events recorded between the kernels.  When the
 * island fits the step's workgroup, the step ranks its parents itself and
 * *close_us is 0 (one partition, untimed, ends the run).  One
 * synchronisation at the end; the population advances by `reps`
 * generations.  Replaces the reference's PerformanceStats timers
 * (core/solver.hpp:143-157, 406-438), which time synchronous launches with
 * std::chrono. */
brkga_status brkga_population_profile(brkga_pop_t pop, int32_t reps, double* step_us, double* close_us,
                                      void* stream);

/* ------------------------------------------------------------ batches --- */

/* Many independent populations (any problem kinds, uneven sizes) evolved
 * together.  The reference has no batched form: it runs one Solver per
 * instance (core/solver.hpp:1517-1556); a batch member's result equals that
 * of the same population evolved alone with brkga_evolve.
 *   BRKGA_BATCH_PLAIN   one member at a time through brkga_population_* /
 *                       brkga_evolve (one launch sequence per member and
 *                       generation); the per-instance form.
 *   BRKGA_BATCH_LAUNCH  all members per generation: one persistent
 *                       queue-driven step launch per size bucket, one
 *                       partition launch, NSGA-II launches per kind.
 *   BRKGA_BATCH_GRAPH   the same launches replayed from a captured graph of
 *                       the requested number of generations. */
typedef struct brkga_batch_s* brkga_batch_t;
typedef enum { BRKGA_BATCH_PLAIN = 0, BRKGA_BATCH_LAUNCH = 1, BRKGA_BATCH_GRAPH = 2 } brkga_batch_mode;

/* count populations, member i over problems[i] with params[i]. */
brkga_status brkga_batch_create(brkga_engine_t engine, int32_t count, const brkga_problem_t* problems,
                                const brkga_params* params, brkga_batch_t* out);
brkga_status brkga_batch_destroy(brkga_batch_t batch);

/* Generation 0 of every member, then `generations` steps of every member. */
brkga_status brkga_batch_initialize(brkga_batch_t batch, int32_t mode, void* stream);
brkga_status brkga_batch_evolve(brkga_batch_t batch, int32_t generations, int32_t mode, void* stream);

/* Picks the workgroup size, slots per work item and coordinate staging of
 * each size bucket by timing `reps` step launches per candidate on the
 * current populations (call after initialize).  Does not advance or change
 * any population.  The choices are kept per engine and (problem kind, size
 * class) and reused by later batches and servers of the engine. */
brkga_status brkga_batch_autotune(brkga_batch_t batch, int32_t reps, void* stream);

/* Member i as a population handle owned by the batch (read, get_best,
 * generation; do not destroy or evolve it on its own while the batch runs). */
brkga_status brkga_batch_population(brkga_batch_t batch, int32_t index, brkga_pop_t* out);

/* Size buckets of the batched launch: count, then per bucket its problem
 * kind, member n range and count, workgroup size and slots per item. */
brkga_status brkga_batch_num_buckets(brkga_batch_t batch, int32_t* count);
brkga_status brkga_batch_bucket_info(brkga_batch_t batch, int32_t bucket, int32_t* kind, int32_t* n_min,
                                     int32_t* n_max, int32_t* members, int32_t* block, int32_t* chunk);
/* Streams the batched launches are spread over: the caller's stream plus
 * those a timed probe found running concurrently with it (1 .. 4), or the
 * count set by the environment variable BRKGA_FLEET_LANES (1 .. 4).  1
 * until the first batched step. */
brkga_status brkga_batch_lanes(brkga_batch_t batch, int32_t* lanes);

/* ------------------------------------------------------------ serving --- */

/* A request loop over one device (no reference counterpart: the reference
 * solves one instance per process run, main.cu:46-305).  Requests of any
 * problem kind and size are admitted first-come first-served while their
 * populations fit the HBM budget (and the resident cap); resident requests
 * advance together, one generation each per tick, through the batched
 * kernels, and each retires at its own generation deadline with its best
 * individual.  Requests that arrive while others run join at the next
 * admission round.  A request's result does not depend on the others it
 * shared the device with. */
typedef struct brkga_server_s* brkga_server_t;

typedef struct {
    uint64_t hbm_budget_bytes; /* device arena for resident populations (0: 8 GiB) */
    int32_t max_resident;      /* most requests on the device at once (0: no cap) */
    int32_t ticks_per_round;   /* generations per resident request between admission rounds (0: 4) */
    int32_t autotune_reps;     /* > 0: tune the launch shape of each (kind, size class) the engine has
                                  not tuned yet, when it first appears, timing this many launches */
} brkga_server_config;

typedef struct {
    double submit_us; /* server clock: microseconds since brkga_server_start */
    double admit_us;
    double done_us;
    int32_t generations; /* generation steps completed */
    int32_t status;      /* brkga_status of the request */
    uint64_t device_bytes; /* arena bytes the request held while resident */
} brkga_request_info;

typedef struct {
    uint64_t submitted, admitted, completed, rejected;
    int32_t resident, pending;
    uint64_t bytes_in_use, peak_bytes, budget_bytes;
    uint64_t ticks;   /* batched ticks run */
    uint64_t rounds;  /* admission rounds */
    double wait_us;   /* worker time spent waiting for ticks to finish */
    double host_us;   /* worker time spent on admission, membership and results */
    int32_t lanes;    /* streams the batched ticks are spread over (as brkga_batch_lanes; 0 while no
                         batched tick has run, e.g. when every round ran its requests on their own streams) */
} brkga_server_stats;

/* Starts the worker thread (config may be NULL for defaults). */
brkga_status brkga_server_start(brkga_engine_t engine, const brkga_server_config* config, brkga_server_t* out);

/* Queues a solve of `problem` with `params` for `generations` steps.
 * own_problem = 1 hands the problem to the server, which destroys it once
 * the request is done; otherwise the caller keeps it alive until then.
 * A request whose population can never fit the budget is rejected with
 * BRKGA_ERR_OOM.  Thread-safe. */
brkga_status brkga_server_submit(brkga_server_t server, brkga_problem_t problem, const brkga_params* params,
                                 int32_t generations, int32_t own_problem, uint64_t* request_id);

/* Collects a finished request: best fitness[num_obj] and decoded
 * solution[chrom_len] (either may be NULL) and its timings.  Waits up to
 * timeout_ms (0: poll, < 0: no limit); BRKGA_ERR_BUSY if still running,
 * BRKGA_ERR_NOT_FOUND for an unknown or already collected id.  Thread-safe. */
brkga_status brkga_server_wait(brkga_server_t server, uint64_t request_id, int32_t timeout_ms, float* fitness,
                               int32_t* decoded, brkga_request_info* info);

brkga_status brkga_server_stats_get(brkga_server_t server, brkga_server_stats* stats);

/* Cancels a request that has not reached its deadline.  A pending request
 * leaves the queue at once; a resident one leaves the device at the start
 * of the worker's next admission round (the generations it ran are dropped
 * and its HBM returns to the budget).  brkga_server_wait then returns
 * BRKGA_ERR_CANCELLED for it, with its info.  A request that already
 * reached its deadline keeps its result (no effect, BRKGA_OK);
 * BRKGA_ERR_NOT_FOUND for an unknown or collected id.  Thread-safe.  No reference counterpart (the
 * reference solves one instance per process run). */
brkga_status brkga_server_cancel(brkga_server_t server, uint64_t request_id);

/* Finishes every submitted request, stops the worker and frees the server;
 * results not collected by then are dropped. */
brkga_status brkga_server_stop(brkga_server_t server);

#ifdef __cplusplus
}
#endif

#endif /* BRKGA_HIP_H */
