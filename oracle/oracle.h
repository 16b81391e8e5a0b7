/*
 * oracle.h - CPU oracle for the BRKGA / NSGA-II generation step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP engine
 * is compared against: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The engine never links, loads or falls back
 * to it.
 *
 * It restates, operator for operator, the reference's algorithms on the
 * engine's counter-based Philox stream (the reference itself is clock-seeded
 * and uses cuRAND, so it cannot be replayed).  Every function names the
 * reference file:line it follows; paths are relative to the reference's
 * brkga/ directory.
 *
 * Pinned semantics (identical in the HIP engine):
 *   RNG     Philox4x32-10 (Salmon et al., SC'11), key = (seed_lo, seed_hi),
 *           counter = (stream, gen, slot, tag<<24 | word_block).
 *   keys    float(u >> 8) * 2^-24, i.e. exact 24-bit values in [0,1).
 *   decode  stable argsort of keys: equal keys keep ascending index order.
 *   sums    every edge/term is an FP32 value computed without FMA
 *           contraction; terms are accumulated as int64 fixed point
 *           llrint(term * 2^S) and converted once, round-to-nearest, to FP32.
 *   sort    partition = stable sort of (fitness, slot) per island.
 */
#ifndef BRKGA_ORACLE_H
#define BRKGA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_TAG_PARENTS = 0,
    ORC_TAG_COINS = 1,
    ORC_TAG_MUTANT = 2,
    ORC_TAG_INIT = 3,
    ORC_TAG_MUTCOIN = 4,
    ORC_TAG_LSEARCH = 5,
    ORC_TAG_DIVERSITY = 6
};

enum {
    ORC_TSP_COORDS = 0,
    ORC_TSP_MATRIX = 1,
    ORC_KNAP_THRESHOLD = 2,
    ORC_KNAP_GREEDY = 3,
    ORC_TSPJ = 4,
    ORC_BITSP_COORDS = 5,
    ORC_BITSP_MATRIX = 6,
    ORC_CONTINUOUS = 7,
    ORC_JOBSHOP = 8,
    ORC_ZDT1 = 9,
    ORC_VRP_DAYS = 10,
    ORC_MOVRP = 11
};
enum { ORC_SPHERE = 0, ORC_RASTRIGIN = 1, ORC_ROSENBROCK = 2 };

typedef struct {
    int kind;
    int n;            /* cities (excluding depot where one exists) or items */
    const float* a;   /* coords [n][2] / matrix [n][n] / TSPJ travel [(n+1)^2] / bi-TSP set 0 */
    const float* b;   /* TSPJ jobs [(n+1)^2] / bi-TSP set 1 */
    const int32_t* w; /* knapsack weights */
    const int32_t* v; /* knapsack values */
    int64_t cap;      /* knapsack capacity / vehicle capacity */
    int fx0, fx1;     /* fixed-point shifts (objective 0 / 1) */
    int ip0;          /* continuous function / machines / stops per day / vehicles */
    float f0, f1;     /* continuous bounds */
    double d0;        /* VRP days: drive limit per day (minutes) */
} orc_problem;

typedef struct {
    int pop;          /* total population (all islands) */
    int elite;        /* total elite (split evenly over islands) */
    int mutants;      /* total mutants (split evenly over islands) */
    double rho;       /* elite_prob */
    int islands;
    int mig_interval;
    int mig_count;
    uint64_t seed;
    int maximize;
    int ls_interval;  /* TSPJ local search every ls_interval generations (0 = off) */
    int ls_count;     /* individuals per island (top of the rank order), capped at the elite */
    int ls_moves;     /* random 2-opt moves per individual */
    int div_interval; /* NSGA-II diversity injection after steps t % div_interval == 0 (0 = off) */
} orc_params;

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
uint32_t orc_philox_word(uint64_t seed, uint32_t stream, uint32_t gen, uint32_t slot,
                         uint32_t tag, uint32_t word);
float orc_u32_to_unit(uint32_t u);

int orc_fixed_shift(double max_abs_term);
int orc_problem_fixed_shift(const orc_problem* p, int objective);
int orc_chrom_len(const orc_problem* p);
int orc_num_obj(const orc_problem* p);
int orc_decoded_len(const orc_problem* p);
int orc_maximize(const orc_problem* p);

void orc_argsort_keys(const float* keys, int n, int32_t* perm);
float orc_tsp_edge_coords(const float* xy, int a, int b);
/* Deterministic FP32 cosine (FP64 Cody-Waite reduction + Taylor series);
 * the engine evaluates the identical operation sequence. */
float orc_cos_det(float y);

void orc_evaluate(const orc_problem* p, const float* keys, int count,
                  float* fitness, int32_t* decoded);

void orc_init_keys(uint64_t seed, uint32_t stream, int slots, int len, float* keys);
void orc_partition(const float* fitness, int segments, int seg_len, int maximize,
                   int32_t* perm);

void orc_brkga_init(const orc_problem* p, const orc_params* q,
                    float* keys, float* fitness, int32_t* perm);
void orc_brkga_step(const orc_problem* p, const orc_params* q, int t,
                    const float* keys_in, const int32_t* perm_in,
                    float* keys_out, float* fitness_out, int32_t* perm_out);
double orc_brkga_run(const orc_problem* p, const orc_params* q, int generations,
                     float* keys, float* fitness, int32_t* perm);
/* TSPJ local search of generation g in place (keys, fitness, rank order);
 * orc_brkga_step applies it to the generation it builds. */
void orc_tspj_local_search(const orc_problem* p, const orc_params* q, int g,
                           float* keys, float* fitness, int32_t* perm);

void orc_nds_ranks(const float* obj, int count, int nobj, int32_t* rank);
void orc_crowding(const float* obj, const int32_t* rank, int count, int nobj, float* crowd);
void orc_nsga2_select(const float* obj, const int32_t* rank, const float* crowd,
                      int count, int nobj, int keep, int32_t* order);
void orc_nsga2_init(const orc_problem* p, const orc_params* q,
                    float* keys, float* obj, int32_t* rank, float* crowd);
/* Returns 1 when the step ended with a diversity injection. */
int orc_nsga2_step(const orc_problem* p, const orc_params* q, int t,
                   const float* keys_in, const float* obj_in,
                   const int32_t* rank_in, const float* crowd_in,
                   float* keys_out, float* obj_out, int32_t* rank_out, float* crowd_out);
/* First pass of ensure_minimum_diversity: indices of the members kept as
 * unique, in order; returns their count. */
int orc_diversity_unique(const float* obj, int count, int nobj, int32_t* kept);
/* The whole injection on one generation of N members (in place); returns 1
 * when it replaced members. */
int orc_nsga2_diversity(const orc_problem* p, const orc_params* q, int t,
                        float* keys, float* obj, int32_t* rank, float* crowd);

#ifdef __cplusplus
}
#endif

#endif
