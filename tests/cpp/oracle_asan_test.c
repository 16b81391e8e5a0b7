/* oracle_asan_test.c - drives every entry point of the CPU oracle on small
 * problems of every kind, for a build with AddressSanitizer and
 * UndefinedBehaviorSanitizer (tests/test_asan.py).  Every buffer handed to the
 * oracle is a heap block of exactly the size the oracle's contract names, so
 * a read or write past it is reported.  Prints one line per case and "done". */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static uint32_t lcg_state = 12345u;

static float frand(float lo, float hi)
{
    lcg_state = lcg_state * 1664525u + 1013904223u;
    return lo + (hi - lo) * (float)(lcg_state >> 8) * (1.0f / 16777216.0f);
}

static float* fbuf(size_t count)
{
    return (float*)malloc(sizeof(float) * (count ? count : 1));
}

static int32_t* ibuf(size_t count)
{
    return (int32_t*)malloc(sizeof(int32_t) * (count ? count : 1));
}

static float* random_floats(size_t count, float lo, float hi)
{
    float* a = fbuf(count);
    for (size_t i = 0; i < count; ++i) a[i] = frand(lo, hi);
    return a;
}

/* symmetric matrix with zero diagonal */
static float* random_matrix(int m, float hi)
{
    float* a = fbuf((size_t)m * m);
    for (int i = 0; i < m; ++i) {
        a[(size_t)i * m + i] = 0.0f;
        for (int j = i + 1; j < m; ++j) a[(size_t)i * m + j] = a[(size_t)j * m + i] = frand(1.0f, hi);
    }
    return a;
}

static void finish_problem(orc_problem* p)
{
    p->fx0 = orc_problem_fixed_shift(p, 0);
    p->fx1 = orc_problem_fixed_shift(p, 1);
}

static orc_params params(int pop, int islands, int elite, int mutants, int maximize)
{
    orc_params q;
    memset(&q, 0, sizeof(q));
    q.pop = pop;
    q.elite = elite;
    q.mutants = mutants;
    q.rho = 0.7;
    q.islands = islands;
    q.mig_interval = islands > 1 ? 2 : 0;
    q.mig_count = islands > 1 ? 2 : 0;
    q.seed = 0x5eed1234abcdull;
    q.maximize = maximize;
    return q;
}

/* evaluate a few random rows with decoded output, then run a short BRKGA */
static void run_brkga(const char* name, const orc_problem* p, orc_params q, int generations)
{
    const int L = orc_chrom_len(p), M = orc_num_obj(p), D = orc_decoded_len(p);
    const int rows = 5;
    float* keys = random_floats((size_t)rows * L, 0.0f, 1.0f);
    float* fit = fbuf((size_t)rows * M);
    int32_t* dec = ibuf((size_t)rows * D);
    orc_evaluate(p, keys, rows, fit, dec);
    free(keys);
    free(dec);

    const int P = q.pop;
    keys = fbuf((size_t)P * L);
    float* fitness = fbuf((size_t)P);
    int32_t* perm = ibuf((size_t)P);
    float* keys2 = fbuf((size_t)P * L);
    float* fitness2 = fbuf((size_t)P);
    int32_t* perm2 = ibuf((size_t)P);
    orc_brkga_init(p, &q, keys, fitness, perm);
    for (int t = 0; t < generations; ++t) {
        orc_brkga_step(p, &q, t, keys, perm, keys2, fitness2, perm2);
        float* tk = keys; keys = keys2; keys2 = tk;
        float* tf = fitness; fitness = fitness2; fitness2 = tf;
        int32_t* tp = perm; perm = perm2; perm2 = tp;
    }
    const double best = orc_brkga_run(p, &q, generations, keys, fitness, perm);
    printf("%s brkga best %.6g fit0 %.6g\n", name, best, (double)fit[0]);
    free(fit);
    free(keys);
    free(fitness);
    free(perm);
    free(keys2);
    free(fitness2);
    free(perm2);
}

static void run_nsga2(const char* name, const orc_problem* p, orc_params q, int generations)
{
    const int N = q.pop, L = orc_chrom_len(p);
    float* keys = fbuf((size_t)N * L);
    float* obj = fbuf((size_t)N * 2);
    int32_t* rank = ibuf((size_t)N);
    float* crowd = fbuf((size_t)N);
    float* keys2 = fbuf((size_t)N * L);
    float* obj2 = fbuf((size_t)N * 2);
    int32_t* rank2 = ibuf((size_t)N);
    float* crowd2 = fbuf((size_t)N);
    orc_nsga2_init(p, &q, keys, obj, rank, crowd);
    int injections = 0;
    for (int t = 0; t < generations; ++t) {
        injections += orc_nsga2_step(p, &q, t, keys, obj, rank, crowd, keys2, obj2, rank2, crowd2);
        float* tk = keys; keys = keys2; keys2 = tk;
        float* to = obj; obj = obj2; obj2 = to;
        int32_t* tr = rank; rank = rank2; rank2 = tr;
        float* tc = crowd; crowd = crowd2; crowd2 = tc;
    }
    int32_t* kept = ibuf((size_t)N);
    const int unique = orc_diversity_unique(obj, N, 2, kept);
    printf("%s nsga2 obj0 %.6g %.6g unique %d injections %d\n", name, (double)obj[0], (double)obj[1], unique,
           injections);
    free(kept);
    free(keys);
    free(obj);
    free(rank);
    free(crowd);
    free(keys2);
    free(obj2);
    free(rank2);
    free(crowd2);
}

int main(void)
{
    orc_problem p;

    /* argsort with ties and NaN keys */
    {
        const int n = 9;
        float* k = fbuf(n);
        const float v[9] = {0.5f, NAN, 0.25f, 0.5f, 0.0f, NAN, 1.0f, -1.0f, 0.25f};
        memcpy(k, v, sizeof(v));
        int32_t* perm = ibuf(n);
        orc_argsort_keys(k, n, perm);
        printf("argsort");
        for (int i = 0; i < n; ++i) printf(" %d", perm[i]);
        printf("\n");
        free(k);
        free(perm);
    }

    /* TSP, coordinates, islands with migration */
    memset(&p, 0, sizeof(p));
    p.kind = ORC_TSP_COORDS;
    p.n = 37;
    float* xy = random_floats((size_t)2 * p.n, 0.0f, 1000.0f);
    p.a = xy;
    finish_problem(&p);
    run_brkga("tsp_coords", &p, params(48, 2, 8, 6, 0), 6);

    /* TSP, matrix */
    memset(&p, 0, sizeof(p));
    p.kind = ORC_TSP_MATRIX;
    p.n = 23;
    float* mat = random_matrix(p.n, 500.0f);
    p.a = mat;
    finish_problem(&p);
    run_brkga("tsp_matrix", &p, params(32, 1, 6, 4, 0), 5);

    /* knapsack, threshold and greedy (maximisation) */
    const int items = 41;
    int32_t* w = ibuf(items);
    int32_t* v = ibuf(items);
    for (int i = 0; i < items; ++i) {
        w[i] = 1 + (int32_t)frand(0.0f, 60.0f);
        v[i] = 1 + (int32_t)frand(0.0f, 90.0f);
    }
    for (int greedy = 0; greedy < 2; ++greedy) {
        memset(&p, 0, sizeof(p));
        p.kind = greedy ? ORC_KNAP_GREEDY : ORC_KNAP_THRESHOLD;
        p.n = items;
        p.w = w;
        p.v = v;
        p.cap = 400;
        finish_problem(&p);
        run_brkga(greedy ? "knapsack_greedy" : "knapsack_threshold", &p, params(40, 2, 8, 6, 1), 5);
    }

    /* TSPJ with the local search */
    memset(&p, 0, sizeof(p));
    p.kind = ORC_TSPJ;
    p.n = 12;
    float* travel = random_matrix(p.n + 1, 40.0f);
    float* jobs = random_floats((size_t)(p.n + 1) * (p.n + 1), 1.0f, 20.0f);
    p.a = travel;
    p.b = jobs;
    finish_problem(&p);
    {
        orc_params q = params(32, 2, 8, 4, 0);
        q.ls_interval = 2;
        q.ls_count = 3;
        q.ls_moves = 5;
        run_brkga("tspj", &p, q, 6);
        const int L = orc_chrom_len(&p), P = q.pop;
        float* keys = fbuf((size_t)P * L);
        float* fitness = fbuf((size_t)P);
        int32_t* perm = ibuf((size_t)P);
        orc_brkga_init(&p, &q, keys, fitness, perm);
        orc_tspj_local_search(&p, &q, 1, keys, fitness, perm);
        printf("tspj local_search best %.6g\n", (double)fitness[perm[0]]);
        free(keys);
        free(fitness);
        free(perm);
    }

    /* bi-objective TSP, coordinates and matrices, with diversity injection */
    memset(&p, 0, sizeof(p));
    p.kind = ORC_BITSP_COORDS;
    p.n = 15;
    float* xy0 = random_floats((size_t)2 * (p.n + 1), 0.0f, 100.0f);
    float* xy1 = random_floats((size_t)2 * (p.n + 1), 0.0f, 100.0f);
    p.a = xy0;
    p.b = xy1;
    finish_problem(&p);
    {
        orc_params q = params(30, 1, 0, 0, 0);
        q.div_interval = 2;
        run_nsga2("bitsp_coords", &p, q, 6);
    }
    memset(&p, 0, sizeof(p));
    p.kind = ORC_BITSP_MATRIX;
    p.n = 9;
    float* m0 = random_matrix(p.n + 1, 50.0f);
    float* m1 = random_matrix(p.n + 1, 50.0f);
    p.a = m0;
    p.b = m1;
    finish_problem(&p);
    run_nsga2("bitsp_matrix", &p, params(24, 1, 0, 0, 0), 4);
    /* three cities: at most six distinct objective pairs among 30 members, so
     * every diversity check injects */
    memset(&p, 0, sizeof(p));
    p.kind = ORC_BITSP_COORDS;
    p.n = 3;
    float* txy0 = random_floats((size_t)2 * (p.n + 1), 0.0f, 100.0f);
    float* txy1 = random_floats((size_t)2 * (p.n + 1), 0.0f, 100.0f);
    p.a = txy0;
    p.b = txy1;
    finish_problem(&p);
    {
        orc_params q = params(30, 1, 0, 0, 0);
        q.div_interval = 1;
        run_nsga2("bitsp_three_cities", &p, q, 5);
    }
    free(txy0);
    free(txy1);

    /* continuous functions */
    for (int f = ORC_SPHERE; f <= ORC_ROSENBROCK; ++f) {
        memset(&p, 0, sizeof(p));
        p.kind = ORC_CONTINUOUS;
        p.n = 10;
        p.ip0 = f;
        p.f0 = -5.12f;
        p.f1 = 5.12f;
        finish_problem(&p);
        run_brkga(f == ORC_SPHERE ? "sphere" : f == ORC_RASTRIGIN ? "rastrigin" : "rosenbrock", &p,
                  params(32, 1, 6, 4, 0), 4);
    }

    /* job scheduling */
    memset(&p, 0, sizeof(p));
    p.kind = ORC_JOBSHOP;
    p.n = 19;
    float* times = random_floats((size_t)p.n, 1.0f, 30.0f);
    p.a = times;
    p.ip0 = 3;
    finish_problem(&p);
    run_brkga("jobshop", &p, params(32, 1, 6, 4, 0), 4);

    /* ZDT1 */
    memset(&p, 0, sizeof(p));
    p.kind = ORC_ZDT1;
    p.n = 8;
    finish_problem(&p);
    run_nsga2("zdt1", &p, params(20, 1, 0, 0, 0), 4);

    /* VRP with day routes */
    memset(&p, 0, sizeof(p));
    p.kind = ORC_VRP_DAYS;
    p.n = 14;
    float* tm = random_matrix(p.n + 1, 60.0f);
    p.a = tm;
    p.ip0 = 4;
    p.d0 = 240.0;
    finish_problem(&p);
    run_brkga("vrp_days", &p, params(32, 1, 6, 4, 0), 4);

    /* bi-objective CVRP */
    memset(&p, 0, sizeof(p));
    p.kind = ORC_MOVRP;
    p.n = 13;
    float* cxy = random_floats((size_t)2 * (p.n + 1), 0.0f, 100.0f);
    int32_t* demand = ibuf((size_t)p.n + 1);
    demand[0] = 0;
    for (int i = 1; i <= p.n; ++i) demand[i] = 1 + (int32_t)frand(0.0f, 9.0f);
    p.a = cxy;
    p.w = demand;
    p.cap = 20;
    p.ip0 = 3;
    finish_problem(&p);
    run_nsga2("movrp", &p, params(24, 1, 0, 0, 0), 4);

    free(xy);
    free(mat);
    free(w);
    free(v);
    free(travel);
    free(jobs);
    free(xy0);
    free(xy1);
    free(m0);
    free(m1);
    free(times);
    free(tm);
    free(cxy);
    free(demand);
    printf("done\n");
    return 0;
}
