/*
 * oracle.c - CPU oracle for the BRKGA / NSGA-II generation step.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): a checker for the HIP engine,
 * never linked into it.  Single-threaded, written for clarity, not speed.
 *
 * Reference anchors (paths relative to the reference's brkga/ directory):
 *   Philox RNG replaces cuRAND     core/cuda_kernels.cuh:10-15,117-129
 *   TSP decode + tour length       configs/tsp_config.hpp:180-225, 359-371
 *   knapsack threshold decoder     configs/knapsack_config.hpp:43-75
 *   TSPJ decode + makespan         configs/tspj_config.hpp:555-642
 *   bi-objective TSP               configs/multi_tsp_config.hpp:54-103
 *   partition (sort by fitness)    core/population.hpp:660-682, core/solver.hpp:416-419
 *   generation step                core/cuda_kernels.cuh:88-138, core/solver.hpp:1109-1135,1244-1258
 *   islands + ring migration       core/solver.hpp:378-476, 517-558
 *   NSGA-II                        core/individual.hpp:99-114, core/population.hpp:334-369,409-517,
 *                                  core/solver.hpp:751-817,1019-1091
 *
 * Build: gcc -O2 -std=c11 -ffp-contract=off (no FMA contraction, SSE float).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ RNG -- */

/* Philox4x32-10 as published by Salmon, Moraes, Dror, Shaw, "Parallel random
 * numbers: as easy as 1, 2, 3" (SC'11): 10 rounds of two 32x32->64 multiplies
 * with the Weyl key schedule. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

uint32_t orc_philox_word(uint64_t seed, uint32_t stream, uint32_t gen, uint32_t slot,
                         uint32_t tag, uint32_t word)
{
    uint32_t ctr[4] = {stream, gen, slot, (tag << 24) | (word >> 2)};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t out[4];
    orc_philox4x32_10(ctr, key, out);
    return out[word & 3];
}

/* 24 high bits -> exact float in [0,1).  Stands in for curand_uniform
 * (cuda_kernels.cuh:117,129) and uniform_real_distribution (individual.hpp:80). */
float orc_u32_to_unit(uint32_t u)
{
    return (float)(u >> 8) * 0x1p-24f;
}

/* ------------------------------------------------------- fixed point sums -- */

/* Shift S so that every term |t| <= max_abs_term maps to |t * 2^S| <= 2^48:
 * up to 2^14 such terms sum without int64 overflow. */
int orc_fixed_shift(double max_abs_term)
{
    if (!(max_abs_term > 0.0) || !isfinite(max_abs_term)) return 30;
    int e;
    frexp(max_abs_term, &e); /* max_abs_term < 2^e */
    int s = 48 - e;
    if (s > 60) s = 60;
    if (s < -60) s = -60;
    return s;
}

static int64_t fx_term(float t, int shift)
{
    return llrintf(ldexpf(t, shift));
}

static float fx_to_float(int64_t acc, int shift)
{
    return (float)acc * ldexpf(1.0f, -shift);
}

/* ------------------------------------------------------------- sorting -- */

typedef int (*orc_less_fn)(int a, int b, const void* ctx);

/* Stable merge sort of an index array under a strict-weak "less". */
static void merge_sort_idx(int32_t* idx, int n, orc_less_fn less, const void* ctx)
{
    if (n < 2) return;
    int32_t* tmp = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    for (int width = 1; width < n; width *= 2) {
        for (int lo = 0; lo < n; lo += 2 * width) {
            int mid = lo + width < n ? lo + width : n;
            int hi = lo + 2 * width < n ? lo + 2 * width : n;
            int i = lo, j = mid, k = lo;
            while (i < mid && j < hi) {
                if (less(idx[j], idx[i], ctx)) tmp[k++] = idx[j++];
                else tmp[k++] = idx[i++];
            }
            while (i < mid) tmp[k++] = idx[i++];
            while (j < hi) tmp[k++] = idx[j++];
        }
        memcpy(idx, tmp, sizeof(int32_t) * (size_t)n);
    }
    free(tmp);
}

/* Key order: numbers ascending, -0 equal to +0, NaN after every number (and
 * equal to another NaN). */
static int key_less(int a, int b, const void* ctx)
{
    const float* k = (const float*)ctx;
    const float x = k[a], y = k[b];
    if (y != y) return x == x;
    return x < y;
}

/* std::sort over pair<float,int>(key, index) (tsp_config.hpp:209-216):
 * ascending key, ties by ascending index == a stable sort of the indices.
 * The reference leaves NaN keys undefined (std::sort needs a strict weak
 * order); here they sort last, in index order. */
void orc_argsort_keys(const float* keys, int n, int32_t* perm)
{
    for (int i = 0; i < n; ++i) perm[i] = i;
    merge_sort_idx(perm, n, key_less, keys);
}

/* --------------------------------------------------------------- decode -- */

/* Edge length exactly as tsp_config.hpp:366-368 builds its matrix, in FP32
 * with every product and sum rounded separately (no FMA). */
float orc_tsp_edge_coords(const float* xy, int a, int b)
{
    float dx = xy[2 * a] - xy[2 * b];
    float dy = xy[2 * a + 1] - xy[2 * b + 1];
    float sx = dx * dx;
    float sy = dy * dy;
    float s = sx + sy;
    return sqrtf(s);
}

int orc_chrom_len(const orc_problem* p)
{
    return p->kind == ORC_TSPJ ? 2 * p->n : p->n;
}

int orc_num_obj(const orc_problem* p)
{
    return (p->kind == ORC_BITSP_COORDS || p->kind == ORC_BITSP_MATRIX || p->kind == ORC_ZDT1 ||
            p->kind == ORC_MOVRP) ? 2 : 1;
}

int orc_decoded_len(const orc_problem* p)
{
    return p->kind == ORC_TSPJ ? 2 * p->n : p->n;
}

int orc_maximize(const orc_problem* p)
{
    return p->kind == ORC_KNAP_THRESHOLD || p->kind == ORC_KNAP_GREEDY;
}

static double max_abs(const float* a, size_t count)
{
    double m = 0.0;
    for (size_t i = 0; i < count; ++i) {
        double v = fabs((double)a[i]);
        if (v > m) m = v;
    }
    return m;
}

static double coords_diameter(const float* xy, int count)
{
    if (count <= 0) return 0.0;
    double x0 = xy[0], x1 = xy[0], y0 = xy[1], y1 = xy[1];
    for (int i = 1; i < count; ++i) {
        if (xy[2 * i] < x0) x0 = xy[2 * i];
        if (xy[2 * i] > x1) x1 = xy[2 * i];
        if (xy[2 * i + 1] < y0) y0 = xy[2 * i + 1];
        if (xy[2 * i + 1] > y1) y1 = xy[2 * i + 1];
    }
    double dx = x1 - x0, dy = y1 - y0;
    return sqrt(dx * dx + dy * dy) * (1.0 + 1e-6);
}

/* Fixed-point shift of one objective, from the largest term it can sum. */
int orc_problem_fixed_shift(const orc_problem* p, int objective)
{
    int n = p->n;
    switch (p->kind) {
    case ORC_TSP_COORDS: return orc_fixed_shift(coords_diameter(p->a, n));
    case ORC_TSP_MATRIX: return orc_fixed_shift(max_abs(p->a, (size_t)n * n));
    case ORC_TSPJ: {
        size_t d = (size_t)(n + 1) * (n + 1);
        double m0 = max_abs(p->a, d), m1 = max_abs(p->b, d);
        return orc_fixed_shift(m0 > m1 ? m0 : m1);
    }
    case ORC_BITSP_COORDS:
        return orc_fixed_shift(coords_diameter(objective ? p->b : p->a, n + 1));
    case ORC_BITSP_MATRIX:
        return orc_fixed_shift(max_abs(objective ? p->b : p->a, (size_t)(n + 1) * (n + 1)));
    case ORC_CONTINUOUS: {
        if (objective) return 30;
        double lo = fabs((double)p->f0), hi = fabs((double)p->f1);
        double m = lo > hi ? lo : hi;
        double bound = p->ip0 == ORC_SPHERE ? m * m
                       : p->ip0 == ORC_RASTRIGIN ? m * m + 10.0
                                                 : 100.0 * (m + m * m) * (m + m * m) + (1.0 + m) * (1.0 + m);
        return orc_fixed_shift(bound * (1.0 + 1e-6));
    }
    case ORC_JOBSHOP: return objective ? 30 : orc_fixed_shift(max_abs(p->a, (size_t)n));
    case ORC_ZDT1: return objective ? 30 : orc_fixed_shift(1.0);
    case ORC_VRP_DAYS: {
        if (objective) return 30;
        int per_day = p->ip0 < 0 ? 0 : (p->ip0 > n ? n : p->ip0);
        double a = (per_day + 1.0) * max_abs(p->a, (size_t)(n + 1) * (n + 1)), b = 2000.0 * n;
        return orc_fixed_shift(a > b ? a : b);
    }
    case ORC_MOVRP: {
        double m = coords_diameter(p->a, n + 1);
        int32_t md = 0;
        for (int i = 0; i <= n; ++i)
            if (p->w[i] > md) md = p->w[i];
        double pd = 100.0 * md, pv = 500.0 * n;
        double pm = pd > pv ? pd : pv;
        return orc_fixed_shift(m > pm ? m : pm);
    }
    default: return 30;
    }
}

/* Closed tour over all n cities (tsp_config.hpp:180-193). */
static float tsp_fitness(const orc_problem* p, const float* keys, int32_t* perm)
{
    int n = p->n;
    orc_argsort_keys(keys, n, perm);
    int64_t acc = 0;
    for (int i = 0; i < n; ++i) {
        int a = perm[i], b = perm[(i + 1) % n];
        float d = p->kind == ORC_TSP_COORDS ? orc_tsp_edge_coords(p->a, a, b)
                                            : p->a[(size_t)a * n + b];
        acc += fx_term(d, p->fx0);
    }
    return fx_to_float(acc, p->fx0);
}

/* Threshold decoder with linear penalty (knapsack_config.hpp:43-62): items
 * with key > 0.5 are packed; overweight costs 1000 per unit.  Integer
 * data is summed exactly and rounded once to FP32. */
static float knap_threshold(const orc_problem* p, const float* keys, int32_t* sel)
{
    int64_t tw = 0, tv = 0;
    for (int i = 0; i < p->n; ++i) {
        int take = keys[i] > 0.5f;
        if (take) { tw += p->w[i]; tv += p->v[i]; }
        if (sel) sel[i] = take;
    }
    int64_t f = tw > p->cap ? tv - 1000 * (tw - p->cap) : tv;
    return (float)f;
}

/* Greedy fill in ascending key order (stable by item index): an item is
 * packed when its weight fits the remaining capacity.  Fitness is the exact
 * packed value.  This decoder has no reference counterpart; it follows the
 * key-order convention of tsp_config.hpp:209-216. */
static float knap_greedy(const orc_problem* p, const float* keys, int32_t* sel)
{
    int n = p->n;
    int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    orc_argsort_keys(keys, n, order);
    int64_t rem = p->cap, val = 0;
    if (sel) memset(sel, 0, sizeof(int32_t) * (size_t)n);
    for (int r = 0; r < n; ++r) {
        int i = order[r];
        if ((int64_t)p->w[i] <= rem) {
            rem -= p->w[i];
            val += p->v[i];
            if (sel) sel[i] = 1;
        }
    }
    free(order);
    return (float)val;
}

/* TSPJ (tspj_config.hpp:555-642): city keys [0,n) and job keys [n,2n) are
 * argsorted separately into 1-based ids; the tour starts and ends at depot 0;
 * the makespan is max(total travel, max over stops of arrival + job time). */
static float tspj_fitness(const orc_problem* p, const float* keys, int32_t* dec)
{
    int n = p->n, dim = n + 1;
    int32_t* tour = (int32_t*)malloc(sizeof(int32_t) * (size_t)(2 * n + 1));
    int32_t* jobs = tour + n;
    orc_argsort_keys(keys, n, tour);
    orc_argsort_keys(keys + n, n, jobs);
    int64_t t = 0, best = 0;
    int cur = 0;
    for (int i = 0; i < n; ++i) {
        int city = tour[i] + 1, job = jobs[i] + 1;
        t += fx_term(p->a[(size_t)cur * dim + city], p->fx0);
        int64_t done = t + fx_term(p->b[(size_t)city * dim + job], p->fx0);
        if (done > best) best = done;
        cur = city;
    }
    int64_t total = t + fx_term(p->a[(size_t)cur * dim + 0], p->fx0);
    if (total > best) best = total;
    if (dec) {
        for (int i = 0; i < n; ++i) {
            dec[i] = tour[i] + 1;
            dec[n + i] = jobs[i] + 1;
        }
    }
    free(tour);
    return fx_to_float(best, p->fx0);
}

/* Bi-objective TSP (multi_tsp_config.hpp:54-103): cities 1..n in key order
 * between two visits of depot 0, measured in two independent metrics. */
static void bitsp_objectives(const orc_problem* p, const float* keys, float* obj, int32_t* perm)
{
    int n = p->n, dim = n + 1;
    orc_argsort_keys(keys, n, perm);
    for (int m = 0; m < 2; ++m) {
        const float* src = m ? p->b : p->a;
        int shift = m ? p->fx1 : p->fx0;
        int64_t acc = 0;
        int cur = 0;
        for (int i = 0; i <= n; ++i) {
            int nxt = i < n ? perm[i] + 1 : 0;
            float d = p->kind == ORC_BITSP_COORDS ? orc_tsp_edge_coords(src, cur, nxt)
                                                  : src[(size_t)cur * dim + nxt];
            acc += fx_term(d, shift);
            cur = nxt;
        }
        obj[m] = fx_to_float(acc, shift);
    }
}

/* Deterministic cosine: y = k pi/2 + r with k = rint(y 2/pi) and pi/2 split
 * into a 33-bit head (k * head exact for |k| < 2^20) and a tail (fdlibm's
 * rem_pio2 constants); Taylor series of cos / sin in r to degree 16 / 17;
 * separately rounded FP64 operations throughout. */
float orc_cos_det(float y)
{
    const double two_over_pi = 0.63661977236758134308;
    const double pio2_hi = 1.57079632673412561417e+00;
    const double pio2_lo = 6.07710050650619224932e-11;
    double yd = (double)y;
    double k = rint(yd * two_over_pi);
    double r = (yd - k * pio2_hi) - k * pio2_lo;
    double z = r * r;
    double c = 1.0 / 20922789888000.0;
    c = c * z - 1.0 / 87178291200.0;
    c = c * z + 1.0 / 479001600.0;
    c = c * z - 1.0 / 3628800.0;
    c = c * z + 1.0 / 40320.0;
    c = c * z - 1.0 / 720.0;
    c = c * z + 1.0 / 24.0;
    c = c * z - 0.5;
    c = c * z + 1.0;
    double sn = 1.0 / 355687428096000.0;
    sn = sn * z - 1.0 / 1307674368000.0;
    sn = sn * z + 1.0 / 6227020800.0;
    sn = sn * z - 1.0 / 39916800.0;
    sn = sn * z + 1.0 / 362880.0;
    sn = sn * z - 1.0 / 5040.0;
    sn = sn * z + 1.0 / 120.0;
    sn = sn * z - 1.0 / 6.0;
    sn = (sn * z) * r + r;
    int q = (int)((long long)k & 3);
    double v = q == 0 ? c : (q == 1 ? -sn : (q == 2 ? -c : sn));
    return (float)v;
}

static float cont_var(float gene, float lo, float hi)
{
    float w = hi - lo;
    float t = gene * w;
    return t + lo;
}

/* Sphere / Rastrigin / Rosenbrock over x_i = key_i (hi - lo) + lo
 * (continuous_opt_gpu_config.hpp:116-160); Rastrigin's constant 10 n is
 * added exactly.  dec holds the FP32 bits of each x_i. */
static float continuous_fitness(const orc_problem* p, const float* keys, int32_t* dec)
{
    int n = p->n, fn = p->ip0;
    int64_t acc = 0;
    for (int i = 0; i < n; ++i) {
        float x = cont_var(keys[i], p->f0, p->f1);
        if (dec) memcpy(&dec[i], &x, sizeof(float));
        float t;
        if (fn == ORC_SPHERE) {
            t = x * x;
        } else if (fn == ORC_RASTRIGIN) {
            float arg = 6.2831854820251465f * x;
            float c = orc_cos_det(arg);
            float xx = x * x;
            float tc = 10.0f * c;
            t = xx - tc;
        } else {
            if (i + 1 >= n) continue;
            float xn = cont_var(keys[i + 1], p->f0, p->f1);
            float xx = x * x;
            float t1 = xn - xx;
            float t2 = 1.0f - x;
            float a = 100.0f * t1;
            a = a * t1;
            float b = t2 * t2;
            t = a + b;
        }
        acc += fx_term(t, p->fx0);
    }
    if (fn == ORC_RASTRIGIN) acc += fx_term(10.0f * (float)n, p->fx0);
    return fx_to_float(acc, p->fx0);
}

static int job_machine(float gene, int M)
{
    float v = gene * (float)M;
    if (!(v >= 0.0f)) return 0;
    if (v >= (float)M) return M - 1;
    int m = (int)v;
    return m < M ? m : M - 1;
}

/* Makespan of the machine assignment int(key M) (job_scheduling_gpu_config
 * .hpp:138-158): largest machine load, empty machines counting 0. */
static float jobshop_fitness(const orc_problem* p, const float* keys, int32_t* dec)
{
    int n = p->n, M = p->ip0;
    int64_t* load = (int64_t*)calloc((size_t)M, sizeof(int64_t));
    for (int i = 0; i < n; ++i) {
        int m = job_machine(keys[i], M);
        load[m] += fx_term(p->a[i], p->fx0);
        if (dec) dec[i] = m;
    }
    int64_t best = 0;
    for (int m = 0; m < M; ++m)
        if (load[m] > best) best = load[m];
    free(load);
    return fx_to_float(best, p->fx0);
}

/* ZDT1 (examples/zdt1.cu:43-63): f1 = x1, g = 1 + 9 sum(x2..xn) / (n - 1)
 * with the key sum exact, f2 = g (1 - sqrt(f1 / g)). */
static void zdt1_objectives(const orc_problem* p, const float* keys, float* obj, int32_t* dec)
{
    int n = p->n;
    int64_t acc = 0;
    for (int i = 1; i < n; ++i) acc += fx_term(keys[i], p->fx0);
    if (dec)
        for (int i = 0; i < n; ++i) memcpy(&dec[i], &keys[i], sizeof(float));
    float f1 = keys[0];
    float g = 1.0f;
    if (n > 1) {
        double sum = (double)acc * ldexp(1.0, -p->fx0);
        g = (float)(1.0 + 9.0 * sum / (double)(n - 1));
    }
    float q = f1 / g;
    float r = sqrtf(q);
    obj[0] = f1;
    obj[1] = (float)((double)g * (1.0 - (double)r));
}

/* Multi-day routes (vrp_blazer_config.hpp:192-296), the reference's loop as
 * written: stops in key order join the current day while it has fewer than
 * ip0 stops and cum + leg + return (FP32) <= d0 (FP64); a stop that fits no
 * empty day is skipped and lands on the unvisited list twice (once in the
 * loop, once in the final sweep over unvisited stops).  Fitness = day
 * drive totals + 1000 per unvisited entry.  dec: day of each stop (0 =
 * skipped). */
static float vrp_days_fitness(const orc_problem* p, const float* keys, int32_t* dec)
{
    int n = p->n, dim = n + 1;
    const float* T = p->a;
    int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    orc_argsort_keys(keys, n, order);
    if (dec) memset(dec, 0, sizeof(int32_t) * (size_t)n);
    int64_t acc = 0;
    int skipped = 0, day = 0, idx = 0;
    while (idx < n) {
        int cur = 0, today = 0;
        float cum = 0.0f;
        while (idx < n && today < p->ip0) {
            int st = order[idx] + 1;
            float tts = T[(size_t)cur * dim + st];
            float ttd = T[(size_t)st * dim];
            float tot = cum + tts;
            tot = tot + ttd;
            if ((double)tot <= p->d0) {
                cum = cum + tts;
                cur = st;
                ++today;
                ++idx;
                if (dec) dec[st - 1] = day + 1;
            } else {
                break;
            }
        }
        if (today > 0) {
            float total = cum + T[(size_t)cur * dim];
            acc += fx_term(total, p->fx0);
            ++day;
        }
        if (today == 0 && idx < n) {
            ++skipped;
            ++idx;
        }
    }
    acc += fx_term(2000.0f * (float)skipped, p->fx0);
    free(order);
    return fx_to_float(acc, p->fx0);
}

/* Bi-objective CVRP (multi_objective_vrp_config.hpp:208-300), as written:
 * keys biased by i 1e-5 + hash 1e-5 with hash = running FP32 sum of
 * key_i (i + 1), sorted (ties by customer), split greedily by capacity (a
 * route always takes its first customer).  Objectives: total distance +
 * 100 per unit of overload; longest route; both + 500 per route beyond
 * ip0 vehicles.  dec: route of each customer (1-based). */
static void movrp_objectives(const orc_problem* p, const float* keys, float* obj, int32_t* dec)
{
    int n = p->n;
    float h = 0.0f;
    for (int i = 0; i < n; ++i) {
        float t = keys[i] * (float)(i + 1);
        h = h + t;
    }
    float* vals = (float*)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) {
        float b0 = (float)i * 0.00001f;
        float b1 = h * 0.00001f;
        float bias = b0 + b1;
        vals[i] = keys[i] + bias;
    }
    orc_argsort_keys(vals, n, order);
    int64_t total = 0, pen = 0, longest = 0, rd = 0, load = 0;
    int routes = 0, last = 0, nonempty = 0;
    for (int r = 0; r <= n; ++r) {
        int c = r < n ? order[r] + 1 : 0;
        int64_t d = r < n ? p->w[c] : 0;
        if (nonempty && (r == n || load + d > p->cap)) {
            rd += fx_term(orc_tsp_edge_coords(p->a, last, 0), p->fx0);
            total += rd;
            if (rd > longest) longest = rd;
            if (load > p->cap) pen += fx_term((float)(load - p->cap) * 100.0f, p->fx0);
            ++routes;
            rd = 0;
            load = 0;
            last = 0;
            nonempty = 0;
        }
        if (r == n) break;
        rd += fx_term(orc_tsp_edge_coords(p->a, last, c), p->fx0);
        load += d;
        last = c;
        nonempty = 1;
        if (dec) dec[c - 1] = routes + 1;
    }
    int64_t veh = routes > p->ip0 ? fx_term((float)(routes - p->ip0) * 500.0f, p->fx0) : 0;
    obj[0] = fx_to_float(total + pen + veh, p->fx0);
    obj[1] = fx_to_float(longest + veh, p->fx0);
    free(vals);
    free(order);
}

void orc_evaluate(const orc_problem* p, const float* keys, int count,
                  float* fitness, int32_t* decoded)
{
    int L = orc_chrom_len(p), D = orc_decoded_len(p), M = orc_num_obj(p);
    int32_t* scratch = (int32_t*)malloc(sizeof(int32_t) * (size_t)(D > 0 ? D : 1));
    for (int r = 0; r < count; ++r) {
        const float* k = keys + (size_t)r * L;
        int32_t* dec = decoded ? decoded + (size_t)r * D : scratch;
        switch (p->kind) {
        case ORC_TSP_COORDS:
        case ORC_TSP_MATRIX: fitness[r] = tsp_fitness(p, k, dec); break;
        case ORC_KNAP_THRESHOLD: fitness[r] = knap_threshold(p, k, dec); break;
        case ORC_KNAP_GREEDY: fitness[r] = knap_greedy(p, k, dec); break;
        case ORC_TSPJ: fitness[r] = tspj_fitness(p, k, dec); break;
        case ORC_BITSP_COORDS:
        case ORC_BITSP_MATRIX: bitsp_objectives(p, k, fitness + (size_t)r * M, dec); break;
        case ORC_CONTINUOUS: fitness[r] = continuous_fitness(p, k, dec); break;
        case ORC_JOBSHOP: fitness[r] = jobshop_fitness(p, k, dec); break;
        case ORC_ZDT1: zdt1_objectives(p, k, fitness + (size_t)r * M, dec); break;
        case ORC_VRP_DAYS: fitness[r] = vrp_days_fitness(p, k, dec); break;
        case ORC_MOVRP: movrp_objectives(p, k, fitness + (size_t)r * M, dec); break;
        default: fitness[r] = 0.0f; break;
        }
    }
    free(scratch);
}

/* ------------------------------------------------------------ partition -- */

/* Order-preserving map of FP32 fitness to u32 (ascending = better first);
 * -0 equals +0, NaN sorts last in both directions. */
static uint32_t fitness_order_key(float f, int maximize)
{
    if (f != f) return 0xFFFFFFFFu;
    if (f == 0.0f) f = 0.0f;
    uint32_t u;
    memcpy(&u, &f, 4);
    u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    return maximize ? ~u : u;
}

typedef struct { const uint32_t* k; } part_ctx;

static int part_less(int a, int b, const void* ctx)
{
    const part_ctx* c = (const part_ctx*)ctx;
    return c->k[a] < c->k[b];
}

/* Stable sort by fitness within each segment (population.hpp:660-675 with
 * the config comparator: '<' minimises, '>' maximises). */
void orc_partition(const float* fitness, int segments, int seg_len, int maximize,
                   int32_t* perm)
{
    uint32_t* k = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(seg_len > 0 ? seg_len : 1));
    for (int s = 0; s < segments; ++s) {
        for (int i = 0; i < seg_len; ++i)
            k[i] = fitness_order_key(fitness[(size_t)s * seg_len + i], maximize);
        part_ctx c = {k};
        int32_t* p = perm + (size_t)s * seg_len;
        for (int i = 0; i < seg_len; ++i) p[i] = i;
        merge_sort_idx(p, seg_len, part_less, &c);
    }
    free(k);
}

/* ---------------------------------------------------------------- BRKGA -- */

void orc_init_keys(uint64_t seed, uint32_t stream, int slots, int len, float* keys)
{
    for (int s = 0; s < slots; ++s)
        for (int j = 0; j < len; ++j)
            keys[(size_t)s * len + j] =
                orc_u32_to_unit(orc_philox_word(seed, stream, 0, (uint32_t)s, ORC_TAG_INIT, (uint32_t)j));
}

/* Generation 0: P random-key rows per island, evaluated and partitioned
 * (population.hpp:107-124). */
void orc_brkga_init(const orc_problem* p, const orc_params* q,
                    float* keys, float* fitness, int32_t* perm)
{
    int L = orc_chrom_len(p);
    int Pi = q->pop / q->islands;
    for (int i = 0; i < q->islands; ++i)
        orc_init_keys(q->seed, (uint32_t)i, Pi, L, keys + (size_t)i * Pi * L);
    orc_evaluate(p, keys, q->pop, fitness, NULL);
    orc_partition(fitness, q->islands, Pi, q->maximize, perm);
}

/* One generation step t (Philox generation counter t+1), per island:
 *   migration slots [Pi-k, Pi)  <- island (i-1)'s top k rows when
 *                                  t > 0 and t % interval == 0
 *                                  (solver.hpp:465-467, 517-546);
 *   elite slots    [0, Ei)      <- rows perm[s] (cuda_kernels.cuh:107-113);
 *   mutant slots   [Pi-Mi, Pi)  <- fresh keys (cuda_kernels.cuh:114-119);
 *   offspring                   <- elite parent perm[u0 % Ei], non-elite
 *                                  parent perm[Ei + u1 % (Pi-Ei)], gene from
 *                                  the elite parent iff coin < rho
 *                                  (cuda_kernels.cuh:120-135).
 * The new population is then evaluated and partitioned. */
void orc_brkga_step(const orc_problem* p, const orc_params* q, int t,
                    const float* keys_in, const int32_t* perm_in,
                    float* keys_out, float* fitness_out, int32_t* perm_out)
{
    int L = orc_chrom_len(p);
    int I = q->islands;
    int Pi = q->pop / I, Ei = q->elite / I, Mi = q->mutants / I;
    int k = q->mig_count;
    uint32_t g = (uint32_t)t + 1u;
    float rho = (float)q->rho;
    int migrate = I > 1 && q->mig_interval > 0 && k > 0 && t > 0 && (t % q->mig_interval) == 0;

    for (int i = 0; i < I; ++i) {
        const int32_t* perm = perm_in + (size_t)i * Pi;
        for (int s = 0; s < Pi; ++s) {
            float* out = keys_out + ((size_t)i * Pi + s) * L;
            if (migrate && s >= Pi - k) {
                int src = (i + I - 1) % I;
                int row = src * Pi + perm_in[(size_t)src * Pi + (s - (Pi - k))];
                memcpy(out, keys_in + (size_t)row * L, sizeof(float) * (size_t)L);
            } else if (s < Ei) {
                int row = i * Pi + perm[s];
                memcpy(out, keys_in + (size_t)row * L, sizeof(float) * (size_t)L);
            } else if (s >= Pi - Mi) {
                for (int j = 0; j < L; ++j)
                    out[j] = orc_u32_to_unit(orc_philox_word(q->seed, (uint32_t)i, g, (uint32_t)s,
                                                             ORC_TAG_MUTANT, (uint32_t)j));
            } else {
                uint32_t u0 = orc_philox_word(q->seed, (uint32_t)i, g, (uint32_t)s, ORC_TAG_PARENTS, 0);
                uint32_t u1 = orc_philox_word(q->seed, (uint32_t)i, g, (uint32_t)s, ORC_TAG_PARENTS, 1);
                int ep = i * Pi + perm[u0 % (uint32_t)Ei];
                int np = i * Pi + perm[Ei + (int)(u1 % (uint32_t)(Pi - Ei))];
                const float* pe = keys_in + (size_t)ep * L;
                const float* pn = keys_in + (size_t)np * L;
                for (int j = 0; j < L; ++j) {
                    float coin = orc_u32_to_unit(orc_philox_word(q->seed, (uint32_t)i, g, (uint32_t)s,
                                                                 ORC_TAG_COINS, (uint32_t)j));
                    out[j] = coin < rho ? pe[j] : pn[j];
                }
            }
        }
    }
    orc_evaluate(p, keys_out, q->pop, fitness_out, NULL);
    orc_partition(fitness_out, I, Pi, q->maximize, perm_out);
    orc_tspj_local_search(p, q, t + 1, keys_out, fitness_out, perm_out);
}

/* TSPJ local search (the hook of tspj_config.hpp:77-135, applied as
 * solver.hpp:458-462 does: generations g > 0 with g % ls_interval == 0, the
 * first min(ls_count, Ei) individuals of every island's rank order).  Each
 * of ls_moves random 2-opt moves draws positions x = u0 % n, y = u1 % n
 * (Philox words 2m, 2m+1 of (island, g, rank, LSEARCH)), orders them and,
 * unless y - x < 2, reverses the decoded city tour between them.  The
 * individual's city keys, sorted ascending, are then handed to the new tour
 * in order (job keys unchanged), and the candidate replaces the individual
 * only if its makespan is strictly smaller (cuda_kernels.cuh:224-246).  The
 * reference reverses key values by position instead (cuda_kernels.cuh:
 * 208-216), which does not reverse the decoded tour; that is not
 * reproduced.  Every island is partitioned again afterwards. */
void orc_tspj_local_search(const orc_problem* p, const orc_params* q, int g,
                           float* keys, float* fitness, int32_t* perm)
{
    if (p->kind != ORC_TSPJ || q->ls_interval <= 0 || q->ls_count <= 0 || g <= 0 || g % q->ls_interval != 0)
        return;
    int n = p->n, L = 2 * n, I = q->islands;
    int Pi = q->pop / I, Ei = q->elite / I;
    int k = q->ls_count < Ei ? q->ls_count : Ei;
    int32_t* tour = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    float* vals = (float*)malloc(sizeof(float) * (size_t)n);
    float* cand = (float*)malloc(sizeof(float) * (size_t)L);
    for (int i = 0; i < I; ++i) {
        for (int r = 0; r < k; ++r) {
            int row = i * Pi + perm[(size_t)i * Pi + r];
            float* key = keys + (size_t)row * L;
            orc_argsort_keys(key, n, tour);
            for (int pos = 0; pos < n; ++pos) vals[pos] = key[tour[pos]];
            for (int m = 0; m < q->ls_moves; ++m) {
                uint32_t u0 = orc_philox_word(q->seed, (uint32_t)i, (uint32_t)g, (uint32_t)r, ORC_TAG_LSEARCH,
                                              (uint32_t)(2 * m));
                uint32_t u1 = orc_philox_word(q->seed, (uint32_t)i, (uint32_t)g, (uint32_t)r, ORC_TAG_LSEARCH,
                                              (uint32_t)(2 * m + 1));
                int x = (int)(u0 % (uint32_t)n), y = (int)(u1 % (uint32_t)n);
                if (x > y) { int tmp = x; x = y; y = tmp; }
                if (y - x < 2) continue;
                while (x < y) { int tmp = tour[x]; tour[x] = tour[y]; tour[y] = tmp; ++x; --y; }
            }
            memcpy(cand, key, sizeof(float) * (size_t)L);
            for (int pos = 0; pos < n; ++pos) cand[tour[pos]] = vals[pos];
            float f = tspj_fitness(p, cand, NULL);
            if (f < fitness[row]) {
                memcpy(key, cand, sizeof(float) * (size_t)n);
                fitness[row] = f;
            }
        }
    }
    free(tour); free(vals); free(cand);
    orc_partition(fitness, I, Pi, q->maximize, perm);
}

/* Init + `generations` steps; returns the best fitness over islands
 * (the first island wins ties). */
double orc_brkga_run(const orc_problem* p, const orc_params* q, int generations,
                     float* keys, float* fitness, int32_t* perm)
{
    int L = orc_chrom_len(p);
    size_t nk = (size_t)q->pop * L;
    float* k2 = (float*)malloc(sizeof(float) * nk);
    float* f2 = (float*)malloc(sizeof(float) * (size_t)q->pop);
    int32_t* p2 = (int32_t*)malloc(sizeof(int32_t) * (size_t)q->pop);
    orc_brkga_init(p, q, keys, fitness, perm);
    for (int t = 0; t < generations; ++t) {
        orc_brkga_step(p, q, t, keys, perm, k2, f2, p2);
        memcpy(keys, k2, sizeof(float) * nk);
        memcpy(fitness, f2, sizeof(float) * (size_t)q->pop);
        memcpy(perm, p2, sizeof(int32_t) * (size_t)q->pop);
    }
    free(k2); free(f2); free(p2);
    int Pi = q->pop / q->islands;
    double best = fitness[perm[0]];
    for (int i = 1; i < q->islands; ++i) {
        double f = fitness[(size_t)i * Pi + perm[(size_t)i * Pi]];
        if (q->maximize ? f > best : f < best) best = f;
    }
    return best;
}

/* -------------------------------------------------------------- NSGA-II -- */

/* Minimisation dominance (individual.hpp:99-114). */
static int dominates(const float* a, const float* b, int nobj)
{
    int better = 0;
    for (int m = 0; m < nobj; ++m) {
        if (a[m] > b[m]) return 0;
        if (a[m] < b[m]) better = 1;
    }
    return better;
}

/* Front index by peeling (population.hpp:334-369, 409-440). */
void orc_nds_ranks(const float* obj, int count, int nobj, int32_t* rank)
{
    int32_t* cnt = (int32_t*)calloc((size_t)(count > 0 ? count : 1), sizeof(int32_t));
    int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * (size_t)(count > 0 ? count : 1));
    int32_t* nxt = (int32_t*)malloc(sizeof(int32_t) * (size_t)(count > 0 ? count : 1));
    for (int i = 0; i < count; ++i) {
        rank[i] = -1;
        for (int j = 0; j < count; ++j)
            if (j != i && dominates(obj + (size_t)j * nobj, obj + (size_t)i * nobj, nobj)) cnt[i]++;
    }
    int nc = 0;
    for (int i = 0; i < count; ++i)
        if (cnt[i] == 0) { rank[i] = 0; cur[nc++] = i; }
    int r = 0;
    while (nc > 0) {
        int nn = 0;
        for (int a = 0; a < nc; ++a) {
            int pidx = cur[a];
            for (int qi = 0; qi < count; ++qi) {
                if (qi != pidx && dominates(obj + (size_t)pidx * nobj, obj + (size_t)qi * nobj, nobj)) {
                    if (--cnt[qi] == 0) { rank[qi] = r + 1; nxt[nn++] = qi; }
                }
            }
        }
        memcpy(cur, nxt, sizeof(int32_t) * (size_t)nn);
        nc = nn;
        ++r;
    }
    free(cnt); free(cur); free(nxt);
}

typedef struct { const float* obj; int nobj; int m; } obj_ctx;

static int obj_less(int a, int b, const void* ctx)
{
    const obj_ctx* c = (const obj_ctx*)ctx;
    return c->obj[(size_t)a * c->nobj + c->m] < c->obj[(size_t)b * c->nobj + c->m];
}

typedef struct { const float* crowd; } crowd_ctx;

static int crowd_greater(int a, int b, const void* ctx)
{
    const crowd_ctx* c = (const crowd_ctx*)ctx;
    return c->crowd[a] > c->crowd[b];
}

static int collect_front(const int32_t* rank, int count, int r, int32_t* out)
{
    int n = 0;
    for (int i = 0; i < count; ++i)
        if (rank[i] == r) out[n++] = i;
    return n;
}

static int max_rank(const int32_t* rank, int count)
{
    int m = -1;
    for (int i = 0; i < count; ++i)
        if (rank[i] > m) m = rank[i];
    return m;
}

/* Crowding distance per front (population.hpp:453-489).  Each objective's
 * order is a stable sort of the ascending-index front by that objective. */
void orc_crowding(const float* obj, const int32_t* rank, int count, int nobj, float* crowd)
{
    int32_t* f = (int32_t*)malloc(sizeof(int32_t) * (size_t)(count > 0 ? count : 1));
    int mr = max_rank(rank, count);
    for (int r = 0; r <= mr; ++r) {
        int n = collect_front(rank, count, r, f);
        if (n == 0) continue;
        if (n <= 2) {
            for (int a = 0; a < n; ++a) crowd[f[a]] = INFINITY;
            continue;
        }
        for (int a = 0; a < n; ++a) crowd[f[a]] = 0.0f;
        for (int m = 0; m < nobj; ++m) {
            collect_front(rank, count, r, f);
            obj_ctx c = {obj, nobj, m};
            merge_sort_idx(f, n, obj_less, &c);
            crowd[f[0]] = INFINITY;
            crowd[f[n - 1]] = INFINITY;
            float range = obj[(size_t)f[n - 1] * nobj + m] - obj[(size_t)f[0] * nobj + m];
            if (range > 0.0f) {
                for (int a = 1; a < n - 1; ++a) {
                    float d = (obj[(size_t)f[a + 1] * nobj + m] - obj[(size_t)f[a - 1] * nobj + m]) / range;
                    crowd[f[a]] += d;
                }
            }
        }
    }
    free(f);
}

/* Survivor selection (population.hpp:491-517): whole fronts in rank order
 * while they fit, then the overflowing front by crowding descending (ties by
 * index).  A whole front of more than two members keeps the order the last
 * crowding pass leaves (last objective ascending, ties by index); smaller
 * fronts keep index order. */
void orc_nsga2_select(const float* obj, const int32_t* rank, const float* crowd,
                      int count, int nobj, int keep, int32_t* order)
{
    int32_t* f = (int32_t*)malloc(sizeof(int32_t) * (size_t)(count > 0 ? count : 1));
    int mr = max_rank(rank, count);
    int got = 0;
    for (int r = 0; r <= mr && got < keep; ++r) {
        int n = collect_front(rank, count, r, f);
        if (n == 0) continue;
        if (got + n <= keep) {
            if (n > 2) {
                obj_ctx c = {obj, nobj, nobj - 1};
                merge_sort_idx(f, n, obj_less, &c);
            }
            for (int a = 0; a < n; ++a) order[got++] = f[a];
        } else {
            crowd_ctx c = {crowd};
            merge_sort_idx(f, n, crowd_greater, &c);
            for (int a = 0; got < keep; ++a) order[got++] = f[a];
        }
    }
    free(f);
}

void orc_nsga2_init(const orc_problem* p, const orc_params* q,
                    float* keys, float* obj, int32_t* rank, float* crowd)
{
    int L = orc_chrom_len(p), M = orc_num_obj(p);
    orc_init_keys(q->seed, 0, q->pop, L, keys);
    orc_evaluate(p, keys, q->pop, obj, NULL);
    orc_nds_ranks(obj, q->pop, M, rank);
    orc_crowding(obj, rank, q->pop, M, crowd);
}

/* Binary tournament (solver.hpp:1042-1056): lower rank wins, then larger
 * crowding; remaining ties go to the second draw. */
static int tournament(const int32_t* rank, const float* crowd, int a, int b)
{
    if (rank[a] < rank[b]) return a;
    if (rank[b] < rank[a]) return b;
    if (crowd[a] > crowd[b]) return a;
    return b;
}

/* One NSGA-II generation (solver.hpp:751-817): N offspring by two
 * tournaments, uniform crossover (coin < rho takes parent 1,
 * solver.hpp:1058-1074) and per-gene resampling with probability 5/L
 * (solver.hpp:1077-1091); then parents + offspring are ranked, crowded and
 * truncated back to N. */
/* ensure_minimum_diversity, first pass (population.hpp:132-157): the members
 * in order; a member is kept when its Euclidean distance in objective space
 * to every member kept before it, in FP64 from the FP32 objectives with
 * std::sqrt as the reference computes it, is not below 0.001. */
int orc_diversity_unique(const float* obj, int count, int nobj, int32_t* kept)
{
    int u = 0;
    for (int i = 0; i < count; ++i) {
        int unique = 1;
        for (int k = 0; k < u && unique; ++k) {
            const float* a = obj + (size_t)i * nobj;
            const float* b = obj + (size_t)kept[k] * nobj;
            double dist = 0.0;
            for (int m = 0; m < nobj; ++m) {
                double diff = (double)a[m] - (double)b[m];
                dist += diff * diff;
            }
            dist = sqrt(dist);
            if (dist < 0.001) unique = 0;
        }
        if (unique) kept[u++] = i;
    }
    return u;
}

/* ensure_minimum_diversity (population.hpp:130-182) on the N members that
 * NSGA-II step t selected, which the reference runs when its generation
 * counter (t) is a multiple of 25 (solver.hpp:812-814).  When fewer than 70%
 * are unique, the unique members stay, in order, in rows [0, U); rows
 * [U, N) are new random individuals (keys from Philox stream (0, t + 1,
 * row, DIVERSITY), in place of the reference's mt19937 randomize); the new
 * rows are evaluated and all N members ranked and crowded again
 * (population.hpp:174-176). */
int orc_nsga2_diversity(const orc_problem* p, const orc_params* q, int t,
                        float* keys, float* obj, int32_t* rank, float* crowd)
{
    int L = orc_chrom_len(p), M = orc_num_obj(p), N = q->pop;
    int32_t* kept = (int32_t*)malloc(sizeof(int32_t) * (size_t)N);
    int u = orc_diversity_unique(obj, N, M, kept);
    double ratio = (double)u / (double)N;
    int ran = 0;
    if (ratio < 0.70) {
        ran = 1;
        for (int k = 0; k < u; ++k) {
            /* kept[k] >= k: moving forward in order never overwrites a row
             * still to be moved */
            if (kept[k] == k) continue;
            memmove(keys + (size_t)k * L, keys + (size_t)kept[k] * L, sizeof(float) * (size_t)L);
            memmove(obj + (size_t)k * M, obj + (size_t)kept[k] * M, sizeof(float) * (size_t)M);
        }
        uint32_t g = (uint32_t)t + 1u;
        for (int r = u; r < N; ++r)
            for (int j = 0; j < L; ++j)
                keys[(size_t)r * L + j] =
                    orc_u32_to_unit(orc_philox_word(q->seed, 0, g, (uint32_t)r, ORC_TAG_DIVERSITY, (uint32_t)j));
        orc_evaluate(p, keys + (size_t)u * L, N - u, obj + (size_t)u * M, NULL);
        orc_nds_ranks(obj, N, M, rank);
        orc_crowding(obj, rank, N, M, crowd);
    }
    free(kept);
    return ran;
}

int orc_nsga2_step(const orc_problem* p, const orc_params* q, int t,
                   const float* keys_in, const float* obj_in,
                   const int32_t* rank_in, const float* crowd_in,
                   float* keys_out, float* obj_out, int32_t* rank_out, float* crowd_out)
{
    int L = orc_chrom_len(p), M = orc_num_obj(p), N = q->pop;
    uint32_t g = (uint32_t)t + 1u;
    float rho = (float)q->rho;
    float mut = (float)(5.0 / (double)L);
    float* ck = (float*)malloc(sizeof(float) * (size_t)2 * N * L);
    float* co = (float*)malloc(sizeof(float) * (size_t)2 * N * M);
    int32_t* cr = (int32_t*)malloc(sizeof(int32_t) * (size_t)2 * N);
    float* cc = (float*)malloc(sizeof(float) * (size_t)2 * N);
    int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (size_t)N);

    memcpy(ck, keys_in, sizeof(float) * (size_t)N * L);
    memcpy(co, obj_in, sizeof(float) * (size_t)N * M);
    for (int c = 0; c < N; ++c) {
        uint32_t r[4];
        for (int w = 0; w < 4; ++w)
            r[w] = orc_philox_word(q->seed, 0, g, (uint32_t)c, ORC_TAG_PARENTS, (uint32_t)w);
        int p1 = tournament(rank_in, crowd_in, (int)(r[0] % (uint32_t)N), (int)(r[1] % (uint32_t)N));
        int p2 = tournament(rank_in, crowd_in, (int)(r[2] % (uint32_t)N), (int)(r[3] % (uint32_t)N));
        float* out = ck + (size_t)(N + c) * L;
        for (int j = 0; j < L; ++j) {
            float coin = orc_u32_to_unit(orc_philox_word(q->seed, 0, g, (uint32_t)c, ORC_TAG_COINS, (uint32_t)j));
            float x = coin < rho ? keys_in[(size_t)p1 * L + j] : keys_in[(size_t)p2 * L + j];
            float mc = orc_u32_to_unit(orc_philox_word(q->seed, 0, g, (uint32_t)c, ORC_TAG_MUTCOIN, (uint32_t)j));
            if (mc < mut)
                x = orc_u32_to_unit(orc_philox_word(q->seed, 0, g, (uint32_t)c, ORC_TAG_MUTANT, (uint32_t)j));
            out[j] = x;
        }
    }
    orc_evaluate(p, ck + (size_t)N * L, N, co + (size_t)N * M, NULL);
    orc_nds_ranks(co, 2 * N, M, cr);
    orc_crowding(co, cr, 2 * N, M, cc);
    orc_nsga2_select(co, cr, cc, 2 * N, M, N, order);
    for (int i = 0; i < N; ++i) {
        int s = order[i];
        memcpy(keys_out + (size_t)i * L, ck + (size_t)s * L, sizeof(float) * (size_t)L);
        memcpy(obj_out + (size_t)i * M, co + (size_t)s * M, sizeof(float) * (size_t)M);
        rank_out[i] = cr[s];
        crowd_out[i] = cc[s];
    }
    free(ck); free(co); free(cr); free(cc); free(order);
    if (q->div_interval > 0 && t % q->div_interval == 0)
        return orc_nsga2_diversity(p, q, t, keys_out, obj_out, rank_out, crowd_out);
    return 0;
}
