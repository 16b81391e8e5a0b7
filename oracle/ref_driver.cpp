// ref_driver.cpp - C entry points over the reference's own CUDA-free headers.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile (target `ref`) from
// the reference sources where they lie (REF_DIR, default the read-only
// reference checkout) into oracle/_ref/libref.so.  Tests use it to pin the
// oracle's restatement of:
//   KnapsackConfig<float>::calculate_knapsack_fitness  configs/knapsack_config.hpp:43-62
//   KnapsackConfig<float>::decode_to_solution          configs/knapsack_config.hpp:64-75
//   MultiTSPConfig<float> objectives + decode_tour     configs/multi_tsp_config.hpp:54-103
//   Individual<float>::dominates                       core/individual.hpp:99-114
//   KnapsackConfig<float>::load_from_file              configs/knapsack_config.hpp:180-304
//   MultiTSPConfig<float>::load_from_file              configs/multi_tsp_config.hpp:165-269
//   VrpBlazerConfig<float> fitness + day routes        configs/vrp_blazer_config.hpp:192-296
//   MultiObjectiveVRPConfig<float> load + objectives   configs/multi_objective_vrp_config.hpp:69-300
// Nothing here is part of the engine.
#include "configs/knapsack_config.hpp"
#include "configs/multi_objective_vrp_config.hpp"
#include "configs/multi_tsp_config.hpp"
#include "configs/vrp_blazer_config.hpp"
#include "core/individual.hpp"

#include <cstdint>
#include <vector>

namespace {

Individual<float> make_individual(const float* keys, int n)
{
    Individual<float> ind(n);
    for (int i = 0; i < n; ++i) ind.chromosomes[0][i] = keys[i];
    return ind;
}

}  // namespace

extern "C" {

// Fitness and 0/1 selection of `count` key rows of length n.
void ref_knapsack_threshold(const float* w, const float* v, int n, float cap,
                            const float* keys, int count, float* fitness, int32_t* sel)
{
    std::vector<float> wv(w, w + n), vv(v, v + n);
    KnapsackConfig<float> cfg(wv, vv, cap);
    for (int r = 0; r < count; ++r) {
        auto ind = make_individual(keys + (size_t)r * n, n);
        fitness[r] = cfg.fitness_function(ind);
        auto dec = cfg.decoder(ind);
        for (int i = 0; i < n; ++i) sel[(size_t)r * n + i] = dec[0][i] > 0.5f ? 1 : 0;
    }
}

// Objectives [count][2] and 0-based city order [count][n] for matrices of
// dimension dim = n + 1 (row/column 0 is the depot).
void ref_multitsp(const float* dist, const float* time, int dim,
                  const float* keys, int count, float* obj, int32_t* perm)
{
    std::vector<std::vector<float>> d(dim, std::vector<float>(dim)), t(dim, std::vector<float>(dim));
    for (int i = 0; i < dim; ++i)
        for (int j = 0; j < dim; ++j) {
            d[i][j] = dist[(size_t)i * dim + j];
            t[i][j] = time[(size_t)i * dim + j];
        }
    MultiTSPConfig<float> cfg(d, t);
    int n = dim - 1;
    for (int r = 0; r < count; ++r) {
        auto ind = make_individual(keys + (size_t)r * n, n);
        obj[2 * r] = cfg.objective_functions[0](ind);
        obj[2 * r + 1] = cfg.objective_functions[1](ind);
        auto tour = cfg.decode_tour(ind);
        for (int i = 0; i < n; ++i) perm[(size_t)r * n + i] = tour[i];
    }
}

// out[i*count + j] = 1 if individual i dominates individual j.
void ref_dominates(const float* obj, int count, int nobj, int32_t* out)
{
    std::vector<Individual<float>> pop;
    pop.reserve(count);
    for (int i = 0; i < count; ++i) {
        Individual<float> ind(std::vector<int>{1}, nobj);
        for (int m = 0; m < nobj; ++m) ind.objectives[m] = obj[(size_t)i * nobj + m];
        pop.push_back(ind);
    }
    for (int i = 0; i < count; ++i)
        for (int j = 0; j < count; ++j) out[(size_t)i * count + j] = pop[i].dominates(pop[j]) ? 1 : 0;
}

// Items of a knapsack file as the reference loads it; returns the item
// count (at most max_items are copied).
int ref_knapsack_load(const char* path, float* w, float* v, int max_items, float* cap)
{
    auto cfg = KnapsackConfig<float>::load_from_file(path);
    const auto& ww = cfg->get_weights();
    const auto& vv = cfg->get_values();
    for (int i = 0; i < (int)ww.size() && i < max_items; ++i) {
        w[i] = ww[i];
        v[i] = vv[i];
    }
    *cap = cfg->get_capacity();
    return (int)ww.size();
}

// Objectives [rows][2] of key rows [rows][n] under the bi-objective TSP the
// reference builds from a TSPLIB file (distance matrix and the generated
// time matrix); returns n (cities after the depot).
int ref_multitsp_load_objectives(const char* path, const float* keys, int rows, float* obj)
{
    auto cfg = MultiTSPConfig<float>::load_from_file(path);
    const int n = cfg->get_num_cities();
    for (int r = 0; r < rows; ++r) {
        auto ind = make_individual(keys + (size_t)r * n, n);
        obj[2 * r] = cfg->objective_functions[0](ind);
        obj[2 * r + 1] = cfg->objective_functions[1](ind);
    }
    return n;
}

// VrpBlazerConfig<float> over time_matrix[dim][dim]: fitness and the day
// (1-based) of each stop, 0 for stops in no day.
void ref_vrp_days(const float* tm, int dim, int max_stops, double max_hours, const float* keys, int count,
                  float* fitness, int32_t* day)
{
    VrpBlazerConfig<float> cfg;
    cfg.max_stops_per_day = max_stops;
    cfg.max_drive_hours_per_day = max_hours;
    cfg.update_derived_params();
    std::vector<std::vector<float>> m(dim, std::vector<float>(dim));
    for (int i = 0; i < dim; ++i)
        for (int j = 0; j < dim; ++j) m[i][j] = tm[(size_t)i * dim + j];
    cfg.set_time_matrix(m);
    const int n = dim - 1;
    for (int r = 0; r < count; ++r) {
        const float* k = keys + (size_t)r * n;
        auto ind = make_individual(k, n);
        fitness[r] = cfg.calculate_fitness(ind);
        auto sol = cfg.decode_solution(k);
        for (int i = 0; i < n; ++i) day[(size_t)r * n + i] = 0;
        for (const auto& d : sol.days)
            for (int st : d.stop_indices) day[(size_t)r * n + st - 1] = d.day_number;
    }
}

// MultiObjectiveVRPConfig<float>::load_from_file(path); returns the number
// of customers, and the vehicle count and capacity it parsed.
int ref_movrp_info(const char* path, int* vehicles, int* capacity)
{
    auto cfg = MultiObjectiveVRPConfig<float>::load_from_file(path);
    *vehicles = cfg->get_num_vehicles();
    *capacity = cfg->get_vehicle_capacity();
    return cfg->get_num_customers();
}

// Objectives [count][2] and the route (1-based) of each customer.
void ref_movrp(const char* path, const float* keys, int count, float* obj, int32_t* route)
{
    auto cfg = MultiObjectiveVRPConfig<float>::load_from_file(path);
    const int n = cfg->get_num_customers();
    for (int r = 0; r < count; ++r) {
        const float* k = keys + (size_t)r * n;
        cfg->decode_solution(k, obj + 2 * (size_t)r);
        auto routes = cfg->decode_routes(k);
        for (size_t ri = 0; ri < routes.size(); ++ri)
            for (int c : routes[ri].customers) route[(size_t)r * n + c - 1] = (int32_t)ri + 1;
    }
}

}  // extern "C"
