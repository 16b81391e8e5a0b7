// facade_test.cpp - drives the C++ facade (include/brkga/) the way a caller
// of the reference's Solver / configs would, on one GPU, and checks it
// against the C ABI it sits on.  Prints FACADE OK and exits 0 on success.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <numeric>
#include <set>
#include <vector>

#include "brkga/configs/continuous_opt_gpu_config.hpp"
#include "brkga/configs/job_scheduling_gpu_config.hpp"
#include "brkga/configs/knapsack_config.hpp"
#include "brkga/configs/multi_objective_vrp_config.hpp"
#include "brkga/configs/multi_tsp_config.hpp"
#include "brkga/configs/tsp_config.hpp"
#include "brkga/configs/tspj_config.hpp"
#include "brkga/configs/vrp_blazer_config.hpp"
#include "brkga/configs/zdt1_config.hpp"
#include "brkga/core/solver.hpp"

#include <cstdlib>
#include <fstream>

static int failures = 0;
#define EXPECT(cond, ...)                        \
    do {                                         \
        if (!(cond)) {                           \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
            std::printf(__VA_ARGS__);            \
            std::printf("\n");                   \
            ++failures;                          \
        }                                        \
    } while (0)

// Small deterministic generator for instance data (not the engine's RNG).
static unsigned lcg(unsigned& s)
{
    s = s * 1664525u + 1013904223u;
    return s >> 8;
}

static bool is_perm(const std::vector<int>& p, int base)
{
    std::vector<int> q(p);
    std::sort(q.begin(), q.end());
    for (std::size_t i = 0; i < q.size(); ++i)
        if (q[i] != (int)i + base) return false;
    return true;
}

// Same population through the raw C ABI: best fitness after G generations.
static float abi_best(BRKGAConfig<float>* cfg, uint64_t seed, int G)
{
    brkga_params p{};
    p.population_size = cfg->population_size;
    p.elite_size = cfg->elite_size;
    p.mutant_size = cfg->mutant_size;
    p.elite_prob = cfg->elite_prob;
    p.num_islands = 1;
    p.seed = seed;
    if (cfg->has_local_search()) {  // the facade's defaults (reference solver.hpp:166)
        p.ls_interval = 2;
        p.ls_individuals = 100;
        p.ls_moves = 5;
    }
    brkga_pop_t pop = nullptr;
    brkga_facade::check(brkga_population_create(cfg->native_problem(0), &p, &pop), "create");
    brkga_facade::check(brkga_population_initialize(pop, nullptr), "init");
    brkga_facade::check(brkga_evolve(pop, G, nullptr), "evolve");
    float f[2] = {0, 0};
    brkga_facade::check(brkga_get_best(pop, f, nullptr, nullptr, nullptr), "best");
    brkga_population_destroy(pop);
    return f[0];
}

static void test_tsp()
{
    unsigned s = 7;
    std::vector<std::pair<float, float>> xy;
    for (int i = 0; i < 60; ++i) xy.push_back({(float)(lcg(s) % 10000), (float)(lcg(s) % 10000)});
    auto cfg = TSPConfig<float>::from_coordinates(xy, "tsp60");
    cfg->population_size = 64;
    cfg->elite_size = 12;
    cfg->mutant_size = 6;
    cfg->max_generations = 20;
    TSPConfig<float>* raw = cfg.get();
    Solver<float> solver(std::move(cfg), false);
    solver.set_seed(99);
    solver.run();
    const Individual<float>& best = solver.get_best_individual();
    EXPECT(solver.get_current_generation() == 20, "generations %d", solver.get_current_generation());
    const float via_abi = abi_best(raw, 99, 20);
    EXPECT(best.fitness == via_abi, "TSP facade %.6f vs C ABI %.6f", best.fitness, via_abi);
    EXPECT(raw->fitness_function(best) == best.fitness, "fitness callback disagrees");
    const auto tour = raw->decode_tour_internal(best);
    EXPECT(is_perm(tour, 0), "decoded tour is not a permutation");
    double len = 0.0;
    for (int i = 0; i < 60; ++i) {
        const auto& a = xy[tour[i]];
        const auto& b = xy[tour[(i + 1) % 60]];
        len += std::sqrt((double)(a.first - b.first) * (a.first - b.first) +
                         (double)(a.second - b.second) * (a.second - b.second));
    }
    EXPECT(std::fabs(len - best.fitness) <= 1e-6 * len, "tour length %.6f vs %.6f", len, best.fitness);
    const auto& hist = solver.get_fitness_history();
    EXPECT(!hist.empty() && hist.back() <= hist.front(), "fitness history does not improve");

    // evaluate_population_gpu on caller-owned device rows
    const int P = 8, L = 60;
    std::vector<float> keys(P * L);
    for (auto& k : keys) k = (float)(lcg(s) & 0xFFFFFF) * (1.0f / 16777216.0f);
    float *dk = nullptr, *df = nullptr;
    EXPECT(hipMalloc(&dk, keys.size() * 4) == hipSuccess && hipMalloc(&df, P * 4) == hipSuccess, "hipMalloc");
    EXPECT(hipMemcpy(dk, keys.data(), keys.size() * 4, hipMemcpyHostToDevice) == hipSuccess, "H2D");
    raw->evaluate_population_gpu(dk, df, P, L);
    std::vector<float> fit(P);
    EXPECT(hipMemcpy(fit.data(), df, P * 4, hipMemcpyDeviceToHost) == hipSuccess, "D2H");
    for (int r = 0; r < P; ++r) {
        Individual<float> ind(std::vector<int>{L});
        for (int j = 0; j < L; ++j) ind.chromosomes[0][j] = keys[r * L + j];
        EXPECT(raw->calculate_tsp_fitness(ind) == fit[r], "row %d: host path vs evaluate_population_gpu", r);
    }
    (void)hipFree(dk);
    (void)hipFree(df);
}

static void test_knapsack(bool greedy)
{
    unsigned s = 11;
    std::vector<float> w, v;
    double tw = 0;
    for (int i = 0; i < 120; ++i) {
        w.push_back((float)(1 + lcg(s) % 100));
        v.push_back(w.back() + (float)(lcg(s) % 11));
        tw += w.back();
    }
    auto cfg = std::make_unique<KnapsackConfig<float>>(w, v, (float)std::floor(tw / 2), "knap120");
    cfg->set_greedy(greedy);
    cfg->population_size = 64;
    cfg->elite_size = 12;
    cfg->mutant_size = 6;
    cfg->max_generations = 15;
    KnapsackConfig<float>* raw = cfg.get();
    Solver<float> solver(std::move(cfg), false);
    solver.set_seed(5);
    solver.run();
    const auto& best = solver.get_best_individual();
    const float via_abi = abi_best(raw, 5, 15);
    EXPECT(best.fitness == via_abi, "knapsack(greedy=%d) facade %.1f vs C ABI %.1f", greedy, best.fitness, via_abi);
    const auto sel = raw->decode_to_selection(best);
    double sw = 0, sv = 0;
    for (int i = 0; i < 120; ++i)
        if (sel[i] > 0.5f) {
            sw += w[i];
            sv += v[i];
        }
    const double expect = sw > raw->get_capacity() ? sv - 1000.0 * (sw - raw->get_capacity()) : sv;
    EXPECT(expect == best.fitness, "knapsack selection value %.1f vs fitness %.1f", expect, best.fitness);
    if (greedy) EXPECT(sw <= raw->get_capacity(), "greedy overfills");
}

static void test_tspj()
{
    unsigned s = 3;
    const int n = 12, dim = n + 1;
    std::vector<std::pair<float, float>> xy;
    for (int i = 0; i < dim; ++i) xy.push_back({(float)(lcg(s) % 1000), (float)(lcg(s) % 1000)});
    std::vector<std::vector<float>> t(dim, std::vector<float>(dim)), j(dim, std::vector<float>(dim, 0.f));
    for (int a = 0; a < dim; ++a)
        for (int b = 0; b < dim; ++b) {
            const float dx = xy[a].first - xy[b].first, dy = xy[a].second - xy[b].second;
            t[a][b] = std::sqrt(dx * dx + dy * dy);
            if (a > 0 && b > 0) j[a][b] = (float)(1 + lcg(s) % 49);
        }
    auto cfg = std::make_unique<TSPJConfig<float>>(t, j, "tspj12");
    cfg->population_size = 64;
    cfg->elite_size = 12;
    cfg->mutant_size = 6;
    cfg->max_generations = 10;
    TSPJConfig<float>* raw = cfg.get();
    Solver<float> solver(std::move(cfg), false);
    solver.set_seed(8);
    solver.run();
    const auto& best = solver.get_best_individual();
    EXPECT(best.fitness == abi_best(raw, 8, 10), "TSPJ facade vs C ABI");
    const auto sol = raw->decode_solution_internal(best);
    EXPECT(is_perm(sol.first, 1) && is_perm(sol.second, 1), "TSPJ decode not 1-based permutations");
}

static void test_multi_tsp()
{
    unsigned s = 21;
    std::vector<std::pair<float, float>> a, b;
    for (int i = 0; i < 31; ++i) {
        a.push_back({(float)(lcg(s) % 1000), (float)(lcg(s) % 1000)});
        b.push_back({(float)(lcg(s) % 1000), (float)(lcg(s) % 1000)});
    }
    auto cfg = MultiTSPConfig<float>::from_coordinates(a, b, "bitsp30");
    cfg->population_size = 64;
    cfg->max_generations = 12;
    MultiTSPConfig<float>* raw = cfg.get();
    Solver<float> solver(std::move(cfg), false);
    solver.set_seed(4);
    solver.run();
    const auto front = solver.get_pareto_front();
    EXPECT(!front.empty(), "empty Pareto front");
    for (const auto& ind : front) {
        EXPECT(ind.objectives.size() == 2, "objectives");
        EXPECT(raw->objective_functions[0](ind) == ind.objectives[0] &&
                   raw->objective_functions[1](ind) == ind.objectives[1],
               "objective callbacks disagree with the device objectives");
        EXPECT(is_perm(raw->decode_tour(ind), 1), "bi-TSP decode not a 1-based permutation");
        for (const auto& other : front) EXPECT(!other.dominates(ind), "front member dominated");
    }
}

static void test_continuous()
{
    for (int f = 0; f < 3; ++f) {
        auto cfg = f == 0 ? ContinuousOptGPUConfig<float>::create_sphere(20)
                   : f == 1 ? ContinuousOptGPUConfig<float>::create_rastrigin(20)
                            : ContinuousOptGPUConfig<float>::create_rosenbrock(20);
        cfg->population_size = 128;
        cfg->elite_size = 20;
        cfg->mutant_size = 12;
        cfg->max_generations = 30;
        ContinuousOptGPUConfig<float>* raw = cfg.get();
        const float want = abi_best(raw, 77, 30);
        Solver<float> solver(std::move(cfg), false);
        solver.set_seed(77);
        solver.set_islands(1, 0, 0);
        solver.run();
        const auto best = solver.get_best_individual();
        EXPECT(best.fitness == want, "continuous %d: facade best %g vs C ABI %g", f, best.fitness, want);
        EXPECT(raw->fitness_function(best) == best.fitness, "continuous %d: callback disagrees", f);
        const auto x = raw->decode_to_solution(best)[0];
        const auto& g = best.get_chromosome();
        for (std::size_t i = 0; i < x.size(); ++i) {
            const float w = raw->get_upper_bound() - raw->get_lower_bound();
            const float t = g[i] * w;
            EXPECT(x[i] == t + raw->get_lower_bound(), "continuous decode x[%zu]", i);
        }
    }
}

static void test_job_scheduling()
{
    auto cfg = JobSchedulingGPUConfig<float>::create_test_instance(50, 5);
    cfg->population_size = 128;
    cfg->elite_size = 20;
    cfg->mutant_size = 12;
    cfg->max_generations = 40;
    JobSchedulingGPUConfig<float>* raw = cfg.get();
    const float want = abi_best(raw, 5, 40);
    Solver<float> solver(std::move(cfg), false);
    solver.set_seed(5);
    solver.set_islands(1, 0, 0);
    solver.run();
    const auto best = solver.get_best_individual();
    EXPECT(best.fitness == want, "job scheduling: facade best %g vs C ABI %g", best.fitness, want);
    const auto m = raw->decode_to_solution(best)[0];
    std::vector<float> load(5, 0.0f);
    for (std::size_t i = 0; i < m.size(); ++i) {
        EXPECT(m[i] >= 0 && m[i] < 5, "machine out of range");
        load[(int)m[i]] += raw->get_job_times()[i];
    }
    EXPECT(*std::max_element(load.begin(), load.end()) == best.fitness, "makespan is not the largest load");
}

static void test_zdt1()
{
    auto cfg = std::make_unique<ZDT1Config<float>>(30);
    cfg->max_generations = 40;
    ZDT1Config<float>* raw = cfg.get();
    Solver<float> solver(std::move(cfg), false);
    solver.set_seed(9);
    solver.run();
    const auto front = solver.get_pareto_front();
    EXPECT(!front.empty(), "ZDT1: empty front");
    for (const auto& ind : front) {
        EXPECT(raw->objective_f1(ind) == ind.objectives[0] && raw->objective_f2(ind) == ind.objectives[1],
               "ZDT1 objective callbacks disagree");
        EXPECT(ind.objectives[1] >= 1.0f - std::sqrt(ind.objectives[0]) - 1e-5f, "ZDT1 below the true front");
        for (const auto& other : front) EXPECT(!other.dominates(ind), "ZDT1 front member dominated");
    }
}

static void test_vrp_blazer()
{
    unsigned s = 41;
    const int dim = 41;
    std::vector<std::pair<float, float>> p;
    for (int i = 0; i < dim; ++i) p.push_back({(float)(lcg(s) % 100), (float)(lcg(s) % 100)});
    std::vector<std::vector<float>> tm(dim, std::vector<float>(dim));
    for (int i = 0; i < dim; ++i)
        for (int j = 0; j < dim; ++j)
            tm[i][j] = std::sqrt((p[i].first - p[j].first) * (p[i].first - p[j].first) +
                                 (p[i].second - p[j].second) * (p[i].second - p[j].second));
    auto cfg = std::make_unique<VrpBlazerConfig<float>>();
    cfg->max_stops_per_day = 6;
    cfg->max_drive_hours_per_day = 4.0;
    cfg->update_derived_params();
    cfg->set_time_matrix(tm);
    cfg->population_size = 128;
    cfg->elite_size = 25;
    cfg->mutant_size = 12;
    cfg->max_generations = 40;
    VrpBlazerConfig<float>* raw = cfg.get();
    const float want = abi_best(raw, 13, 40);
    Solver<float> solver(std::move(cfg), false);
    solver.set_seed(13);
    solver.set_islands(1, 0, 0);
    solver.run();
    const auto best = solver.get_best_individual();
    EXPECT(best.fitness == want, "VRP days: facade best %g vs C ABI %g", best.fitness, want);
    EXPECT(raw->calculate_fitness(best) == best.fitness, "VRP days callback disagrees");
    const auto sol = raw->decode_solution(best.get_chromosome().data());
    std::set<int> seen;
    for (const auto& d : sol.days) {
        EXPECT((int)d.stop_indices.size() <= 6 && !d.stop_indices.empty(), "day size");
        EXPECT(d.total_drive_minutes <= 240.0f + 1e-3f, "day over the drive limit");
        for (int st : d.stop_indices) seen.insert(st);
    }
    EXPECT((int)seen.size() + (int)sol.unvisited_stops.size() / 2 == dim - 1, "stops lost");
    const float approx = sol.total_drive_minutes + 1000.0f * sol.unvisited_stops.size();
    EXPECT(std::fabs(approx - best.fitness) <= 1e-5f * best.fitness, "fitness vs listed days: %g vs %g", approx,
           best.fitness);
    raw->export_solution_json(best, "/tmp/facade_vrp_days.json", 1.0);
    std::ifstream in("/tmp/facade_vrp_days.json");
    EXPECT(in.good(), "VRP JSON export missing");
}

static void test_movrp()
{
    unsigned s = 61;
    const int dim = 32;
    const char* path = "/tmp/facade_movrp.vrp";
    {
        std::ofstream f(path);
        f << "NAME : A-n32-k5\nTYPE : CVRP\nDIMENSION : " << dim << "\nCAPACITY : 100\nNODE_COORD_SECTION\n";
        for (int i = 0; i < dim; ++i) f << (i + 1) << " " << (lcg(s) % 100) << " " << (lcg(s) % 100) << "\n";
        f << "DEMAND_SECTION\n";
        for (int i = 0; i < dim; ++i) f << (i + 1) << " " << (i == 0 ? 0 : 1 + (int)(lcg(s) % 29)) << "\n";
        f << "DEPOT_SECTION\n1\n-1\nEOF\n";
    }
    auto cfg = MultiObjectiveVRPConfig<float>::load_from_file(path);
    EXPECT(cfg->get_num_customers() == 31 && cfg->get_num_vehicles() == 5 && cfg->get_vehicle_capacity() == 100,
           "MO-VRP loader");
    cfg->population_size = 64;
    cfg->max_generations = 20;
    MultiObjectiveVRPConfig<float>* raw = cfg.get();
    Solver<float> solver(std::move(cfg), false);
    solver.set_seed(3);
    solver.run();
    const auto front = solver.get_pareto_front();
    EXPECT(!front.empty(), "MO-VRP: empty front");
    for (const auto& ind : front) {
        float obj[2];
        raw->decode_solution(ind.get_chromosome().data(), obj);
        EXPECT(obj[0] == ind.objectives[0] && obj[1] == ind.objectives[1], "MO-VRP decode_solution disagrees");
        const auto routes = raw->decode_routes(ind.get_chromosome().data());
        std::set<int> seen;
        float longest = 0.0f;
        for (const auto& r : routes) {
            EXPECT(r.load <= 100, "route over capacity");
            for (int c : r.customers) seen.insert(c);
            longest = std::max(longest, r.distance);
        }
        EXPECT((int)seen.size() == 31, "customers lost");
        const float pen = routes.size() > 5 ? (routes.size() - 5) * 500.0f : 0.0f;
        EXPECT(std::fabs(longest + pen - ind.objectives[1]) <= 1e-5f * ind.objectives[1], "longest route");
        for (const auto& other : front) EXPECT(!other.dominates(ind), "MO-VRP front member dominated");
    }
    raw->export_solution(front[0], "/tmp/facade_movrp.txt");
    std::ifstream in("/tmp/facade_movrp.txt");
    EXPECT(in.good(), "MO-VRP export missing");
}

int main()
{
    int count = 0;
    if (brkga_device_count(&count) != BRKGA_OK || count == 0) {
        std::printf("no GPU\n");
        return 2;
    }
    test_tsp();
    test_knapsack(false);
    test_knapsack(true);
    test_tspj();
    test_multi_tsp();
    test_continuous();
    test_job_scheduling();
    test_zdt1();
    test_vrp_blazer();
    test_movrp();
    if (failures) {
        std::printf("%d failure(s)\n", failures);
        return 1;
    }
    std::printf("FACADE OK\n");
    return 0;
}
