// host_config_test.cpp - configs with host callbacks only, written the way
// the reference's code generator is told to write them
// (llm_solver/context/context_for_llm_minimal.txt:13-30): fitness_function /
// decoder / comparator lambdas, no engine decoder.  They compile unchanged
// against the facade and run on the device with host evaluation.  Checks:
//   * a greedy-knapsack config in host lambdas follows, bit for bit, the
//     trajectory of the facade's KnapsackConfig with the device's greedy
//     decoder (same seed, islands, migration; integer fitness);
//   * a bi-objective config in host objective_functions runs NSGA-II; the
//     host-view non-dominated sort gives the device's front ranks;
//   * the host-view operations of Population and the Solver reports run.
// Prints HOST CONFIGS OK and exits 0 on success.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <numeric>
#include <vector>

#include "brkga/configs/knapsack_config.hpp"
#include "brkga/core/solver.hpp"

static int failures = 0;
#define EXPECT(cond, ...)                                    \
    do {                                                     \
        if (!(cond)) {                                       \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
            std::printf(__VA_ARGS__);                        \
            std::printf("\n");                               \
            ++failures;                                      \
        }                                                    \
    } while (0)

// ---- a config in the generator's template form -------------------------
template<typename T>
class GreedyPackConfig : public BRKGAConfig<T> {
public:
    GreedyPackConfig(const std::vector<int>& w, const std::vector<int>& v, int cap)
        : BRKGAConfig<T>({(int)w.size()}), weights(w), values(v), capacity(cap) {
        this->fitness_function = [this](const Individual<T>& ind) {
            const auto& chromosome = ind.get_chromosome();
            std::vector<int> order;
            order.reserve(chromosome.size());
            for (int i = 0; i < (int)chromosome.size(); ++i) order.push_back(i);
            std::stable_sort(order.begin(), order.end(),
                             [&](int a, int b) { return chromosome[a] < chromosome[b]; });
            int left = capacity;
            T value = 0;
            for (int i : order)
                if (weights[i] <= left) {
                    left -= weights[i];
                    value += values[i];
                }
            return value;
        };
        this->decoder = [this](const Individual<T>& ind) {
            std::vector<std::vector<T>> result(1);
            const auto& chromosome = ind.get_chromosome();
            result[0].reserve(chromosome.size());
            for (T k : chromosome) result[0].push_back(k < T(0.5) ? T(1) : T(0));
            return result;
        };
        this->comparator = [](T a, T b) { return a > b; };
    }

private:
    std::vector<int> weights, values;
    int capacity;
};

// ---- a bi-objective config in host objective functions -----------------
template<typename T>
class TwoObjectiveConfig : public BRKGAConfig<T> {
public:
    explicit TwoObjectiveConfig(int n) : BRKGAConfig<T>({n}, 2) {
        this->objective_functions[0] = [](const Individual<T>& ind) { return ind.get_chromosome()[0]; };
        this->objective_functions[1] = [n](const Individual<T>& ind) {
            const auto& c = ind.get_chromosome();
            T s = 0;
            for (int i = 1; i < n; ++i) s += c[i];
            const T g = T(1) + T(9) * s / T(n - 1);
            return g * (T(1) - std::sqrt(c[0] / g));
        };
    }
};

static unsigned lcg(unsigned& s)
{
    s = s * 1664525u + 1013904223u;
    return s >> 8;
}

int main()
{
    const int n = 150, P = 128, G = 12;
    unsigned seed = 99;
    std::vector<int> w(n), v(n);
    long long total = 0;
    for (int i = 0; i < n; ++i) {
        w[i] = 1 + (int)(lcg(seed) % 100);
        v[i] = w[i] + (int)(lcg(seed) % 11);
        total += w[i];
    }
    const int cap = (int)(total / 2);

    // host lambdas vs the device's greedy decoder
    auto host_cfg = std::make_unique<GreedyPackConfig<float>>(w, v, cap);
    std::vector<float> wf(w.begin(), w.end()), vf(v.begin(), v.end());
    auto dev_cfg = std::make_unique<KnapsackConfig<float>>(wf, vf, (float)cap);
    dev_cfg->set_greedy(true);
    for (BRKGAConfig<float>* c : {(BRKGAConfig<float>*)host_cfg.get(), (BRKGAConfig<float>*)dev_cfg.get()}) {
        c->population_size = P;
        c->elite_size = 24;
        c->mutant_size = 12;
        c->max_generations = G;
    }
    EXPECT(host_cfg->native_kind() < 0, "the template config has no engine decoder");
    Solver<float> hs(std::move(host_cfg), false);
    Solver<float> ds(std::move(dev_cfg), false);
    for (Solver<float>* s : {&hs, &ds}) {
        s->set_seed(0xABCD);
        s->set_islands(2, 3, 2);
    }
    hs.set_host_threads(4);
    hs.run();
    ds.run();
    EXPECT(hs.get_population().host_evaluated() && !ds.get_population().host_evaluated(), "evaluation paths");
    const auto& hi = hs.get_population().get_individuals();
    const auto& di = ds.get_population().get_individuals();
    EXPECT(hi.size() == di.size() && (int)hi.size() == P, "population sizes");
    int diff = 0;
    for (std::size_t i = 0; i < hi.size() && i < di.size(); ++i)
        if (hi[i].fitness != di[i].fitness || hi[i].flatten() != di[i].flatten()) ++diff;
    EXPECT(diff == 0, "%d members differ between host lambdas and the device decoder", diff);
    EXPECT(hs.get_best_individual().fitness == ds.get_best_individual().fitness, "best fitness %f vs %f",
           hs.get_best_individual().fitness, ds.get_best_individual().fitness);
    EXPECT(hs.get_fitness_history() == ds.get_fitness_history(), "fitness histories differ");

    // host-view operations
    auto& pop = hs.get_population();
    pop.evaluate_all();
    const auto elite = pop.get_elite(), rest = pop.get_non_elite();
    EXPECT(elite.size() == 24 && rest.size() == (std::size_t)P - 24, "elite split %zu / %zu", elite.size(), rest.size());
    for (const auto& r : rest) EXPECT(r.fitness <= elite.back().fitness, "non-elite member better than the elite");
    pop.update_best();
    EXPECT(pop.get_best().fitness == elite.front().fitness, "update_best");
    hs.print_multi_gpu_performance();
    hs.benchmark_modes();

    // bi-objective host callbacks on NSGA-II
    auto mo = std::make_unique<TwoObjectiveConfig<float>>(12);
    mo->population_size = 64;
    mo->max_generations = 10;
    Solver<float> ms(std::move(mo), false);
    ms.set_seed(7);
    ms.run();
    auto& mp = ms.get_population();
    std::vector<int> dev_rank;
    for (const auto& ind : mp.get_individuals()) dev_rank.push_back(ind.rank);
    mp.fast_non_dominated_sort();
    std::vector<int> host_rank;
    for (const auto& ind : mp.get_individuals()) host_rank.push_back(ind.rank);
    EXPECT(dev_rank == host_rank, "host-view fronts differ from the device ranks");
    const auto front = ms.get_pareto_front();
    EXPECT(!front.empty(), "empty front");
    for (const auto& a : front)
        for (const auto& b : front) EXPECT(!a.dominates(b), "front member dominated");
    const float hv = mp.calculate_hypervolume_2d(1.0f, 10.0f);
    const double spread = mp.calculate_spread();
    EXPECT(hv > 0.0f && std::isfinite(spread), "hypervolume %f spread %f", hv, spread);
    mp.calculate_crowding_distance();
    mp.select_next_generation_nsga2();
    EXPECT(mp.get_individuals().size() == 64, "selection keeps the population size");

    if (failures) {
        std::printf("%d failure(s)\n", failures);
        return 1;
    }
    std::printf("HOST CONFIGS OK\n");
    return 0;
}
