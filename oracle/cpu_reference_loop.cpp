// cpu_reference_loop.cpp - CPU baseline of the benchmark (test and
// measurement infrastructure, never linked into the engine).
//
// The reference's own host loop cannot be compiled here (population.hpp,
// solver.hpp and tsp_config.hpp include cuda_runtime.h), so this program
// restates it with the same data structures and the same work per
// generation, for timing only:
//   * individuals are objects holding vector<vector<float>> chromosomes, a
//     fitness and an evaluated flag (core/individual.hpp:10-226);
//   * fitness = closed-tour length over an n x n FP32 distance matrix built
//     from the coordinates (configs/tsp_config.hpp:359-371), the tour
//     decoded by std::sort of (key, city) pairs (tsp_config.hpp:180-225);
//   * a generation (core/solver.hpp:234-273) resets the next generation,
//     copies the elite (population.hpp:700-714), copies the elite and the
//     non-elite out (get_elite / get_non_elite, population.hpp:787-803),
//     builds offspring with mt19937 draws: one uniform_int per parent and one
//     uniform_real<double> coin per gene (solver.hpp:1109-1135), refills the
//     mutant slots (solver.hpp:1244-1258), swaps, evaluates every member not
//     yet evaluated (evaluate_all_cpu, population.hpp:626-653) and sorts the
//     Individual objects by fitness (population.hpp:660-682);
//   * islands: each island is such a population; every migration_interval
//     generations island i's last k members are replaced by copies of
//     island (i - 1)'s best k (core/solver.hpp:517-558).
// The evaluation runs serially (as evaluate_all_cpu does) or, with
// --threads > 1, with OpenMP over the members of each island.
//
// Usage: cpu_reference_loop COORDS_F32 N POP_PER_ISLAND ISLANDS ELITE MUTANTS
//        RHO GENERATIONS MIG_INTERVAL MIG_K THREADS RUNS SEED
// COORDS_F32: N x 2 float32 coordinates (raw little-endian).
// Prints one JSON line: per-run seconds, median, generations per second.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct Member {
    std::vector<std::vector<float>> chromosomes;
    float fitness = 0.0f;
    bool evaluated = false;
};

struct Instance {
    int n = 0;
    std::vector<float> dist;  // n x n, zero diagonal
};

Instance build_instance(const std::vector<float>& xy, int n)
{
    Instance in;
    in.n = n;
    in.dist.assign((size_t)n * n, 0.0f);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            if (i == j) continue;
            const float dx = xy[2 * i] - xy[2 * j];
            const float dy = xy[2 * i + 1] - xy[2 * j + 1];
            in.dist[(size_t)i * n + j] = std::sqrt(dx * dx + dy * dy);
        }
    return in;
}

float tour_length(const Instance& in, const Member& m)
{
    const std::vector<float>& keys = m.chromosomes[0];
    std::vector<std::pair<float, int>> order;
    order.reserve(keys.size());
    for (int i = 0; i < (int)keys.size(); ++i) order.emplace_back(keys[i], i);
    std::sort(order.begin(), order.end());
    float total = 0.0f;
    for (int i = 0; i + 1 < in.n; ++i) total += in.dist[(size_t)order[i].second * in.n + order[i + 1].second];
    if (in.n > 1) total += in.dist[(size_t)order[in.n - 1].second * in.n + order[0].second];
    return total;
}

struct Island {
    std::vector<Member> individuals, next_generation;
};

struct Run {
    const Instance* inst;
    int P, E, M, islands, mig_interval, mig_k, threads;
    double rho;
    std::mt19937 rng;
    std::vector<Island> isl;

    void randomize(Member& m)
    {
        std::uniform_real_distribution<float> u(0.0f, 1.0f);
        for (auto& c : m.chromosomes)
            for (auto& g : c) g = u(rng);
        m.evaluated = false;
    }

    void evaluate(Island& is)
    {
        const int P_ = (int)is.individuals.size();
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads) if (threads > 1)
#endif
        for (int i = 0; i < P_; ++i) {
            Member& m = is.individuals[i];
            if (!m.evaluated) {
                m.fitness = tour_length(*inst, m);
                m.evaluated = true;
            }
        }
    }

    static void sort_members(Island& is)
    {
        std::sort(is.individuals.begin(), is.individuals.end(),
                  [](const Member& a, const Member& b) { return a.fitness < b.fitness; });
    }

    void initialize()
    {
        isl.assign(islands, Island());
        for (auto& is : isl) {
            is.individuals.assign(P, Member());
            is.next_generation.assign(P, Member());
            for (int i = 0; i < P; ++i) {
                is.individuals[i].chromosomes.assign(1, std::vector<float>(inst->n));
                is.next_generation[i].chromosomes.assign(1, std::vector<float>(inst->n));
                randomize(is.individuals[i]);
            }
            evaluate(is);
            sort_members(is);
        }
    }

    void generation(Island& is)
    {
        for (auto& m : is.next_generation) m.evaluated = false;
        for (int i = 0; i < E; ++i) is.next_generation[i] = is.individuals[i];
        std::vector<Member> elite(is.individuals.begin(), is.individuals.begin() + E);
        std::vector<Member> non_elite(is.individuals.begin() + E, is.individuals.end());
        std::uniform_real_distribution<double> coin(0.0, 1.0);
        std::uniform_int_distribution<int> pick_e(0, E - 1);
        std::uniform_int_distribution<int> pick_n(0, (int)non_elite.size() - 1);
        const int offspring = P - E - M;
        for (int c = 0; c < offspring; ++c) {
            const int a = pick_e(rng), b = pick_n(rng);
            Member& child = is.next_generation[E + c];
            for (size_t comp = 0; comp < child.chromosomes.size(); ++comp)
                for (size_t j = 0; j < child.chromosomes[comp].size(); ++j)
                    child.chromosomes[comp][j] =
                        coin(rng) < rho ? elite[a].chromosomes[comp][j] : non_elite[b].chromosomes[comp][j];
            child.evaluated = false;
        }
        for (int i = P - M; i < P; ++i) randomize(is.next_generation[i]);
        std::swap(is.individuals, is.next_generation);
        evaluate(is);
        sort_members(is);
    }

    void migrate()
    {
        std::vector<std::vector<Member>> best(islands);
        for (int i = 0; i < islands; ++i)
            best[i].assign(isl[i].individuals.begin(), isl[i].individuals.begin() + mig_k);
        for (int i = 0; i < islands; ++i) {
            const auto& src = best[(i + islands - 1) % islands];
            for (int k = 0; k < mig_k; ++k) isl[i].individuals[P - mig_k + k] = src[k];
            sort_members(isl[i]);
        }
    }
};

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 14) {
        std::fprintf(stderr,
                     "usage: %s COORDS_F32 N POP_PER_ISLAND ISLANDS ELITE MUTANTS RHO GENERATIONS MIG_INTERVAL MIG_K "
                     "THREADS RUNS SEED\n",
                     argv[0]);
        return 2;
    }
    const int n = std::atoi(argv[2]);
    std::vector<float> xy((size_t)2 * n);
    FILE* f = std::fopen(argv[1], "rb");
    if (!f || std::fread(xy.data(), sizeof(float), xy.size(), f) != xy.size()) {
        std::fprintf(stderr, "cannot read %d coordinate pairs from %s\n", n, argv[1]);
        return 2;
    }
    std::fclose(f);
    const Instance inst = build_instance(xy, n);
    Run r;
    r.inst = &inst;
    r.P = std::atoi(argv[3]);
    r.islands = std::atoi(argv[4]);
    r.E = std::atoi(argv[5]);
    r.M = std::atoi(argv[6]);
    r.rho = std::atof(argv[7]);
    const int G = std::atoi(argv[8]);
    r.mig_interval = std::atoi(argv[9]);
    r.mig_k = std::atoi(argv[10]);
    r.threads = std::max(1, std::atoi(argv[11]));
    const int runs = std::max(1, std::atoi(argv[12]));
    const unsigned seed = (unsigned)std::strtoul(argv[13], nullptr, 0);
    if (r.P <= 0 || r.islands <= 0 || r.E <= 0 || r.M < 0 || r.E + r.M > r.P || G <= 0) {
        std::fprintf(stderr, "invalid sizes\n");
        return 2;
    }
    std::vector<double> secs;
    float best = 0.0f;
    // run 0 is the warm-up; runs 1..runs are timed
    for (int run = 0; run <= runs; ++run) {
        r.rng.seed(seed + (unsigned)run);
        r.initialize();
        const auto t0 = std::chrono::steady_clock::now();
        for (int g = 1; g <= G; ++g) {
            for (auto& is : r.isl) r.generation(is);
            if (r.islands > 1 && r.mig_interval > 0 && r.mig_k > 0 && g % r.mig_interval == 0) r.migrate();
        }
        const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (run > 0) secs.push_back(dt);
        best = r.isl[0].individuals[0].fitness;
        for (const auto& is : r.isl) best = std::min(best, is.individuals[0].fitness);
    }
    std::vector<double> sorted = secs;
    std::sort(sorted.begin(), sorted.end());
    const double med = sorted[sorted.size() / 2];
    std::printf("{\"runs_s\": [");
    for (size_t i = 0; i < secs.size(); ++i) std::printf("%s%.6f", i ? ", " : "", secs[i]);
    std::printf("], \"median_s\": %.6f, \"generations\": %d, \"gens_per_s\": %.4f, \"threads\": %d, "
                "\"best\": %.3f}\n",
                med, G, G / med, r.threads, best);
    return 0;
}
