// brkga/core/population.hpp - device-resident population of the C++ facade.
//
// Interface of the reference's Population<T> (core/population.hpp:21-1097)
// and of its device-resident sibling GPUPopulation<T>
// (core/gpu_population.cuh:52-508), backed by an engine population
// (brkga_pop_t).  The keys, fitness and order live in device memory; host
// views (get_individuals, get_best, get_pareto_front) are materialised on
// demand, as GPUPopulation::sync_to_host does.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <iomanip>
#include <iostream>
#include <limits>
#include <numeric>
#include <set>
#include <thread>
#include <utility>
#include <vector>

#include "config.hpp"

struct BRKGAEngineOptions {
    int device = 0;
    int num_islands = 1;         // solver.hpp:378-476 (islands on one GPU)
    int migration_interval = 0;  // solver.hpp:164
    int num_migrants = 0;        // solver.hpp:164
    uint64_t seed = 0x5EED;      // the reference seeds from the clock
    bool use_graph = true;       // replay generation loops from a HIP graph
    // local search of configs with has_local_search() (TSPJ), with the
    // reference's defaults (solver.hpp:166)
    bool local_search = true;
    int ls_interval = 2;
    int ls_individuals = 100;
    int ls_moves = 5;
    // multi-objective configs: Population::ensure_minimum_diversity after
    // every 25th generation step, as Solver does (solver.hpp:812-814); 0 = off
    // (configs without an engine decoder run without it)
    int diversity_interval = 25;
    // configs without an engine decoder: host threads that run the config's
    // fitness_function / objective_functions over a generation's new rows
    // (0: one per hardware thread; 1: serially, as evaluate_all_cpu does,
    // population.hpp:626-653, for callbacks that are not thread-safe)
    int host_threads = 0;
};

template <typename T>
class Population {
public:
    // A config with an engine decoder (native_kind() >= 0) runs the whole
    // generation on the device.  Any other config (a host fitness_function /
    // objective_functions, like the reference's generated configs) runs with
    // its keys built, partitioned and selected on the device and its
    // callbacks evaluated on the host between the two halves of each
    // generation (brkga_population_host_begin / _end).
    Population(BRKGAConfig<T>* cfg, const BRKGAEngineOptions& opt = BRKGAEngineOptions()) : config_(cfg), opt_(opt)
    {
        problem_ = cfg->native_problem(opt.device);
        if (!problem_) {
            const int M = cfg->num_objectives;
            if (M == 1 ? !cfg->fitness_function : (int)cfg->objective_functions.size() < M)
                throw std::runtime_error("config has neither an engine decoder nor host fitness callbacks");
            // comparator a > b: larger fitness is better (knapsack_config.hpp:37)
            const int maximize = (M == 1 && cfg->comparator && cfg->comparator(T(1), T(0))) ? 1 : 0;
            brkga_facade::check(brkga_problem_create_host(brkga_facade::engine_for_device(opt.device),
                                                          cfg->get_total_chromosome_length(), M, maximize,
                                                          &host_problem_),
                                "brkga_problem_create_host");
            problem_ = host_problem_;
            opt_.diversity_interval = 0;
        }
        brkga_params p{};
        p.population_size = cfg->population_size;
        p.elite_size = cfg->elite_size;
        p.mutant_size = cfg->mutant_size;
        p.elite_prob = cfg->elite_prob;
        p.num_islands = opt.num_islands;
        p.migration_interval = opt.migration_interval;
        p.num_migrants = opt.num_migrants;
        p.seed = opt.seed;
        p.diversity_interval = opt_.diversity_interval;
        if (opt.local_search && cfg->has_local_search()) {
            p.ls_interval = opt.ls_interval;
            p.ls_individuals = opt.ls_individuals;
            p.ls_moves = opt.ls_moves;
        }
        brkga_facade::check(brkga_population_create(problem_, &p, &pop_), "brkga_population_create");
    }

    ~Population()
    {
        if (pop_) brkga_population_destroy(pop_);
        if (host_problem_) brkga_problem_destroy(host_problem_);
    }

    // Whether the config's callbacks are evaluated on the host.
    bool host_evaluated() const { return host_problem_ != nullptr; }

    Population(const Population&) = delete;
    Population& operator=(const Population&) = delete;

    // population.hpp:107-124
    void initialize()
    {
        if (host_problem_) host_generation(true);
        else brkga_facade::check(brkga_population_initialize(pop_, nullptr), "brkga_population_initialize");
        fitness_history_.clear();
        dirty_ = true;
        record_fitness();
    }

    // One or more generation steps on the device.
    void evolve(int generations)
    {
        if (generations <= 0) return;
        if (host_problem_) {
            for (int g = 0; g < generations; ++g) host_generation(false);
        } else if (opt_.use_graph)
            brkga_facade::check(brkga_evolve_graph(pop_, generations, nullptr), "brkga_evolve_graph");
        else
            brkga_facade::check(brkga_evolve(pop_, generations, nullptr), "brkga_evolve");
        dirty_ = true;
    }

    // Appends the current best fitness to the history (population.hpp:740-742).
    void record_fitness() { fitness_history_.push_back(get_best().fitness); }

    const Individual<T>& get_best() const
    {
        sync();
        return best_;
    }

    const std::vector<Individual<T>>& get_individuals() const
    {
        sync();
        return individuals_;
    }

    // Members of front 0 (multi-objective) or the best individual.
    std::vector<Individual<T>> get_pareto_front() const
    {
        sync();
        std::vector<Individual<T>> out;
        if (config_->num_objectives > 1) {
            for (const auto& ind : individuals_)
                if (ind.rank == 0) out.push_back(ind);
        } else {
            out.push_back(best_);
        }
        return out;
    }

    std::vector<std::vector<int>> get_fronts() const
    {
        sync();
        std::vector<std::vector<int>> fronts;
        for (std::size_t i = 0; i < individuals_.size(); ++i) {
            const int r = individuals_[i].rank;
            if (r < 0) continue;
            if ((int)fronts.size() <= r) fronts.resize(r + 1);
            fronts[r].push_back((int)i);
        }
        return fronts;
    }

    const std::vector<T>& get_fitness_history() const { return fitness_history_; }
    int generation() const
    {
        int32_t g = 0;
        brkga_facade::check(brkga_population_generation(pop_, &g), "brkga_population_generation");
        return g;
    }

    T get_average_fitness() const
    {
        sync();
        double s = 0.0;
        for (const auto& ind : individuals_) s += ind.fitness;
        return individuals_.empty() ? T(0) : T(s / individuals_.size());
    }

    // Mean pairwise Euclidean key distance (population.hpp:819-843); O(P^2 L),
    // for reporting only.
    double get_diversity() const
    {
        sync();
        const std::size_t P = individuals_.size();
        if (P < 2) return 0.0;
        double total = 0.0;
        std::size_t pairs = 0;
        std::vector<std::vector<T>> rows;
        for (const auto& ind : individuals_) rows.push_back(ind.flatten());
        for (std::size_t i = 0; i < P; ++i)
            for (std::size_t j = i + 1; j < P; ++j) {
                double d = 0.0;
                for (std::size_t k = 0; k < rows[i].size(); ++k) {
                    const double t = (double)rows[i][k] - (double)rows[j][k];
                    d += t * t;
                }
                total += std::sqrt(d);
                ++pairs;
            }
        return total / pairs;
    }

    void print_statistics(int gen) const
    {
        sync();
        if (config_->num_objectives > 1) {
            auto fronts = get_fronts();
            std::set<std::pair<T, T>> uniq;
            for (const auto& ind : individuals_)
                if (ind.rank == 0) uniq.insert({ind.objectives[0], ind.objectives[1]});
            std::cout << "Gen " << std::setw(4) << gen << ": Fronts=" << fronts.size()
                      << ", Pareto=" << (fronts.empty() ? 0 : fronts[0].size()) << ", Unique=" << uniq.size()
                      << " [GPU-Resident]" << std::endl;
        } else {
            std::cout << "Generation " << std::setw(4) << gen << ": Best=" << std::fixed << std::setprecision(2)
                      << best_.fitness << " [GPU-Resident]" << std::endl;
        }
    }

    brkga_pop_t handle() const { return pop_; }

    // ----- host-view operations of the reference's Population<T> -----
    // The device population evaluates, ranks and selects every generation
    // itself; these methods recompute the same quantities on the host view
    // (get_individuals), for callers that drive or inspect a population the
    // way the reference's host loop does.

    // population.hpp:523-529: every member of a completed generation is
    // evaluated already; this brings the host view up to date.
    void evaluate_all() { sync(); }

    // population.hpp:409-440: front index of every member of the host view
    // (minimisation dominance, individual.hpp:99-114), fronts by peeling.
    void fast_non_dominated_sort()
    {
        sync();
        const std::size_t n = individuals_.size();
        std::vector<int> count(n, 0);
        std::vector<std::vector<int>> dominated(n);
        for (std::size_t a = 0; a < n; ++a)
            for (std::size_t b = 0; b < n; ++b) {
                if (a == b) continue;
                if (individuals_[a].dominates(individuals_[b])) dominated[a].push_back((int)b);
                else if (individuals_[b].dominates(individuals_[a])) ++count[a];
            }
        fronts_.clear();
        std::vector<int> cur;
        for (std::size_t i = 0; i < n; ++i)
            if (count[i] == 0) cur.push_back((int)i);
        int r = 0;
        while (!cur.empty()) {
            std::vector<int> next;
            for (int i : cur) {
                individuals_[i].rank = r;
                for (int j : dominated[i])
                    if (--count[j] == 0) next.push_back(j);
            }
            fronts_.push_back(cur);
            std::sort(next.begin(), next.end());
            cur.swap(next);
            ++r;
        }
    }

    // population.hpp:442-489: crowding distance within each front of the
    // last fast_non_dominated_sort (fronts of two or fewer are infinite).
    void calculate_crowding_distance()
    {
        if (fronts_.empty()) fast_non_dominated_sort();
        const int M = config_->num_objectives;
        for (auto& front : fronts_) {
            for (int i : front) individuals_[i].crowding_distance = T(0);
            if (front.size() <= 2) {
                for (int i : front) individuals_[i].crowding_distance = std::numeric_limits<T>::infinity();
                continue;
            }
            std::vector<int> order = front;
            for (int m = 0; m < M; ++m) {
                std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
                    return individuals_[a].objectives[m] < individuals_[b].objectives[m];
                });
                const T lo = individuals_[order.front()].objectives[m];
                const T hi = individuals_[order.back()].objectives[m];
                individuals_[order.front()].crowding_distance = std::numeric_limits<T>::infinity();
                individuals_[order.back()].crowding_distance = std::numeric_limits<T>::infinity();
                const T range = hi - lo;
                if (!(range > T(0))) continue;
                for (std::size_t k = 1; k + 1 < order.size(); ++k) {
                    T& cd = individuals_[order[k]].crowding_distance;
                    cd += (individuals_[order[k + 1]].objectives[m] - individuals_[order[k - 1]].objectives[m]) / range;
                }
            }
        }
    }

    // population.hpp:491-517: the host view truncated to population_size
    // members, whole fronts first, the cut front by crowding (largest first).
    void select_next_generation_nsga2()
    {
        if (fronts_.empty()) calculate_crowding_distance();
        const std::size_t keep = (std::size_t)config_->population_size;
        std::vector<Individual<T>> next;
        for (auto& front : fronts_) {
            if (next.size() >= keep) break;
            std::vector<int> members = front;
            if (next.size() + members.size() > keep)
                std::stable_sort(members.begin(), members.end(), [&](int a, int b) {
                    return individuals_[a].crowding_distance > individuals_[b].crowding_distance;
                });
            for (int i : members) {
                if (next.size() >= keep) break;
                next.push_back(individuals_[i]);
            }
        }
        individuals_.swap(next);
        fronts_.clear();
    }

    // population.hpp:660-675: the host view ordered best first (rank, then
    // crowding for multi-objective configs; the config's comparator, or
    // ascending fitness, otherwise).
    void sort()
    {
        sync();
        if (config_->num_objectives > 1) {
            std::stable_sort(individuals_.begin(), individuals_.end(), [](const Individual<T>& a, const Individual<T>& b) {
                return a.rank != b.rank ? a.rank < b.rank : a.crowding_distance > b.crowding_distance;
            });
        } else {
            const auto& cmp = config_->comparator;
            std::stable_sort(individuals_.begin(), individuals_.end(), [&](const Individual<T>& a, const Individual<T>& b) {
                return cmp ? cmp(a.fitness, b.fitness) : a.fitness < b.fitness;
            });
        }
    }

    // population.hpp:677-682
    void update_best()
    {
        sort();
        if (!individuals_.empty()) best_ = individuals_.front();
    }

    // population.hpp:787-803: the elite_size best members and the rest.
    std::vector<Individual<T>> get_elite() const
    {
        const auto ranked = ranked_view();
        const std::size_t e = std::min(ranked.size(), (std::size_t)std::max(config_->elite_size, 0));
        return std::vector<Individual<T>>(ranked.begin(), ranked.begin() + e);
    }
    std::vector<Individual<T>> get_non_elite() const
    {
        const auto ranked = ranked_view();
        const std::size_t e = std::min(ranked.size(), (std::size_t)std::max(config_->elite_size, 0));
        return std::vector<Individual<T>>(ranked.begin() + e, ranked.end());
    }

    // population.hpp:925-953: area dominated by the front-0 points that lie
    // within the reference point, swept in ascending first objective.
    T calculate_hypervolume_2d(T ref_x = T(1), T ref_y = T(1)) const
    {
        if (config_->num_objectives != 2) return T(0);
        std::vector<std::pair<T, T>> pts;
        for (const auto& ind : get_pareto_front())
            if (ind.objectives[0] <= ref_x && ind.objectives[1] <= ref_y)
                pts.emplace_back(ind.objectives[0], ind.objectives[1]);
        std::sort(pts.begin(), pts.end());
        T area = T(0), x_prev = T(0);
        for (const auto& q : pts) {
            area += (q.first - x_prev) * (ref_y - q.second);
            x_prev = q.first;
        }
        return area;
    }

    // population.hpp:955-989: spread of front 0 in the order get_pareto_front
    // returns it: mean absolute deviation of consecutive objective-space
    // distances over their mean.
    double calculate_spread() const
    {
        if (config_->num_objectives < 2) return 0.0;
        const auto front = get_pareto_front();
        if (front.size() < 2) return 0.0;
        std::vector<double> gaps;
        for (std::size_t i = 0; i + 1 < front.size(); ++i) {
            double s2 = 0.0;
            for (int m = 0; m < config_->num_objectives; ++m) {
                const double d = (double)front[i].objectives[m] - (double)front[i + 1].objectives[m];
                s2 += d * d;
            }
            gaps.push_back(std::sqrt(s2));
        }
        const double mean = std::accumulate(gaps.begin(), gaps.end(), 0.0) / gaps.size();
        double dev = 0.0;
        for (double g : gaps) dev += std::fabs(g - mean);
        return dev / (gaps.size() * mean);
    }

private:
    // Host view ordered best first, as the reference's individuals are
    // after its per-generation sort.
    std::vector<Individual<T>> ranked_view() const
    {
        sync();
        std::vector<Individual<T>> v = individuals_;
        if (config_->num_objectives > 1) {
            std::stable_sort(v.begin(), v.end(), [](const Individual<T>& a, const Individual<T>& b) {
                return a.rank != b.rank ? a.rank < b.rank : a.crowding_distance > b.crowding_distance;
            });
        } else {
            const auto& cmp = config_->comparator;
            std::stable_sort(v.begin(), v.end(), [&](const Individual<T>& a, const Individual<T>& b) {
                return cmp ? cmp(a.fitness, b.fitness) : a.fitness < b.fitness;
            });
        }
        return v;
    }

    // One generation of a host-evaluated config: keys from the device, the
    // config's callbacks over the rows that need them (evaluate_all_cpu
    // skips members already evaluated: elite and migrant copies), fitness
    // back to the device, which partitions or selects.
    void host_generation(bool init)
    {
        const int P = config_->population_size;
        const int L = config_->get_total_chromosome_length();
        const int M = config_->num_objectives;
        std::vector<float> keys((std::size_t)P * L), fit((std::size_t)P * M);
        std::vector<int32_t> needs(P);
        brkga_facade::check(brkga_population_host_begin(pop_, init ? 1 : 0, keys.data(), fit.data(), needs.data(),
                                                        nullptr),
                            "brkga_population_host_begin");
        std::vector<int> rows;
        for (int i = 0; i < P; ++i)
            if (needs[i]) rows.push_back(i);
        auto work = [&](std::size_t lo, std::size_t hi) {
            for (std::size_t k = lo; k < hi; ++k) {
                const int i = rows[k];
                Individual<T> ind = M > 1 ? Individual<T>(config_->component_lengths, M)
                                          : Individual<T>(config_->component_lengths);
                ind.unflatten(std::vector<T>(keys.begin() + (std::size_t)i * L, keys.begin() + (std::size_t)(i + 1) * L));
                if (M > 1)
                    for (int m = 0; m < M; ++m) fit[(std::size_t)i * M + m] = (float)config_->objective_functions[m](ind);
                else
                    fit[i] = (float)config_->fitness_function(ind);
            }
        };
        std::size_t threads = opt_.host_threads > 0 ? (std::size_t)opt_.host_threads
                                                    : std::max(1u, std::thread::hardware_concurrency());
        threads = std::min(threads, std::max<std::size_t>(1, rows.size() / 8));
        if (threads <= 1) {
            work(0, rows.size());
        } else {
            std::vector<std::thread> pool;
            const std::size_t per = (rows.size() + threads - 1) / threads;
            for (std::size_t t = 0; t < threads; ++t) {
                const std::size_t lo = t * per, hi = std::min(rows.size(), lo + per);
                if (lo < hi) pool.emplace_back(work, lo, hi);
            }
            for (auto& th : pool) th.join();
        }
        brkga_facade::check(brkga_population_host_end(pop_, fit.data(), nullptr), "brkga_population_host_end");
    }

    void sync() const
    {
        if (!dirty_) return;
        const int P = config_->population_size;
        const int L = config_->get_total_chromosome_length();
        const int M = config_->num_objectives;
        std::vector<float> keys((std::size_t)P * L), fit((std::size_t)P * M), crowd(P);
        std::vector<int32_t> order(P), rank(P, -1);
        brkga_facade::check(brkga_population_read(pop_, keys.data(), fit.data(), order.data(), rank.data(),
                                                  crowd.data(), nullptr),
                            "brkga_population_read");
        individuals_.clear();
        for (int i = 0; i < P; ++i) {
            Individual<T> ind = M > 1 ? Individual<T>(config_->component_lengths, M)
                                      : Individual<T>(config_->component_lengths);
            std::vector<T> row(keys.begin() + (std::size_t)i * L, keys.begin() + (std::size_t)(i + 1) * L);
            ind.unflatten(row);
            if (M > 1) {
                std::vector<T> obj(fit.begin() + (std::size_t)i * M, fit.begin() + (std::size_t)(i + 1) * M);
                ind.set_objectives(obj);
                ind.rank = rank[i];
                ind.crowding_distance = crowd[i];
            } else {
                ind.set_fitness(fit[i]);
            }
            individuals_.push_back(ind);
        }
        std::vector<float> bf(M), bk(L);
        brkga_facade::check(brkga_get_best(pop_, bf.data(), bk.data(), nullptr, nullptr), "brkga_get_best");
        best_ = M > 1 ? Individual<T>(config_->component_lengths, M) : Individual<T>(config_->component_lengths);
        best_.unflatten(std::vector<T>(bk.begin(), bk.end()));
        if (M > 1) best_.set_objectives(std::vector<T>(bf.begin(), bf.end()));
        else best_.set_fitness(bf[0]);
        dirty_ = false;
    }

    BRKGAConfig<T>* config_;
    BRKGAEngineOptions opt_;
    brkga_problem_t problem_ = nullptr;
    brkga_problem_t host_problem_ = nullptr;  // owned: configs evaluated on the host
    brkga_pop_t pop_ = nullptr;
    std::vector<std::vector<int>> fronts_;    // of the last host-view fast_non_dominated_sort
    std::vector<T> fitness_history_;
    mutable bool dirty_ = true;
    mutable std::vector<Individual<T>> individuals_;
    mutable Individual<T> best_{std::vector<int>{1}};
};
