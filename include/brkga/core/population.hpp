// brkga/core/population.hpp - device-resident population of the C++ facade.
//
// Interface of the reference's Population<T> (core/population.hpp:21-1097)
// and of its device-resident sibling GPUPopulation<T>
// (core/gpu_population.cuh:52-508), backed by an engine population
// (brkga_pop_t).  The keys, fitness and order live in device memory; host
// views (get_individuals, get_best, get_pareto_front) are materialised on
// demand, as GPUPopulation::sync_to_host does.
#pragma once

#include <cmath>
#include <cstdint>
#include <iomanip>
#include <iostream>
#include <set>
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
    int diversity_interval = 25;
};

template <typename T>
class Population {
public:
    Population(BRKGAConfig<T>* cfg, const BRKGAEngineOptions& opt = BRKGAEngineOptions()) : config_(cfg), opt_(opt)
    {
        problem_ = cfg->native_problem(opt.device);
        if (!problem_)
            throw std::runtime_error("config has no engine decoder (native_kind() < 0); the engine runs the "
                                     "TSP / TSPJ / knapsack / bi-objective TSP configs");
        brkga_params p{};
        p.population_size = cfg->population_size;
        p.elite_size = cfg->elite_size;
        p.mutant_size = cfg->mutant_size;
        p.elite_prob = cfg->elite_prob;
        p.num_islands = opt.num_islands;
        p.migration_interval = opt.migration_interval;
        p.num_migrants = opt.num_migrants;
        p.seed = opt.seed;
        p.diversity_interval = opt.diversity_interval;
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
    }

    Population(const Population&) = delete;
    Population& operator=(const Population&) = delete;

    // population.hpp:107-124
    void initialize()
    {
        brkga_facade::check(brkga_population_initialize(pop_, nullptr), "brkga_population_initialize");
        fitness_history_.clear();
        dirty_ = true;
        record_fitness();
    }

    // One or more generation steps on the device.
    void evolve(int generations)
    {
        if (generations <= 0) return;
        if (opt_.use_graph)
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

private:
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
    brkga_pop_t pop_ = nullptr;
    std::vector<T> fitness_history_;
    mutable bool dirty_ = true;
    mutable std::vector<Individual<T>> individuals_;
    mutable Individual<T> best_{std::vector<int>{1}};
};
