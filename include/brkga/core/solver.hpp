// brkga/core/solver.hpp - Solver<T> of the C++ facade.
//
// Constructor, run() / initialize() / evolve_generation() and the result
// accessors of the reference's Solver<T> (core/solver.hpp:160-231,
// 1517-1669).  The generation loop runs entirely on the GPU: BRKGA for
// single-objective configs, NSGA-II for multi-objective ones; the island
// model and ring migration run on one device.  Settings the reference
// hard-codes (seed, islands, migration interval / count) have setters.
#pragma once

#include <chrono>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "population.hpp"

template <typename T>
class Solver {
public:
    Solver(std::unique_ptr<BRKGAConfig<T>> cfg, bool verbose = true, int print_freq = 50)
        : config_(std::move(cfg)), verbose_(verbose), print_frequency_(print_freq > 0 ? print_freq : 50)
    {
        if (!config_ || !config_->is_valid()) throw std::runtime_error("Invalid configuration");
    }

    void set_seed(uint64_t seed) { options_.seed = seed; }
    void set_device(int device)
    {
        options_.device = device;
        config_->set_device(device);
    }
    void set_islands(int islands, int migration_interval, int migrants)
    {
        options_.num_islands = islands;
        options_.migration_interval = migration_interval;
        options_.num_migrants = migrants;
    }
    void set_use_graph(bool on) { options_.use_graph = on; }
    // local search of configs that have one (solver.hpp:74-77,166)
    void set_local_search(bool on, int interval = 2, int individuals = 100, int moves = 5)
    {
        options_.local_search = on;
        options_.ls_interval = interval;
        options_.ls_individuals = individuals;
        options_.ls_moves = moves;
    }

    void initialize()
    {
        population_.reset();
        population_ = std::make_unique<Population<T>>(config_.get(), options_);
        population_->initialize();
        current_generation_ = 0;
        if (verbose_) {
            if (config_->is_multi_objective())
                std::cout << "Initial Pareto front size: " << population_->get_pareto_front().size() << std::endl;
            else
                std::cout << "Initial best fitness: " << std::fixed << std::setprecision(2)
                          << population_->get_best().fitness << std::endl;
        }
    }

    void evolve_generation()
    {
        if (!population_) initialize();
        population_->evolve(1);
        ++current_generation_;
        population_->record_fitness();
    }

    // Solver::run (core/solver.hpp:1517-1556): the loop is issued in chunks
    // of print_frequency generations; statistics are printed and the fitness
    // history recorded at chunk boundaries.
    void run()
    {
        const auto t0 = std::chrono::steady_clock::now();
        initialize();
        const int G = config_->max_generations;
        int gen = 0;
        while (gen < G) {
            const int chunk = std::min(print_frequency_, G - gen);
            population_->evolve(chunk);
            gen += chunk;
            current_generation_ = gen;
            population_->record_fitness();
            if (verbose_) population_->print_statistics(gen - 1);
        }
        elapsed_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (verbose_) print_final_results(elapsed_);
    }

    void print_final_results(double execution_time) const
    {
        std::cout << "Final Results:\n  Execution time: " << std::fixed << std::setprecision(2) << execution_time
                  << " seconds\n  Generations: " << current_generation_ << "/" << config_->max_generations << "\n";
        if (config_->is_multi_objective())
            std::cout << "  Pareto front size: " << population_->get_pareto_front().size() << std::endl;
        else
            std::cout << "  Best fitness: " << std::setprecision(6) << population_->get_best().fitness << std::endl;
    }

    const Individual<T>& get_best_individual() const { return population_->get_best(); }
    std::vector<Individual<T>> get_pareto_front() const { return population_->get_pareto_front(); }
    const std::vector<T>& get_fitness_history() const { return population_->get_fitness_history(); }
    BRKGAConfig<T>* get_config() { return config_.get(); }
    Population<T>& get_population() { return *population_; }
    int get_current_generation() const { return current_generation_; }
    double get_elapsed_seconds() const { return elapsed_; }

    // Solver::export_pareto_front (core/solver.hpp:1641-1669).
    void export_pareto_front(const std::string& filename)
    {
        if (!config_->is_multi_objective()) {
            std::cout << "Not a multi-objective problem" << std::endl;
            return;
        }
        auto front = get_pareto_front();
        std::ofstream f(filename);
        f << "# Pareto Front\n# Objectives: " << config_->num_objectives << "\n# Solutions: " << front.size()
          << "\n# Generated by NSGA-II on one GPU\n";
        for (const auto& ind : front) {
            for (std::size_t i = 0; i < ind.objectives.size(); ++i)
                f << ind.objectives[i] << (i + 1 < ind.objectives.size() ? " " : "");
            f << "\n";
        }
    }

private:
    std::unique_ptr<BRKGAConfig<T>> config_;
    bool verbose_;
    int print_frequency_;
    BRKGAEngineOptions options_;
    std::unique_ptr<Population<T>> population_;
    int current_generation_ = 0;
    double elapsed_ = 0.0;
};
