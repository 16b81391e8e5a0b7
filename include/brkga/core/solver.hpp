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
    // multi-objective configs: diversity injection every `interval` steps
    // (solver.hpp:812-814 uses 25; 0 = off)
    void set_diversity_interval(int interval) { options_.diversity_interval = interval; }
    // configs evaluated on the host: threads for the callbacks (0 = one per
    // hardware thread, 1 = serial)
    void set_host_threads(int threads) { options_.host_threads = threads; }
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
        const auto t0 = std::chrono::steady_clock::now();
        population_->evolve(1);
        population_->generation();  // waits for the step
        step_seconds_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        ++steps_timed_;
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
            const auto c0 = std::chrono::steady_clock::now();
            population_->evolve(chunk);
            population_->generation();  // waits for the chunk
            step_seconds_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - c0).count();
            steps_timed_ += chunk;
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

    // Solver::benchmark_modes (solver.hpp:1671-1724) compares one GPU with
    // several over 10 generations.  This build runs on one GPU, so the two
    // forms compared are its own: the generations launched one by one and
    // replayed from a captured HIP graph (configs evaluated on the host have
    // one form only).  The population is initialised again afterwards.
    void benchmark_modes()
    {
        std::cout << "\n=== Launch-form benchmark (one GPU) ===" << std::endl;
        const bool keep = options_.use_graph;
        double secs[2] = {0.0, 0.0};
        for (int graph = 0; graph < 2; ++graph) {
            options_.use_graph = graph != 0;
            population_.reset();
            population_ = std::make_unique<Population<T>>(config_.get(), options_);
            population_->initialize();
            population_->evolve(1);  // warm-up (captures the graph)
            population_->generation();
            const auto t0 = std::chrono::steady_clock::now();
            population_->evolve(10);
            population_->generation();
            secs[graph] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (population_->host_evaluated()) {
                secs[1] = secs[0];
                break;
            }
        }
        std::cout << "Benchmark Results (10 generations):\n  Launches:    " << std::fixed << std::setprecision(3)
                  << secs[0] << " seconds\n  Graph:       " << secs[1] << " seconds\n  Speedup:     "
                  << std::setprecision(2) << (secs[1] > 0.0 ? secs[0] / secs[1] : 0.0) << "x" << std::endl;
        std::cout << "===========================" << std::endl;
        options_.use_graph = keep;
        initialize();
    }

    // Solver::print_multi_gpu_performance (solver.hpp:1591-1625): time
    // spent in generation steps so far, in total and per generation, on the
    // one GPU this build runs on (evaluation, crossover and sorting are one
    // fused step on the device, so they are not split).
    void print_multi_gpu_performance() const
    {
        std::cout << "\n=== GPU Performance Statistics ===\nActive GPUs: 1 (device " << options_.device << ")\n";
        std::cout << "Generation steps: " << steps_timed_ << ", " << std::fixed << std::setprecision(4) << step_seconds_
                  << " s";
        if (steps_timed_ > 0)
            std::cout << " (" << std::setprecision(4) << step_seconds_ / steps_timed_ * 1000.0 << " ms per generation)";
        std::cout << "\n=========================================" << std::endl;
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
    double step_seconds_ = 0.0;  // wall time of generation steps (run, evolve_generation)
    long long steps_timed_ = 0;
};
