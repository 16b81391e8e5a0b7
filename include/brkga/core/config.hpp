// brkga/core/config.hpp - problem plug-in base of the C++ facade.
//
// Same fields, constructors, virtual hooks and helpers as the reference's
// BRKGAConfig<T> (core/config.hpp:14-229), so a config written against that
// interface compiles unchanged.  What changes is underneath:
//   * native configs (TSPConfig, KnapsackConfig, TSPJConfig, MultiTSPConfig)
//     hand the engine a device-resident problem through native_problem();
//     the whole generation step then runs on the GPU;
//   * evaluate_population_gpu() of a native config calls brkga_evaluate()
//     (include/brkga_hip.h) instead of a CUDA kernel.
// Errors reported by the engine are raised as std::runtime_error, as the
// reference core does (core/cuda_kernels.cuh:250-262).
#pragma once

#include <algorithm>
#include <fstream>
#include <functional>
#include <iostream>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../brkga_hip.h"
#include "individual.hpp"

namespace brkga_facade {

inline void check(brkga_status st, const char* what)
{
    if (st != BRKGA_OK) throw std::runtime_error(std::string(what) + ": " + brkga_last_error());
}

// One engine per device for the whole process.
inline brkga_engine_t engine_for_device(int device)
{
    static std::mutex mu;
    static std::map<int, brkga_engine_t> engines;
    std::lock_guard<std::mutex> lock(mu);
    auto it = engines.find(device);
    if (it != engines.end()) return it->second;
    brkga_engine_t e = nullptr;
    check(brkga_engine_create(device, &e), "brkga_engine_create");
    engines[device] = e;
    return e;
}

}  // namespace brkga_facade

template <typename T>
class BRKGAConfig {
public:
    std::vector<int> component_lengths;
    int num_components;
    int population_size;
    int elite_size;
    int mutant_size;
    double elite_prob;
    int max_generations;
    int threads_per_block;
    int blocks_per_grid;
    std::function<T(const Individual<T>&)> fitness_function;
    std::function<std::vector<std::vector<T>>(const Individual<T>&)> decoder;
    std::function<bool(T, T)> comparator;
    int num_objectives;
    std::vector<std::function<T(const Individual<T>&)>> objective_functions;
    void* d_fitness_func;
    void* d_decoder_func;

    explicit BRKGAConfig(int chrom_len) : BRKGAConfig(std::vector<int>{chrom_len}) {}

    explicit BRKGAConfig(const std::vector<int>& comp_lengths)
        : component_lengths(comp_lengths), num_components((int)comp_lengths.size()), population_size(100),
          elite_size(20), mutant_size(10), elite_prob(0.7), max_generations(1000), threads_per_block(256),
          blocks_per_grid(32), num_objectives(1), d_fitness_func(nullptr), d_decoder_func(nullptr)
    {
    }

    BRKGAConfig(const std::vector<int>& comp_lengths, int n_objectives) : BRKGAConfig(comp_lengths)
    {
        num_objectives = n_objectives;
        objective_functions.resize(n_objectives);
    }

    virtual ~BRKGAConfig()
    {
        for (auto& kv : native_) brkga_problem_destroy(kv.second);
    }

    // ----- GPU hooks (same signatures as the reference) -----
    virtual bool has_gpu_evaluation() const { return native_kind() >= 0; }

    // d_population[pop_size][chrom_len] and d_fitness[pop_size][num_objectives]
    // are device pointers on the current HIP device.
    virtual void evaluate_population_gpu(T* d_population, T* d_fitness, int pop_size, int chrom_len)
    {
        static_assert(sizeof(T) == sizeof(float), "the engine evaluates FP32 keys");
        brkga_problem_t p = native_problem(device_);
        if (!p) return;
        brkga_facade::check(brkga_evaluate(p, reinterpret_cast<const float*>(d_population),
                                           reinterpret_cast<float*>(d_fitness), pop_size, chrom_len, nullptr),
                            "brkga_evaluate");
    }

    virtual bool has_local_search() const { return false; }
    virtual int apply_local_search_gpu(T*, T*, T*, T*, void*, int, int, int, int) { return 0; }
    virtual void initialize_device_functions() {}
    virtual void cleanup_device_functions() {}

    // ----- engine hook -----
    // Device the config's data lives on (the reference follows the caller's
    // current CUDA device; the facade names it explicitly).
    void set_device(int device) { device_ = device; }
    int device() const { return device_; }

    // Kind of the engine decoder this config maps to, or -1 for a config the
    // engine cannot run (its fitness_function is then host code only).
    virtual int native_kind() const { return -1; }

    // Device-resident problem for `device`, created on first use.
    brkga_problem_t native_problem(int device)
    {
        if (native_kind() < 0) return nullptr;
        std::lock_guard<std::mutex> lock(native_mu_);
        auto it = native_.find(device);
        if (it != native_.end()) return it->second;
        brkga_problem_t p = create_native_problem(brkga_facade::engine_for_device(device));
        native_[device] = p;
        return p;
    }

    // ----- problem behaviour (overridden by configs) -----
    virtual void print_solution(const Individual<T>& ind)
    {
        if (num_objectives > 1) {
            std::cout << "Objectives: ";
            for (std::size_t i = 0; i < ind.objectives.size(); ++i)
                std::cout << ind.objectives[i] << (i + 1 < ind.objectives.size() ? ", " : "");
            std::cout << std::endl;
        } else {
            std::cout << "Best fitness: " << ind.fitness << std::endl;
        }
    }

    virtual bool validate_solution(const Individual<T>&) { return true; }

    virtual void export_solution(const Individual<T>& ind, const std::string& filename)
    {
        std::ofstream f(filename);
        if (num_objectives > 1) {
            f << "Objectives: ";
            for (std::size_t i = 0; i < ind.objectives.size(); ++i)
                f << ind.objectives[i] << (i + 1 < ind.objectives.size() ? " " : "");
            f << std::endl;
        } else {
            f << "Fitness: " << ind.fitness << std::endl;
        }
        f << "Chromosome: ";
        for (int c = 0; c < num_components; ++c) {
            const auto& comp = ind.get_component(c);
            f << "Component " << c << ": ";
            for (std::size_t i = 0; i < comp.size(); ++i) f << comp[i] << (i + 1 < comp.size() ? " " : "");
            if (c + 1 < num_components) f << " | ";
        }
        f << std::endl;
    }

    // core/config.hpp:159-179
    virtual bool is_valid() const
    {
        const bool lengths_ok = !component_lengths.empty() &&
                                std::all_of(component_lengths.begin(), component_lengths.end(),
                                            [](int len) { return len > 0; });
        if (!lengths_ok || population_size <= 0 || elite_prob < 0.0 || elite_prob > 1.0 || max_generations <= 0 ||
            num_objectives <= 0)
            return false;
        if (num_objectives == 1) return elite_size > 0 && mutant_size >= 0 && elite_size + mutant_size <= population_size;
        return true;
    }

    int get_total_chromosome_length() const
    {
        int t = 0;
        for (int len : component_lengths) t += len;
        return t;
    }
    int chromosome_length() const { return get_total_chromosome_length(); }
    int get_offspring_size() const { return population_size - elite_size - mutant_size; }
    void update_cuda_grid_size() { blocks_per_grid = (population_size + threads_per_block - 1) / threads_per_block; }
    bool is_multi_objective() const { return num_objectives > 1; }

    void print_config() const
    {
        std::cout << "Configuration:\n  Mode: " << (is_multi_objective() ? "Multi-objective" : "Single-objective")
                  << "\n  Components: " << num_components << "\n  Total chromosome length: "
                  << get_total_chromosome_length() << "\n  Population size: " << population_size
                  << "\n  Elite size: " << elite_size << "\n  Mutant size: " << mutant_size
                  << "\n  Offspring size: " << get_offspring_size() << "\n  Elite probability: " << elite_prob
                  << "\n  Max generations: " << max_generations << std::endl;
    }

    // Decode + fitness of one host individual on the device (backs the
    // fitness_function / decoder callbacks of the native configs; the
    // reference runs them on the host).  fitness gets num_objectives values,
    // decoded the decoder's integer output (chromosome length).
    void host_evaluate(const Individual<T>& ind, std::vector<float>* fitness, std::vector<int32_t>* decoded)
    {
        brkga_problem_t p = native_problem(device_);
        if (!p) throw std::runtime_error("config has no engine decoder");
        const std::vector<T> row = ind.flatten();
        std::vector<float> keys(row.begin(), row.end());
        std::vector<float> fit(num_objectives);
        std::vector<int32_t> dec(keys.size());
        brkga_facade::check(brkga_evaluate_host(p, keys.data(), 1, fit.data(), dec.data()), "brkga_evaluate_host");
        if (fitness) *fitness = fit;
        if (decoded) *decoded = dec;
    }

protected:
    // Native configs upload their data here (called once per device).
    virtual brkga_problem_t create_native_problem(brkga_engine_t) { return nullptr; }

private:
    int device_ = 0;
    std::mutex native_mu_;
    std::map<int, brkga_problem_t> native_;
};
