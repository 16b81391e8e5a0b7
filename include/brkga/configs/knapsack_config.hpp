// brkga/configs/knapsack_config.hpp - KnapsackConfig<T> of the C++ facade.
//
// Constructor, callbacks, size presets and accessors of the reference's
// KnapsackConfig (configs/knapsack_config.hpp:11-380).  The default decoder
// is the reference's: select key > 0.5, penalty 1000 per unit over
// capacity (configs/knapsack_config.hpp:43-75).  set_greedy(true) switches
// to packing items in ascending key order while they fit.  The engine sums
// exact integers: weights and values must be integral, and so must the
// capacity for the threshold decoder.
#pragma once

#include <cmath>
#include <iostream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../core/config.hpp"
#include "../utils/file_utils.hpp"

template <typename T>
class KnapsackConfig : public BRKGAConfig<T> {
public:
    KnapsackConfig(const std::vector<T>& w, const std::vector<T>& v, T cap, const std::string& name = "Knapsack")
        : BRKGAConfig<T>(std::vector<int>{static_cast<int>(w.size())}), weights_(w), values_(v), capacity_(cap),
          instance_name_(name)
    {
        if (weights_.size() != values_.size()) throw std::invalid_argument("Weights and values must have same size");
        for (std::size_t i = 0; i < w.size(); ++i)
            if (w[i] != std::floor(w[i]) || v[i] != std::floor(v[i]))
                throw std::invalid_argument("the engine's knapsack needs integral weights and values");

        this->fitness_function = [this](const Individual<T>& ind) { return calculate_knapsack_fitness(ind); };
        this->decoder = [this](const Individual<T>& ind) { return decode_to_solution(ind); };
        this->comparator = [](T a, T b) { return a > b; };  // maximisation
        this->threads_per_block = 256;
        this->update_cuda_grid_size();
    }

    // Greedy decoder (items in ascending key order, taken when they fit).
    // Must be chosen before the first population is created.
    void set_greedy(bool on) { greedy_ = on; }
    bool is_greedy() const { return greedy_; }

    // configs/knapsack_config.hpp:43-62
    T calculate_knapsack_fitness(const Individual<T>& individual)
    {
        std::vector<float> f;
        this->host_evaluate(individual, &f, nullptr);
        return static_cast<T>(f[0]);
    }

    // configs/knapsack_config.hpp:64-75: the 0/1 selection as one component
    std::vector<std::vector<T>> decode_to_solution(const Individual<T>& individual)
    {
        std::vector<std::vector<T>> out(1);
        out[0] = decode_to_selection(individual);
        return out;
    }

    std::vector<T> decode_to_selection(const Individual<T>& individual)
    {
        std::vector<int32_t> d;
        this->host_evaluate(individual, nullptr, &d);
        return std::vector<T>(d.begin(), d.end());
    }

    // configs/knapsack_config.hpp:307-333
    static void configure_for_size(KnapsackConfig<T>* config, int num_items)
    {
        if (num_items <= 50) {
            config->population_size = 200;
            config->elite_size = 40;
            config->mutant_size = 20;
            config->max_generations = 300;
        } else if (num_items <= 100) {
            config->population_size = 400;
            config->elite_size = 80;
            config->mutant_size = 40;
            config->max_generations = 500;
        } else if (num_items <= 500) {
            config->population_size = 800;
            config->elite_size = 160;
            config->mutant_size = 80;
            config->max_generations = 1000;
        } else {
            config->population_size = 1200;
            config->elite_size = 240;
            config->mutant_size = 120;
            config->max_generations = 1500;
        }
        config->elite_prob = 0.7;
        config->update_cuda_grid_size();
    }

    // configs/knapsack_config.hpp:180-192: ".csv" files as CSV, anything
    // else as the text format
    static std::unique_ptr<KnapsackConfig<T>> load_from_file(const std::string& filename)
    {
        if (!FileUtils::file_exists(filename)) throw std::runtime_error("File does not exist: " + filename);
        return FileUtils::get_extension(filename) == "csv" ? load_from_csv(filename) : load_from_text_file(filename);
    }

    // CSV (configs/knapsack_config.hpp:195-248): "weight,value[,capacity]"
    // lines ('#' comments); the first parsable third field is the capacity,
    // otherwise the capacity is half the total weight.
    static std::unique_ptr<KnapsackConfig<T>> load_from_csv(const std::string& filename)
    {
        std::vector<T> w, v;
        T cap = T(0);
        bool have_cap = false;
        for (const std::string& line : FileUtils::read_lines(filename)) {
            if (line.empty() || line[0] == '#') continue;
            const std::vector<std::string> f = FileUtils::parse_csv_line(line);
            if (f.size() < 2) continue;
            if (!have_cap && f.size() >= 3) {
                try {
                    cap = static_cast<T>(std::stod(f[2]));
                    have_cap = true;
                } catch (...) {
                }
            }
            try {
                const T a = static_cast<T>(std::stod(f[0]));
                const T b = static_cast<T>(std::stod(f[1]));
                w.push_back(a);
                v.push_back(b);
            } catch (...) {
            }
        }
        if (w.empty()) throw std::runtime_error("No valid items found in CSV file");
        if (!have_cap) {
            T total = T(0);
            for (T x : w) total += x;
            cap = total * 0.5;
        }
        return std::make_unique<KnapsackConfig<T>>(w, v, cap, FileUtils::get_basename(filename));
    }

    // Text (configs/knapsack_config.hpp:250-304): "num_items capacity", then
    // "weight value" lines ('#' comments).
    static std::unique_ptr<KnapsackConfig<T>> load_from_text_file(const std::string& filename)
    {
        std::vector<T> w, v;
        T cap = T(0);
        int count = 0;
        bool header = false;
        for (const std::string& line : FileUtils::read_lines(filename)) {
            if (line.empty() || line[0] == '#') continue;
            std::istringstream in(line);
            if (!header) {
                if (in >> count >> cap) header = true;
                continue;
            }
            T a, b;
            if (in >> a >> b) {
                w.push_back(a);
                v.push_back(b);
            }
        }
        if (w.empty()) throw std::runtime_error("No items found in file");
        if ((int)w.size() != count)
            std::cout << "Warning: Item count (" << w.size() << ") doesn't match expected (" << count << ")"
                      << std::endl;
        return std::make_unique<KnapsackConfig<T>>(w, v, cap, FileUtils::get_basename(filename));
    }

    const std::vector<T>& get_weights() const { return weights_; }
    const std::vector<T>& get_values() const { return values_; }
    T get_capacity() const { return capacity_; }
    const std::string& get_instance_name() const { return instance_name_; }

    void print_instance_info() const
    {
        std::cout << "=== Knapsack Instance Info ===\nName: " << instance_name_ << "\nItems: " << weights_.size()
                  << "\nCapacity: " << capacity_ << "\n==============================" << std::endl;
    }

    int native_kind() const override { return greedy_ ? BRKGA_KNAPSACK_GREEDY : BRKGA_KNAPSACK_THRESHOLD; }

protected:
    // With integral weights the greedy fill only compares sums against the
    // capacity, so floor(capacity) is exact; the threshold decoder's penalty
    // uses the capacity itself, which must then be integral.
    brkga_problem_t create_native_problem(brkga_engine_t engine) override
    {
        if (!greedy_ && capacity_ != std::floor(capacity_))
            throw std::invalid_argument("the threshold knapsack decoder needs an integral capacity");
        std::vector<int32_t> w(weights_.begin(), weights_.end()), v(values_.begin(), values_.end());
        brkga_problem_t p = nullptr;
        brkga_facade::check(brkga_problem_create_knapsack(engine, w.data(), v.data(), (int32_t)w.size(),
                                                          (int64_t)std::floor(capacity_), greedy_ ? 1 : 0, &p),
                            "brkga_problem_create_knapsack");
        return p;
    }

private:
    std::vector<T> weights_, values_;
    T capacity_;
    std::string instance_name_;
    bool greedy_ = false;
};
