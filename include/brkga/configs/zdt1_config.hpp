// brkga/configs/zdt1_config.hpp - ZDT1Config<T> of the C++ facade.
//
// The reference's ZDT1 example config (examples/zdt1.cu:14-100): n
// variables in [0, 1] (the keys), objectives f1 = x1 and
// f2 = g (1 - sqrt(f1 / g)) with g = 1 + 9 sum(x2..xn) / (n - 1), both
// minimised on NSGA-II; true front f2 = 1 - sqrt(f1).
#pragma once

#include <iomanip>
#include <iostream>
#include <vector>

#include "../core/config.hpp"

template <typename T>
class ZDT1Config : public BRKGAConfig<T> {
public:
    explicit ZDT1Config(int n_vars = 30) : BRKGAConfig<T>(std::vector<int>{n_vars}, 2), num_variables_(n_vars)
    {
        this->population_size = 100;
        this->elite_size = 0;   // not used by NSGA-II
        this->mutant_size = 0;  // not used by NSGA-II
        this->max_generations = 5;
        this->elite_prob = 0.5;  // crossover bias
        this->objective_functions[0] = [this](const Individual<T>& ind) { return objective_f1(ind); };
        this->objective_functions[1] = [this](const Individual<T>& ind) { return objective_f2(ind); };
        this->update_cuda_grid_size();
    }

    T objective_f1(const Individual<T>& individual) { return objectives(individual)[0]; }
    T objective_f2(const Individual<T>& individual) { return objectives(individual)[1]; }

    void print_solution(const Individual<T>& individual) override
    {
        std::cout << "\n=== Solution ===\nObjective 1 (f1): " << std::fixed << std::setprecision(6)
                  << individual.objectives[0] << "\nObjective 2 (f2): " << individual.objectives[1]
                  << "\n================" << std::endl;
    }

    // examples/zdt1.cu:75-82: every variable in [0, 1]
    bool validate_solution(const Individual<T>& individual) override
    {
        for (const auto& v : individual.get_chromosome())
            if (v < 0.0 || v > 1.0) return false;
        return true;
    }

    void print_instance_info() const
    {
        std::cout << "=== ZDT1 Benchmark Problem ===\nVariables: " << num_variables_
                  << "\nObjectives: 2 (minimize)\nTrue Pareto front: f2 = 1 - sqrt(f1)\n"
                  << "===============================" << std::endl;
    }

    int native_kind() const override { return BRKGA_ZDT1; }

protected:
    brkga_problem_t create_native_problem(brkga_engine_t engine) override
    {
        brkga_problem_t p = nullptr;
        brkga_facade::check(brkga_problem_create_zdt1(engine, num_variables_, &p), "brkga_problem_create_zdt1");
        return p;
    }

private:
    std::vector<float> objectives(const Individual<T>& individual)
    {
        std::vector<float> f;
        this->host_evaluate(individual, &f, nullptr);
        return f;
    }

    int num_variables_;
};
