// brkga/configs/continuous_opt_gpu_config.hpp - ContinuousOptGPUConfig<T>
// of the C++ facade.
//
// Constructor, factories and accessors of the reference's
// ContinuousOptGPUConfig (configs/continuous_opt_gpu_config.hpp:21-225):
// sphere, Rastrigin and Rosenbrock over x_i = key_i (upper - lower) + lower,
// minimised.  All three run on the engine (the reference evaluates
// Rosenbrock on the host).  Rastrigin's cosine is the engine's deterministic
// FP32 cosine, within half an ulp of the exact value.
#pragma once

#include <algorithm>
#include <cstring>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "../core/config.hpp"

enum class OptimizationFunction { SPHERE, RASTRIGIN, ROSENBROCK };

template <typename T>
class ContinuousOptGPUConfig : public BRKGAConfig<T> {
public:
    ContinuousOptGPUConfig(int num_vars, T lower, T upper, OptimizationFunction func = OptimizationFunction::SPHERE,
                           const std::string& name = "ContinuousOpt_GPU")
        : BRKGAConfig<T>(std::vector<int>{num_vars}), num_variables_(num_vars), lower_bound_(lower),
          upper_bound_(upper), function_type_(func), instance_name_(name)
    {
        this->fitness_function = [this](const Individual<T>& ind) { return calculate_continuous_fitness(ind); };
        this->decoder = [this](const Individual<T>& ind) { return decode_to_solution(ind); };
        this->comparator = [](T a, T b) { return a < b; };  // minimisation
        this->threads_per_block = 256;
        this->update_cuda_grid_size();
    }

    // configs/continuous_opt_gpu_config.hpp:116-130
    T calculate_continuous_fitness(const Individual<T>& individual)
    {
        std::vector<float> f;
        this->host_evaluate(individual, &f, nullptr);
        return static_cast<T>(f[0]);
    }

    // configs/continuous_opt_gpu_config.hpp:162-175: the decision variables
    std::vector<std::vector<T>> decode_to_solution(const Individual<T>& individual)
    {
        std::vector<int32_t> bits;
        this->host_evaluate(individual, nullptr, &bits);
        std::vector<std::vector<T>> out(1);
        out[0].reserve(bits.size());
        for (int32_t b : bits) {
            float x;
            std::memcpy(&x, &b, sizeof(float));
            out[0].push_back(static_cast<T>(x));
        }
        return out;
    }

    void print_solution(const Individual<T>& individual) override
    {
        const auto vars = decode_to_solution(individual)[0];
        const char* names[] = {"Sphere", "Rastrigin", "Rosenbrock"};
        std::cout << "\n=== " << names[static_cast<int>(function_type_)] << " GPU Solution ===\nInstance: "
                  << instance_name_ << "\nVariables: " << vars.size() << "\nRange: [" << lower_bound_ << ", "
                  << upper_bound_ << "]\nFitness: " << individual.fitness << "\n\nDecision variables (first 10):"
                  << std::endl;
        for (std::size_t i = 0; i < std::min<std::size_t>(10, vars.size()); ++i)
            std::cout << "  x[" << i << "] = " << vars[i] << std::endl;
        if (vars.size() > 10) std::cout << "  ... (" << (vars.size() - 10) << " more variables)" << std::endl;
        std::cout << "GPU Evaluation: Enabled\n=====================================" << std::endl;
    }

    // configs/continuous_opt_gpu_config.hpp:199-225
    static std::unique_ptr<ContinuousOptGPUConfig<T>> create_sphere(int num_vars = 30)
    {
        return std::make_unique<ContinuousOptGPUConfig<T>>(num_vars, T(-5.12), T(5.12), OptimizationFunction::SPHERE,
                                                           "Sphere_GPU_" + std::to_string(num_vars));
    }
    static std::unique_ptr<ContinuousOptGPUConfig<T>> create_rastrigin(int num_vars = 30)
    {
        return std::make_unique<ContinuousOptGPUConfig<T>>(num_vars, T(-5.12), T(5.12),
                                                           OptimizationFunction::RASTRIGIN,
                                                           "Rastrigin_GPU_" + std::to_string(num_vars));
    }
    static std::unique_ptr<ContinuousOptGPUConfig<T>> create_rosenbrock(int num_vars = 30)
    {
        return std::make_unique<ContinuousOptGPUConfig<T>>(num_vars, T(-5.0), T(10.0),
                                                           OptimizationFunction::ROSENBROCK,
                                                           "Rosenbrock_GPU_" + std::to_string(num_vars));
    }

    bool is_gpu_available() const { return true; }
    OptimizationFunction get_function() const { return function_type_; }
    T get_lower_bound() const { return lower_bound_; }
    T get_upper_bound() const { return upper_bound_; }

    int native_kind() const override { return BRKGA_CONTINUOUS; }

protected:
    brkga_problem_t create_native_problem(brkga_engine_t engine) override
    {
        brkga_problem_t p = nullptr;
        brkga_facade::check(brkga_problem_create_continuous(engine, static_cast<int32_t>(function_type_),
                                                            num_variables_, static_cast<float>(lower_bound_),
                                                            static_cast<float>(upper_bound_), &p),
                            "brkga_problem_create_continuous");
        return p;
    }

private:
    int num_variables_;
    T lower_bound_, upper_bound_;
    OptimizationFunction function_type_;
    std::string instance_name_;
};
