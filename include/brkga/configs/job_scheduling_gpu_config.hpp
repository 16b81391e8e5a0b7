// brkga/configs/job_scheduling_gpu_config.hpp - JobSchedulingGPUConfig<T>
// of the C++ facade.
//
// Constructor, test-instance factory and accessors of the reference's
// JobSchedulingGPUConfig (configs/job_scheduling_gpu_config.hpp:17-245): job
// i runs on machine int(key_i * machines) (clamped), fitness = largest
// machine load, minimised.  Any machine count up to 4096 runs on the engine
// (the reference's GPU kernel keeps 32 machine loads).
#pragma once

#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../core/config.hpp"

template <typename T>
class JobSchedulingGPUConfig : public BRKGAConfig<T> {
public:
    JobSchedulingGPUConfig(const std::vector<T>& jobs, int machines, const std::string& name = "JobScheduling_GPU")
        : BRKGAConfig<T>(std::vector<int>{static_cast<int>(jobs.size())}), job_times_(jobs), num_machines_(machines),
          instance_name_(name)
    {
        if (num_machines_ <= 0) throw std::invalid_argument("Number of machines must be positive");
        this->fitness_function = [this](const Individual<T>& ind) { return calculate_scheduling_fitness(ind); };
        this->decoder = [this](const Individual<T>& ind) { return decode_to_solution(ind); };
        this->comparator = [](T a, T b) { return a < b; };  // makespan
        this->threads_per_block = 256;
        this->update_cuda_grid_size();
    }

    // configs/job_scheduling_gpu_config.hpp:138-158
    T calculate_scheduling_fitness(const Individual<T>& individual)
    {
        std::vector<float> f;
        this->host_evaluate(individual, &f, nullptr);
        return static_cast<T>(f[0]);
    }

    // configs/job_scheduling_gpu_config.hpp:160-174: machine of every job
    std::vector<std::vector<T>> decode_to_solution(const Individual<T>& individual)
    {
        std::vector<int32_t> m;
        this->host_evaluate(individual, nullptr, &m);
        std::vector<std::vector<T>> out(1);
        out[0].assign(m.begin(), m.end());
        return out;
    }

    // configs/job_scheduling_gpu_config.hpp:176-212
    void print_solution(const Individual<T>& individual) override
    {
        const auto machine = decode_to_solution(individual)[0];
        std::vector<T> load(num_machines_, T(0));
        std::vector<std::vector<int>> jobs(num_machines_);
        for (std::size_t i = 0; i < machine.size(); ++i) {
            const int m = static_cast<int>(machine[i]);
            load[m] += job_times_[i];
            jobs[m].push_back(static_cast<int>(i));
        }
        std::cout << "\n=== Job Scheduling GPU Solution ===\nInstance: " << instance_name_
                  << "\nMakespan: " << individual.fitness << std::endl;
        for (int m = 0; m < num_machines_; ++m) {
            std::cout << "Machine " << m << " (load " << load[m] << "): ";
            for (std::size_t j = 0; j < jobs[m].size(); ++j)
                std::cout << jobs[m][j] << (j + 1 < jobs[m].size() ? ", " : "");
            if (load[m] == individual.fitness) std::cout << " <- CRITICAL";
            std::cout << std::endl;
        }
    }

    // configs/job_scheduling_gpu_config.hpp:214-228
    static std::unique_ptr<JobSchedulingGPUConfig<T>> create_test_instance(int num_jobs = 50, int num_machines = 5)
    {
        std::vector<T> jobs;
        jobs.reserve(num_jobs);
        for (int i = 0; i < num_jobs; ++i) jobs.push_back(static_cast<T>(10 + (i % 40)));
        return std::make_unique<JobSchedulingGPUConfig<T>>(
            jobs, num_machines, "Test_GPU_J" + std::to_string(num_jobs) + "_M" + std::to_string(num_machines));
    }

    bool is_gpu_available() const { return true; }
    int get_num_machines() const { return num_machines_; }
    const std::vector<T>& get_job_times() const { return job_times_; }

    int native_kind() const override { return BRKGA_JOB_SCHEDULING; }

protected:
    brkga_problem_t create_native_problem(brkga_engine_t engine) override
    {
        std::vector<float> t(job_times_.begin(), job_times_.end());
        brkga_problem_t p = nullptr;
        brkga_facade::check(brkga_problem_create_job_scheduling(engine, t.data(), (int32_t)t.size(), num_machines_, &p),
                            "brkga_problem_create_job_scheduling");
        return p;
    }

private:
    std::vector<T> job_times_;
    int num_machines_;
    std::string instance_name_;
};
