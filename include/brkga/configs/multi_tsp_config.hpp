// brkga/configs/multi_tsp_config.hpp - MultiTSPConfig<T> of the C++ facade.
//
// Bi-objective TSP of the reference (configs/multi_tsp_config.hpp:17-290):
// depot 0 + n cities over a distance matrix and a time matrix, both
// objectives minimised, solved by NSGA-II.  from_coordinates() takes two
// coordinate sets instead (objective m = tour length in set m).
#pragma once

#include <cmath>
#include <memory>
#include <random>
#include <string>
#include <utility>
#include <vector>

#include "../core/config.hpp"
#include "tsp_config.hpp"

template <typename T>
class MultiTSPConfig : public BRKGAConfig<T> {
public:
    MultiTSPConfig(const std::vector<std::vector<T>>& distances, const std::vector<std::vector<T>>& times,
                   const std::string& name = "MultiTSP")
        : BRKGAConfig<T>(std::vector<int>{static_cast<int>(distances.size() - 1)}, 2), distance_matrix_(distances),
          time_matrix_(times), num_cities_(static_cast<int>(distances.size() - 1)), instance_name_(name)
    {
        setup();
    }

    static std::unique_ptr<MultiTSPConfig<T>> from_coordinates(const std::vector<std::pair<T, T>>& xy0,
                                                               const std::vector<std::pair<T, T>>& xy1,
                                                               const std::string& name = "MultiTSP")
    {
        if (xy0.size() != xy1.size() || xy0.size() < 2)
            throw std::invalid_argument("two coordinate sets of n + 1 points (depot first)");
        std::unique_ptr<MultiTSPConfig<T>> c(new MultiTSPConfig<T>(static_cast<int>(xy0.size() - 1), name));
        c->xy0_ = xy0;
        c->xy1_ = xy1;
        return c;
    }

    // configs/multi_tsp_config.hpp:54-67
    T calculate_tour_distance(const Individual<T>& individual)
    {
        std::vector<float> f;
        this->host_evaluate(individual, &f, nullptr);
        return static_cast<T>(f[0]);
    }

    // configs/multi_tsp_config.hpp:69-82
    T calculate_tour_time(const Individual<T>& individual)
    {
        std::vector<float> f;
        this->host_evaluate(individual, &f, nullptr);
        return static_cast<T>(f[1]);
    }

    // configs/multi_tsp_config.hpp:84-103: 1-based city order after the depot
    std::vector<int> decode_tour(const Individual<T>& individual)
    {
        std::vector<int32_t> d;
        this->host_evaluate(individual, nullptr, &d);
        std::vector<int> tour;
        for (int32_t c : d) tour.push_back(c + 1);
        return tour;
    }

    // configs/multi_tsp_config.hpp:272-285
    static void configure_for_size(MultiTSPConfig<T>* config, int num_cities)
    {
        if (num_cities <= 20) {
            config->population_size = 100;
            config->max_generations = 250;
        } else if (num_cities <= 50) {
            config->population_size = 200;
            config->max_generations = 500;
        } else {
            config->population_size = 300;
            config->max_generations = 1000;
        }
        config->update_cuda_grid_size();
    }

    // configs/multi_tsp_config.hpp:190-202: TSPLIB coordinates (city 0 is
    // the depot) -> distance matrix; the time matrix scales every off-
    // diagonal distance by a speed factor drawn from std::mt19937(42) and
    // std::uniform_real_distribution<T>(0.7, 1.8) in row-major order.
    static std::unique_ptr<MultiTSPConfig<T>> load_from_file(const std::string& filename)
    {
        const auto dist = load_tsplib(filename).first;
        return std::make_unique<MultiTSPConfig<T>>(dist, generate_time_matrix(dist, T(0.7), T(1.8)),
                                                   FileUtils::get_basename(filename));
    }

    // configs/multi_tsp_config.hpp:204-269: FP32 Euclidean matrix (zero
    // diagonal) of the TSPLIB coordinates, and the NAME field.
    static std::pair<std::vector<std::vector<T>>, std::string> load_tsplib(const std::string& filename)
    {
        std::string name;
        std::vector<std::pair<T, T>> xy;
        TSPConfig<T>::read_tsplib_coords(filename, &name, &xy);
        const std::size_t n = xy.size();
        std::vector<std::vector<T>> d(n, std::vector<T>(n, T(0)));
        for (std::size_t i = 0; i < n; ++i)
            for (std::size_t j = 0; j < n; ++j)
                if (i != j) {
                    const T dx = xy[i].first - xy[j].first;
                    const T dy = xy[i].second - xy[j].second;
                    const T sx = dx * dx;
                    const T sy = dy * dy;
                    d[i][j] = std::sqrt(sx + sy);
                }
        return {d, name};
    }

    // configs/multi_tsp_config.hpp:165-188
    static std::vector<std::vector<T>> generate_time_matrix(const std::vector<std::vector<T>>& dist,
                                                            T speed_min = T(0.8), T speed_max = T(1.5))
    {
        const std::size_t n = dist.size();
        std::vector<std::vector<T>> t(n, std::vector<T>(n, T(0)));
        std::mt19937 rng(42);
        std::uniform_real_distribution<T> speed(speed_min, speed_max);
        for (std::size_t i = 0; i < n; ++i)
            for (std::size_t j = 0; j < n; ++j)
                if (i != j) t[i][j] = dist[i][j] * speed(rng);
        return t;
    }

    int get_num_cities() const { return num_cities_; }
    const std::string& get_instance_name() const { return instance_name_; }

    void print_instance_info() const
    {
        std::cout << "=== Multi-Objective TSP ===\nName: " << instance_name_ << "\nCities: " << num_cities_
                  << "\nObjectives: distance, time\n===========================" << std::endl;
    }

    int native_kind() const override { return xy0_.empty() ? BRKGA_BITSP_MATRIX : BRKGA_BITSP_COORDS; }

protected:
    brkga_problem_t create_native_problem(brkga_engine_t engine) override
    {
        brkga_problem_t p = nullptr;
        if (!xy0_.empty()) {
            std::vector<float> a, b;
            for (const auto& c : xy0_) {
                a.push_back(static_cast<float>(c.first));
                a.push_back(static_cast<float>(c.second));
            }
            for (const auto& c : xy1_) {
                b.push_back(static_cast<float>(c.first));
                b.push_back(static_cast<float>(c.second));
            }
            brkga_facade::check(brkga_problem_create_bitsp_coords(engine, a.data(), b.data(), num_cities_, &p),
                                "brkga_problem_create_bitsp_coords");
        } else {
            std::vector<float> a, b;
            for (const auto& row : distance_matrix_)
                for (T v : row) a.push_back(static_cast<float>(v));
            for (const auto& row : time_matrix_)
                for (T v : row) b.push_back(static_cast<float>(v));
            brkga_facade::check(brkga_problem_create_bitsp_matrix(engine, a.data(), b.data(), num_cities_ + 1, &p),
                                "brkga_problem_create_bitsp_matrix");
        }
        return p;
    }

private:
    MultiTSPConfig(int n, const std::string& name)
        : BRKGAConfig<T>(std::vector<int>{n}, 2), num_cities_(n), instance_name_(name)
    {
        setup();
    }

    void setup()
    {
        this->population_size = 100;
        this->elite_size = 0;
        this->mutant_size = 0;
        this->max_generations = 500;
        this->elite_prob = 0.7;
        this->objective_functions[0] = [this](const Individual<T>& ind) { return calculate_tour_distance(ind); };
        this->objective_functions[1] = [this](const Individual<T>& ind) { return calculate_tour_time(ind); };
        this->update_cuda_grid_size();
    }

    std::vector<std::vector<T>> distance_matrix_, time_matrix_;
    std::vector<std::pair<T, T>> xy0_, xy1_;
    int num_cities_;
    std::string instance_name_;
};
