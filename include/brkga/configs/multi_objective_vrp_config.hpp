// brkga/configs/multi_objective_vrp_config.hpp - MultiObjectiveVRPConfig
// <Float> of the C++ facade.
//
// Loader, size presets, route decode, objectives and export of the
// reference's MultiObjectiveVRPConfig (configs/multi_objective_vrp_config
// .hpp:15-449): customers in biased key order (key + i 1e-5 + hash 1e-5,
// hash = running FP32 sum of key_i (i + 1)) split greedily into routes by
// vehicle capacity; objectives total distance (+100 per unit of overload)
// and longest route, each +500 per route beyond the vehicle count; NSGA-II.
//
// Objectives and the route of each customer come from the engine
// (brkga_problem_create_movrp); decode_routes orders each route's customers
// by the same biased keys on the host and sums its FP32 legs for listings.
#pragma once

#include <algorithm>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../core/config.hpp"

template <typename Float>
class MultiObjectiveVRPConfig : public BRKGAConfig<Float> {
public:
    struct Customer {
        int id;
        Float x, y;
        int demand;
    };

    struct Route {
        std::vector<int> customers;
        Float distance;
        int load;
    };

    MultiObjectiveVRPConfig() : BRKGAConfig<Float>(std::vector<int>{31}, 2) {}

    int get_num_customers() const { return num_customers_; }
    int get_num_vehicles() const { return num_vehicles_; }
    int get_vehicle_capacity() const { return vehicle_capacity_; }
    const std::vector<Customer>& get_customers() const { return customers_; }
    const std::vector<std::vector<Float>>& get_distances() const { return distances_; }

    // configs/multi_objective_vrp_config.hpp:69-175 (CVRPLIB: NAME,
    // DIMENSION, CAPACITY, NODE_COORD_SECTION, DEMAND_SECTION; vehicles from
    // "-k<v>" in the name, else max(3, n / 6))
    static std::unique_ptr<MultiObjectiveVRPConfig<Float>> load_from_file(const std::string& filename)
    {
        auto config = std::make_unique<MultiObjectiveVRPConfig<Float>>();
        std::ifstream file(filename);
        if (!file.is_open()) throw std::runtime_error("Cannot open VRP file: " + filename);
        std::string line;
        bool coords = false, demands = false;
        int dimension = 0;
        while (std::getline(file, line)) {
            if (line.find("NAME") != std::string::npos) {
                const std::size_t pos = line.find(':');
                if (pos != std::string::npos) {
                    config->instance_name_ = line.substr(pos + 1);
                    config->instance_name_.erase(0, config->instance_name_.find_first_not_of(" \t"));
                }
            } else if (line.find("DIMENSION") != std::string::npos) {
                std::istringstream iss(line);
                std::string key;
                char colon;
                iss >> key >> colon >> dimension;
                config->num_customers_ = dimension - 1;
            } else if (line.find("CAPACITY") != std::string::npos) {
                std::istringstream iss(line);
                std::string key;
                char colon;
                iss >> key >> colon >> config->vehicle_capacity_;
            } else if (line.find("NODE_COORD_SECTION") != std::string::npos) {
                coords = true;
                config->customers_.resize(dimension);
                break;
            }
        }
        if (coords)
            for (int i = 0; i < dimension; ++i) {
                int id;
                Float x, y;
                file >> id >> x >> y;
                config->customers_[id - 1] = {id - 1, x, y, 0};
            }
        while (std::getline(file, line))
            if (line.find("DEMAND_SECTION") != std::string::npos) {
                demands = true;
                break;
            }
        if (demands)
            for (int i = 0; i < dimension; ++i) {
                int id, demand;
                file >> id >> demand;
                config->customers_[id - 1].demand = demand;
            }
        const std::size_t k = config->instance_name_.find("-k");
        config->num_vehicles_ = k != std::string::npos ? std::stoi(config->instance_name_.substr(k + 2))
                                                       : std::max(3, config->num_customers_ / 6);
        config->build_distance_matrix();
        config->component_lengths = {config->num_customers_};
        config->num_components = 1;
        config->objective_functions.resize(2);
        MultiObjectiveVRPConfig<Float>* self = config.get();
        config->objective_functions[0] = [self](const Individual<Float>& ind) {
            return self->calculate_total_distance(ind);
        };
        config->objective_functions[1] = [self](const Individual<Float>& ind) {
            return self->calculate_max_route_length(ind);
        };
        return config;
    }

    // configs/multi_objective_vrp_config.hpp:188-206
    static void configure_for_size(MultiObjectiveVRPConfig<Float>* config, int num_customers)
    {
        if (num_customers <= 50) {
            config->population_size = 1000;
            config->max_generations = 1500;
        } else if (num_customers <= 100) {
            config->population_size = 1500;
            config->max_generations = 2000;
        } else {
            config->population_size = 2000;
            config->max_generations = 2500;
        }
        config->elite_size = config->population_size / 5;
        config->mutant_size = config->population_size / 6;
        config->elite_prob = 0.10f;
        config->update_cuda_grid_size();
    }

    // configs/multi_objective_vrp_config.hpp:208-262: routes in visiting order
    std::vector<Route> decode_routes(const Float* chromosome)
    {
        std::vector<int32_t> rid;
        this->host_evaluate(individual_of(chromosome), nullptr, &rid);
        const int n = num_customers_;
        float hash = 0.0f;
        for (int i = 0; i < n; ++i) hash = hash + (float)chromosome[i] * (float)(i + 1);
        std::vector<float> val(n);
        for (int i = 0; i < n; ++i) {
            const float bias = (float)i * 0.00001f + hash * 0.00001f;
            val[i] = (float)chromosome[i] + bias;
        }
        std::vector<int> order(n);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return val[a] < val[b]; });
        int routes = 0;
        for (int r : rid) routes = std::max(routes, (int)r);
        std::vector<Route> out(routes, Route{{}, Float(0), 0});
        std::vector<int> last(routes, 0);
        for (int i : order) {
            Route& r = out[rid[i] - 1];
            const int c = i + 1;
            r.distance += distances_[last[rid[i] - 1]][c];
            r.load += customers_[c].demand;
            r.customers.push_back(c);
            last[rid[i] - 1] = c;
        }
        for (int r = 0; r < routes; ++r) out[r].distance += distances_[last[r]][0];
        return out;
    }

    Float calculate_total_distance(const Individual<Float>& individual) { return objectives(individual)[0]; }
    Float calculate_max_route_length(const Individual<Float>& individual) { return objectives(individual)[1]; }

    // configs/multi_objective_vrp_config.hpp:276-300
    void decode_solution(const Float* chromosome, Float* objectives_out)
    {
        const std::vector<float> f = objectives(individual_of(chromosome));
        objectives_out[0] = f[0];
        objectives_out[1] = f[1];
    }

    // configs/multi_objective_vrp_config.hpp:302-356
    void export_solution(const Individual<Float>& individual, const std::string& filename) override
    {
        std::ofstream out(filename);
        if (!out.is_open()) throw std::runtime_error("Cannot open output file: " + filename);
        const auto routes = decode_routes(individual.get_chromosome().data());
        out << "Multi-Objective VRP Solution for: " << instance_name_ << "\n";
        out << "Objective 1 (Total Distance): " << std::fixed << std::setprecision(2) << individual.objectives[0]
            << "\n";
        out << "Objective 2 (Max Route Length): " << std::fixed << std::setprecision(2) << individual.objectives[1]
            << "\n";
        out << "Number of routes: " << routes.size() << "\n\n";
        Float total = 0.0f, longest = 0.0f;
        for (std::size_t r = 0; r < routes.size(); ++r) {
            const auto& route = routes[r];
            out << "Route " << (r + 1) << ": 0";
            for (int c : route.customers) out << " -> " << c;
            out << " -> 0\n  Distance: " << std::fixed << std::setprecision(2) << route.distance << "\n  Load: "
                << route.load << " / " << vehicle_capacity_ << "\n  Utilization: " << std::fixed
                << std::setprecision(1) << (100.0 * route.load / vehicle_capacity_) << "%\n\n";
            total += route.distance;
            longest = std::max(longest, route.distance);
        }
        out << "=== Summary ===\nTotal distance: " << std::fixed << std::setprecision(2) << total
            << "\nLongest route: " << std::fixed << std::setprecision(2) << longest
            << "\nAverage route: " << std::fixed << std::setprecision(2) << (total / routes.size())
            << "\nBalance ratio (max/avg): " << std::fixed << std::setprecision(2)
            << (longest / (total / routes.size())) << "\n";
    }

    // configs/multi_objective_vrp_config.hpp:177-192
    void print_instance_info() const
    {
        std::cout << "\n=== VRP Instance Information ===\nInstance: " << instance_name_
                  << "\nCustomers: " << num_customers_ << "\nVehicles: " << num_vehicles_
                  << "\nVehicle capacity: " << vehicle_capacity_ << "\nChromosome size: " << num_customers_
                  << "\nOptimization: Multi-Objective (Distance vs. Max Route)\n"
                  << "  Objective 1: Total distance (minimize)\n"
                  << "  Objective 2: Longest single route (minimize - minimax)\n\nTrade-off explanation:\n"
                  << "  - Minimize total: May create very long routes for some vehicles\n"
                  << "  - Minimize longest: Ensures no vehicle is overworked (fairness)\n"
                  << "================================" << std::endl;
    }

    // configs/multi_objective_vrp_config.hpp:369-386: one solution file per
    // front member plus the route data file, in `directory`
    void export_all_pareto_solutions(const std::vector<Individual<Float>>& pareto_front, const std::string& directory)
    {
        std::filesystem::create_directories(directory);
        std::cout << "\nExporting " << pareto_front.size() << " Pareto solutions..." << std::endl;
        for (std::size_t i = 0; i < pareto_front.size(); ++i) {
            std::stringstream name;
            name << directory << "/solution_" << std::setw(3) << std::setfill('0') << i << ".txt";
            export_solution(pareto_front[i], name.str());
        }
        export_pareto_routes_data(pareto_front, directory + "/routes_data.txt");
        std::cout << "All solutions exported to: " << directory << "/" << std::endl;
    }

    // configs/multi_objective_vrp_config.hpp:389-446
    void export_pareto_routes_data(const std::vector<Individual<Float>>& pareto_front, const std::string& filename)
    {
        std::ofstream out(filename);
        if (!out.is_open()) throw std::runtime_error("Cannot open output file: " + filename);
        out << "# VRP Routes Data for Visualization\n# Instance: " << instance_name_ << "\n# Customers: "
            << num_customers_ << "\n# Vehicles: " << num_vehicles_ << "\n#\n# Format:\n"
            << "# SOLUTION <idx> <obj1> <obj2>\n# ROUTE <route_idx> <distance> <load>\n"
            << "# CUSTOMER <customer_id> <x> <y> <demand>\n#\n\n";
        const Customer& depot = customers_[0];
        for (std::size_t si = 0; si < pareto_front.size(); ++si) {
            const auto& ind = pareto_front[si];
            const auto routes = decode_routes(ind.get_chromosome().data());
            out << "SOLUTION " << si << " " << std::fixed << std::setprecision(2) << ind.objectives[0] << " "
                << ind.objectives[1] << "\n";
            for (std::size_t r = 0; r < routes.size(); ++r) {
                out << "ROUTE " << r << " " << std::fixed << std::setprecision(2) << routes[r].distance << " "
                    << routes[r].load << "\n";
                out << "CUSTOMER 0 " << depot.x << " " << depot.y << " " << depot.demand << "\n";
                for (int c : routes[r].customers)
                    out << "CUSTOMER " << c << " " << customers_[c].x << " " << customers_[c].y << " "
                        << customers_[c].demand << "\n";
                out << "CUSTOMER 0 " << depot.x << " " << depot.y << " " << depot.demand << "\nEND_ROUTE\n";
            }
            out << "END_SOLUTION\n\n";
        }
    }

    int native_kind() const override { return BRKGA_MOVRP; }

protected:
    brkga_problem_t create_native_problem(brkga_engine_t engine) override
    {
        const int dim = num_customers_ + 1;
        std::vector<float> xy(2 * (std::size_t)dim);
        std::vector<int32_t> dem(dim);
        for (int i = 0; i < dim; ++i) {
            xy[2 * i] = customers_[i].x;
            xy[2 * i + 1] = customers_[i].y;
            dem[i] = customers_[i].demand;
        }
        brkga_problem_t p = nullptr;
        brkga_facade::check(brkga_problem_create_movrp(engine, xy.data(), dem.data(), dim, vehicle_capacity_,
                                                       num_vehicles_, &p),
                            "brkga_problem_create_movrp");
        return p;
    }

private:
    // configs/multi_objective_vrp_config.hpp:38-56 (FP32, 0 on the diagonal)
    void build_distance_matrix()
    {
        const int dim = num_customers_ + 1;
        distances_.assign(dim, std::vector<Float>(dim));
        for (int i = 0; i < dim; ++i)
            for (int j = 0; j < dim; ++j) {
                if (i == j) {
                    distances_[i][j] = 0.0;
                    continue;
                }
                const Float dx = customers_[i].x - customers_[j].x;
                const Float dy = customers_[i].y - customers_[j].y;
                distances_[i][j] = std::sqrt(dx * dx + dy * dy);
            }
    }

    Individual<Float> individual_of(const Float* chromosome) const
    {
        Individual<Float> ind(num_customers_);
        for (int i = 0; i < num_customers_; ++i) ind.chromosomes[0][i] = chromosome[i];
        return ind;
    }

    std::vector<float> objectives(const Individual<Float>& individual)
    {
        std::vector<float> f;
        this->host_evaluate(individual, &f, nullptr);
        return f;
    }

    std::vector<Customer> customers_;
    int num_customers_ = 0;
    int num_vehicles_ = 0;
    int vehicle_capacity_ = 0;
    std::string instance_name_;
    std::vector<std::vector<Float>> distances_;
};
