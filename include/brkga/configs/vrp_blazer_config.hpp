// brkga/configs/vrp_blazer_config.hpp - VrpBlazerConfig<Float> of the C++
// facade.
//
// Fields, loaders, size presets, decode and JSON export of the reference's
// VrpBlazerConfig (configs/vrp_blazer_config.hpp:17-381): stops in key order
// fill days of at most max_stops_per_day stops whose drive time (legs plus
// the return to depot 0) stays within max_drive_hours_per_day; fitness =
// total drive minutes + 1000 per entry of the unvisited list (which, as in
// the reference, names every skipped stop twice), minimised.
//
// The day assignment comes from the engine (brkga_problem_create_vrp_days);
// decode_solution lays it out on the host with the reference's FP32
// running sums for the exported listing.  The drive-limit fields are read
// when the engine problem is first created on a device.
#pragma once

#include <algorithm>
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
class VrpBlazerConfig : public BRKGAConfig<Float> {
public:
    int max_stops_per_day;
    double max_drive_hours_per_day;
    double stop_time_hours;
    int depot_index;
    double max_drive_minutes_per_day;
    double stop_time_minutes;

    struct DayRoute {
        int day_number;
        std::vector<int> stop_indices;  // locations (0 = depot)
        std::vector<int> stop_orders;   // 1-based position in the day
        std::vector<Float> drive_times_from_prev;
        std::vector<Float> cumulative_drive_times;
        Float total_drive_minutes;
        Float total_stop_minutes;
    };

    struct VrpSolutionData {
        std::vector<DayRoute> days;
        Float total_drive_minutes;
        Float total_stop_minutes;
        int total_stops;
        std::vector<int> unvisited_stops;
    };

    VrpBlazerConfig()
        : BRKGAConfig<Float>(std::vector<int>{1}, 1), max_stops_per_day(12), max_drive_hours_per_day(8.0),
          stop_time_hours(0.5), depot_index(0)
    {
        update_derived_params();
    }

    void update_derived_params()
    {
        max_drive_minutes_per_day = max_drive_hours_per_day * 60.0;
        stop_time_minutes = stop_time_hours * 60.0;
    }

    // configs/vrp_blazer_config.hpp:73-88
    static std::unique_ptr<VrpBlazerConfig<Float>> load_from_matrix_file(const std::string& matrix_file,
                                                                         const std::string& config_file)
    {
        auto config = std::make_unique<VrpBlazerConfig<Float>>();
        config->load_config_params(config_file);
        config->load_time_matrix(matrix_file);
        config->setup_brkga_params();
        return config;
    }

    // configs/vrp_blazer_config.hpp:90-127: "Key = value" lines; a missing
    // file keeps the defaults
    void load_config_params(const std::string& filename)
    {
        std::ifstream file(filename);
        if (!file.is_open()) {
            std::cerr << "Warning: Cannot open config file, using defaults: " << filename << std::endl;
            return;
        }
        std::string line;
        while (std::getline(file, line)) {
            std::istringstream iss(line);
            std::string key;
            char equals;
            if (!(iss >> key >> equals)) continue;
            if (key == "MaxStopsPerDay") iss >> max_stops_per_day;
            else if (key == "MaxDriveHoursPerDay") iss >> max_drive_hours_per_day;
            else if (key == "StopTimeHours") iss >> stop_time_hours;
            else if (key == "DepotIndex") iss >> depot_index;
            else if (key == "PopulationSize") iss >> this->population_size;
            else if (key == "MaxGenerations") iss >> this->max_generations;
        }
        update_derived_params();
    }

    // configs/vrp_blazer_config.hpp:129-149: size, then the full matrix
    void load_time_matrix(const std::string& filename)
    {
        std::ifstream file(filename);
        if (!file.is_open()) throw std::runtime_error("Cannot open time matrix file: " + filename);
        file >> num_locations_;
        num_stops_ = num_locations_ - 1;
        time_matrix_.assign(num_locations_, std::vector<Float>(num_locations_));
        for (int i = 0; i < num_locations_; ++i)
            for (int j = 0; j < num_locations_; ++j) file >> time_matrix_[i][j];
    }

    void set_time_matrix(const std::vector<std::vector<Float>>& matrix)
    {
        time_matrix_ = matrix;
        num_locations_ = (int)matrix.size();
        num_stops_ = num_locations_ - 1;
        setup_brkga_params();
    }

    // configs/vrp_blazer_config.hpp:158-189
    void setup_brkga_params()
    {
        this->component_lengths = {num_stops_};
        this->num_components = 1;
        if (num_stops_ <= 50) {
            this->population_size = 500;
            this->max_generations = 1000;
        } else if (num_stops_ <= 100) {
            this->population_size = 800;
            this->max_generations = 1500;
        } else if (num_stops_ <= 200) {
            this->population_size = 1000;
            this->max_generations = 2000;
        } else {
            this->population_size = 1500;
            this->max_generations = 2500;
        }
        this->elite_size = this->population_size / 5;
        this->mutant_size = this->population_size / 10;
        this->elite_prob = 0.7;
        this->update_cuda_grid_size();
        this->fitness_function = [this](const Individual<Float>& ind) { return calculate_fitness(ind); };
        this->comparator = [](Float a, Float b) { return a < b; };
    }

    // configs/vrp_blazer_config.hpp:192-282, laid out from the engine's day
    // of each stop
    VrpSolutionData decode_solution(const Float* chromosome)
    {
        std::vector<int32_t> day;
        this->host_evaluate(individual_of(chromosome), nullptr, &day);
        std::vector<int> order(num_stops_);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return chromosome[a] < chromosome[b]; });
        VrpSolutionData sol{};
        sol.total_drive_minutes = 0;
        sol.total_stop_minutes = 0;
        sol.total_stops = 0;
        for (int i : order) {
            const int stop = i + 1;
            if (day[i] == 0) {
                sol.unvisited_stops.push_back(stop);
                continue;
            }
            if (sol.days.empty() || sol.days.back().day_number != day[i]) {
                if (!sol.days.empty()) close_day(sol);
                DayRoute d{};
                d.day_number = day[i];
                d.total_drive_minutes = 0;
                d.total_stop_minutes = 0;
                sol.days.push_back(d);
            }
            DayRoute& d = sol.days.back();
            const int prev = d.stop_indices.empty() ? depot_index : d.stop_indices.back();
            const Float leg = time_matrix_[prev][stop];
            const Float cum = (d.cumulative_drive_times.empty() ? Float(0) : d.cumulative_drive_times.back()) + leg;
            d.stop_indices.push_back(stop);
            d.stop_orders.push_back((int)d.stop_indices.size());
            d.drive_times_from_prev.push_back(leg);
            d.cumulative_drive_times.push_back(cum);
        }
        if (!sol.days.empty()) close_day(sol);
        // the reference's final sweep lists every stop no day visited again
        for (int i = 0; i < num_stops_; ++i)
            if (day[i] == 0) sol.unvisited_stops.push_back(i + 1);
        return sol;
    }

    // configs/vrp_blazer_config.hpp:285-296
    Float calculate_fitness(const Individual<Float>& individual)
    {
        std::vector<float> f;
        this->host_evaluate(individual, &f, nullptr);
        return static_cast<Float>(f[0]);
    }

    // configs/vrp_blazer_config.hpp:298-366
    void export_solution_json(const Individual<Float>& individual, const std::string& filename,
                              double solve_time_seconds)
    {
        std::ofstream out(filename);
        if (!out.is_open()) throw std::runtime_error("Cannot open output file: " + filename);
        const auto sol = decode_solution(individual.get_chromosome().data());
        out << "{\n  \"SolverName\": \"BRKGA-CUDA\",\n  \"SolveTimeSeconds\": " << std::fixed << std::setprecision(3)
            << solve_time_seconds << ",\n  \"TotalDriveMinutes\": " << std::setprecision(2) << sol.total_drive_minutes
            << ",\n  \"TotalStopMinutes\": " << sol.total_stop_minutes << ",\n  \"TotalStops\": " << sol.total_stops
            << ",\n  \"UnvisitedStops\": [";
        for (std::size_t i = 0; i < sol.unvisited_stops.size(); ++i)
            out << sol.unvisited_stops[i] << (i + 1 < sol.unvisited_stops.size() ? ", " : "");
        out << "],\n  \"Days\": [\n";
        for (std::size_t d = 0; d < sol.days.size(); ++d) {
            const auto& day = sol.days[d];
            out << "    {\n      \"DayNumber\": " << day.day_number << ",\n      \"DriveTimeMinutes\": "
                << std::setprecision(2) << day.total_drive_minutes << ",\n      \"StopTimeMinutes\": "
                << day.total_stop_minutes << ",\n      \"Stops\": [\n";
            for (std::size_t s = 0; s < day.stop_indices.size(); ++s) {
                out << "        {\n          \"Index\": " << day.stop_indices[s] << ",\n          \"Order\": "
                    << day.stop_orders[s] << ",\n          \"DriveTimeFromPrevious\": " << std::setprecision(2)
                    << day.drive_times_from_prev[s] << ",\n          \"CumulativeDriveTime\": "
                    << day.cumulative_drive_times[s] << "\n        }" << (s + 1 < day.stop_indices.size() ? "," : "")
                    << "\n";
            }
            out << "      ]\n    }" << (d + 1 < sol.days.size() ? "," : "") << "\n";
        }
        out << "  ]\n}\n";
    }

    void print_instance_info() const
    {
        std::cout << "=== VRP Blazer Instance ===\nLocations: " << num_locations_ << " (1 depot + " << num_stops_
                  << " stops)\nMax stops/day: " << max_stops_per_day << "\nMax drive hours/day: "
                  << max_drive_hours_per_day << "\nStop time (hours): " << stop_time_hours << std::endl;
    }

    int get_num_stops() const { return num_stops_; }
    int get_num_locations() const { return num_locations_; }
    const std::vector<std::vector<Float>>& get_time_matrix() const { return time_matrix_; }

    int native_kind() const override { return BRKGA_VRP_DAYS; }

protected:
    brkga_problem_t create_native_problem(brkga_engine_t engine) override
    {
        std::vector<float> flat;
        flat.reserve((std::size_t)num_locations_ * num_locations_);
        for (const auto& row : time_matrix_) flat.insert(flat.end(), row.begin(), row.end());
        brkga_problem_t p = nullptr;
        brkga_facade::check(brkga_problem_create_vrp_days(engine, flat.data(), num_locations_, max_stops_per_day,
                                                          max_drive_minutes_per_day, &p),
                            "brkga_problem_create_vrp_days");
        return p;
    }

private:
    Individual<Float> individual_of(const Float* chromosome) const
    {
        Individual<Float> ind(num_stops_);
        for (int i = 0; i < num_stops_; ++i) ind.chromosomes[0][i] = chromosome[i];
        return ind;
    }

    void close_day(VrpSolutionData& sol) const
    {
        DayRoute& d = sol.days.back();
        const int last = d.stop_indices.back();
        d.total_drive_minutes = d.cumulative_drive_times.back() + time_matrix_[last][depot_index];
        d.total_stop_minutes = (Float)(d.stop_indices.size() * stop_time_minutes);
        sol.total_drive_minutes += d.total_drive_minutes;
        sol.total_stop_minutes += d.total_stop_minutes;
        sol.total_stops += (int)d.stop_indices.size();
    }

    std::vector<std::vector<Float>> time_matrix_;
    int num_locations_ = 1;
    int num_stops_ = 0;
};
