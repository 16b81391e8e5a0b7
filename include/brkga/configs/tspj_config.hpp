// brkga/configs/tspj_config.hpp - TSPJConfig<T> of the C++ facade.
//
// Constructor, callbacks, size presets and accessors of the reference's
// TSPJConfig (configs/tspj_config.hpp:27-870): travel[dim][dim] and
// jobs[dim][dim] with row / column 0 the depot, chromosome = n city keys |
// n job keys, makespan minimised (configs/tspj_config.hpp:555-642).  The
// local search (configs/tspj_config.hpp:77-135: random 2-opt moves on the
// elites, kept when the makespan improves) runs on the device inside the
// generation loop; the solver's options set its schedule.
#pragma once

#include <fstream>
#include <iomanip>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../core/config.hpp"
#include "../utils/file_utils.hpp"

template <typename T>
class TSPJConfig : public BRKGAConfig<T> {
public:
    TSPJConfig(const std::vector<std::vector<T>>& travel, const std::vector<std::vector<T>>& jobs,
               const std::string& name = "TSPJ")
        : BRKGAConfig<T>(std::vector<int>{static_cast<int>(travel.size() - 1), static_cast<int>(travel.size() - 1)}),
          travel_(travel), jobs_(jobs), num_cities_(static_cast<int>(travel.size() - 1)),
          num_jobs_(static_cast<int>(travel.size() - 1)), instance_name_(name)
    {
        this->fitness_function = [this](const Individual<T>& ind) { return calculate_tspj_fitness(ind); };
        this->decoder = [this](const Individual<T>& ind) { return decode_to_solution(ind); };
        this->comparator = [](T a, T b) { return a < b; };
        this->threads_per_block = 256;
        this->update_cuda_grid_size();
    }

    // configs/tspj_config.hpp:555-585
    T calculate_tspj_fitness(const Individual<T>& individual)
    {
        std::vector<float> f;
        this->host_evaluate(individual, &f, nullptr);
        return static_cast<T>(f[0]);
    }

    // configs/tspj_config.hpp:587-604: [1-based cities], [1-based jobs]
    std::vector<std::vector<T>> decode_to_solution(const Individual<T>& individual)
    {
        const auto s = decode_solution_internal(individual);
        std::vector<std::vector<T>> out(2);
        for (int c : s.first) out[0].push_back(static_cast<T>(c));
        for (int j : s.second) out[1].push_back(static_cast<T>(j));
        return out;
    }

    std::pair<std::vector<int>, std::vector<int>> decode_solution_internal(const Individual<T>& individual)
    {
        std::vector<int32_t> d;
        this->host_evaluate(individual, nullptr, &d);
        return {std::vector<int>(d.begin(), d.begin() + num_cities_),
                std::vector<int>(d.begin() + num_cities_, d.end())};
    }

    // configs/tspj_config.hpp:805-830
    static void configure_for_size(TSPJConfig<T>* config, int num_cities)
    {
        if (num_cities <= 10) {
            config->population_size = 300;
            config->elite_size = 60;
            config->mutant_size = 30;
            config->max_generations = 500;
        } else if (num_cities <= 20) {
            config->population_size = 600;
            config->elite_size = 120;
            config->mutant_size = 60;
            config->max_generations = 1000;
        } else if (num_cities <= 50) {
            config->population_size = 1000;
            config->elite_size = 200;
            config->mutant_size = 100;
            config->max_generations = 2000;
        } else {
            config->population_size = 2000;
            config->elite_size = 350;
            config->mutant_size = 200;
            config->max_generations = 30000;
        }
        config->elite_prob = 0.75;
        config->update_cuda_grid_size();
    }

    // configs/tspj_config.hpp:727-769: a pair of CSV matrices named
    // <name>_TT.csv / <name>_JT.csv or <name>_cost_table_by_coordinates.csv /
    // <name>_tasktime_table.csv; either file of the pair may be given.
    static std::unique_ptr<TSPJConfig<T>> load_from_file(const std::string& filename)
    {
        const std::string base = FileUtils::get_basename(filename);
        const std::string dir = FileUtils::get_directory(filename);
        std::string tt, jt, name;
        std::size_t at;
        if ((at = base.find("_TT")) != std::string::npos) {
            name = base.substr(0, at);
            tt = filename;
            jt = dir + "/" + name + "_JT.csv";
        } else if ((at = base.find("_JT")) != std::string::npos) {
            name = base.substr(0, at);
            jt = filename;
            tt = dir + "/" + name + "_TT.csv";
        } else if ((at = base.find("_cost_table_by_coordinates")) != std::string::npos) {
            name = base.substr(0, at);
            tt = filename;
            jt = dir + "/" + name + "_tasktime_table.csv";
        } else if ((at = base.find("_tasktime_table")) != std::string::npos) {
            name = base.substr(0, at);
            jt = filename;
            tt = dir + "/" + name + "_cost_table_by_coordinates.csv";
        } else {
            throw std::runtime_error("TSPJ CSV files must be named with '_TT'/'_JT' or "
                                     "'_cost_table_by_coordinates'/'_tasktime_table' suffix");
        }
        if (!FileUtils::file_exists(tt)) throw std::runtime_error("Travel times file not found: " + tt);
        if (!FileUtils::file_exists(jt)) throw std::runtime_error("Job times file not found: " + jt);
        return std::make_unique<TSPJConfig<T>>(load_csv_matrix(tt), load_csv_matrix(jt), name);
    }

    // configs/tspj_config.hpp:771-802: comma-separated rows ('#' comments);
    // empty cells are skipped, unparsable ones read as 0.
    static std::vector<std::vector<T>> load_csv_matrix(const std::string& filename)
    {
        std::vector<std::vector<T>> m;
        for (const std::string& line : FileUtils::read_lines(filename)) {
            if (line.empty() || line[0] == '#') continue;
            std::vector<T> row;
            for (const std::string& cell : FileUtils::parse_csv_line(line)) {
                if (cell.empty()) continue;
                try {
                    row.push_back(static_cast<T>(std::stod(cell)));
                } catch (...) {
                    row.push_back(T(0));
                }
            }
            if (!row.empty()) m.push_back(row);
        }
        return m;
    }

    // configs/tspj_config.hpp:695-725
    void export_solution(const Individual<T>& individual, const std::string& filename) override
    {
        const auto s = decode_solution_internal(individual);
        FileUtils::ensure_directory(FileUtils::get_directory(filename));
        std::ofstream f(filename);
        if (!f) throw std::runtime_error("Cannot create solution file: " + filename);
        f << "TSPJ Solution for " << instance_name_ << std::endl;
        f << "Makespan: " << std::fixed << std::setprecision(6) << individual.fitness << std::endl;
        f << "City sequence: ";
        for (std::size_t i = 0; i < s.first.size(); ++i) f << s.first[i] << (i + 1 < s.first.size() ? " " : "");
        f << std::endl << "Job sequence: ";
        for (std::size_t i = 0; i < s.second.size(); ++i) f << s.second[i] << (i + 1 < s.second.size() ? " " : "");
        f << std::endl;
    }

    int get_num_cities() const { return num_cities_; }
    int get_num_jobs() const { return num_jobs_; }
    const std::vector<std::vector<T>>& get_travel_times() const { return travel_; }
    const std::vector<std::vector<T>>& get_job_durations() const { return jobs_; }
    const std::string& get_instance_name() const { return instance_name_; }

    void print_instance_info() const
    {
        std::cout << "=== TSPJ Instance Info ===\nName: " << instance_name_ << "\nCities: " << num_cities_
                  << "\nJobs: " << num_jobs_ << "\n==========================" << std::endl;
    }

    int native_kind() const override { return BRKGA_TSPJ; }
    bool has_local_search() const override { return true; }

protected:
    brkga_problem_t create_native_problem(brkga_engine_t engine) override
    {
        const int dim = num_cities_ + 1;
        std::vector<float> t, j;
        for (const auto& row : travel_)
            for (T v : row) t.push_back(static_cast<float>(v));
        for (const auto& row : jobs_)
            for (T v : row) j.push_back(static_cast<float>(v));
        if ((int)t.size() != dim * dim || (int)j.size() != dim * dim)
            throw std::invalid_argument("TSPJ matrices must both be (n + 1) x (n + 1)");
        brkga_problem_t p = nullptr;
        brkga_facade::check(brkga_problem_create_tspj(engine, t.data(), j.data(), dim, &p),
                            "brkga_problem_create_tspj");
        return p;
    }

private:
    std::vector<std::vector<T>> travel_, jobs_;
    int num_cities_, num_jobs_;
    std::string instance_name_;
};
