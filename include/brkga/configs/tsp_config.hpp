// brkga/configs/tsp_config.hpp - TSPConfig<T> of the C++ facade.
//
// Constructor, decoder / fitness callbacks, size presets and accessors of
// the reference's TSPConfig (configs/tsp_config.hpp:25-300).  The device
// problem is either the N x N matrix the reference takes (engine kind
// BRKGA_TSP_MATRIX) or, through from_coordinates(), the Euclidean
// coordinates themselves (BRKGA_TSP_COORDS: edges computed on the fly with
// the reference's formula, configs/tsp_config.hpp:366-368, no matrix).
#pragma once

#include <cctype>
#include <memory>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "../core/config.hpp"
#include "../utils/file_utils.hpp"

template <typename T>
class TSPConfig : public BRKGAConfig<T> {
public:
    // configs/tsp_config.hpp:38-61: visit all N cities of the matrix
    TSPConfig(const std::vector<std::vector<T>>& distances, const std::string& name = "TSP")
        : BRKGAConfig<T>(std::vector<int>{static_cast<int>(distances.size())}), distance_matrix_(distances),
          num_cities_(static_cast<int>(distances.size())), instance_name_(name)
    {
        setup();
    }

    // Euclidean instance from (x, y) coordinates; no matrix is formed.
    static std::unique_ptr<TSPConfig<T>> from_coordinates(const std::vector<std::pair<T, T>>& xy,
                                                          const std::string& name = "TSP")
    {
        std::unique_ptr<TSPConfig<T>> c(new TSPConfig<T>(static_cast<int>(xy.size()), name));
        c->coords_ = xy;
        return c;
    }

    // configs/tsp_config.hpp:180-193, evaluated by the engine's decoder
    T calculate_tsp_fitness(const Individual<T>& individual)
    {
        std::vector<float> f;
        this->host_evaluate(individual, &f, nullptr);
        return static_cast<T>(f[0]);
    }

    // configs/tsp_config.hpp:195-204: the tour as one component
    std::vector<std::vector<T>> decode_to_solution(const Individual<T>& individual)
    {
        const std::vector<int> tour = decode_tour_internal(individual);
        std::vector<std::vector<T>> out(1);
        for (int c : tour) out[0].push_back(static_cast<T>(c));
        return out;
    }

    // configs/tsp_config.hpp:206-225: stable argsort of the keys
    std::vector<int> decode_tour_internal(const Individual<T>& individual)
    {
        std::vector<int32_t> d;
        this->host_evaluate(individual, nullptr, &d);
        return std::vector<int>(d.begin(), d.end());
    }

    // configs/tsp_config.hpp:267-287
    static void configure_for_size(TSPConfig<T>* config, int num_cities)
    {
        if (num_cities <= 20) {
            config->population_size = 200;
            config->elite_size = 40;
            config->mutant_size = 20;
            config->max_generations = 100;
        } else if (num_cities <= 50) {
            config->population_size = 1000;
            config->elite_size = 800;
            config->mutant_size = 400;
            config->max_generations = 2000;
        } else {
            config->population_size = 8000;
            config->elite_size = 1600;
            config->mutant_size = 800;
            config->max_generations = 5000;
        }
        config->elite_prob = 0.7;
        config->update_cuda_grid_size();
    }

    // configs/tsp_config.hpp:260-265
    static std::unique_ptr<TSPConfig<T>> load_from_file(const std::string& filename)
    {
        if (!FileUtils::file_exists(filename)) throw std::runtime_error("File does not exist: " + filename);
        return load_from_tsplib(filename);
    }

    // TSPLIB NODE_COORD_SECTION instance (configs/tsp_config.hpp:301-380):
    // NAME, DIMENSION, then "id x y" lines with 1-based ids until EOF.  The
    // reference turns the coordinates into an n x n matrix with the edge
    // formula the engine evaluates from coordinates directly, so the
    // instance comes back as a coordinate instance (no matrix).
    static std::unique_ptr<TSPConfig<T>> load_from_tsplib(const std::string& filename)
    {
        std::string name;
        std::vector<std::pair<T, T>> xy;
        read_tsplib_coords(filename, &name, &xy);
        if (xy.empty()) throw std::runtime_error("No coordinates found in TSPLIB file");
        if (name.empty()) name = FileUtils::get_basename(filename);
        return from_coordinates(xy, name);
    }

    // TSPLIB header keyword of a line ("NAME : x" -> "NAME"), or "".
    static std::string tsplib_key(const std::string& line, std::string* value)
    {
        const std::size_t a = line.find_first_not_of(" \t");
        if (a == std::string::npos) return std::string();
        std::size_t b = a;
        while (b < line.size() && (std::isalnum((unsigned char)line[b]) || line[b] == '_')) ++b;
        const std::string key = line.substr(a, b - a);
        if (value) {
            std::size_t c = line.find(':', b);
            c = c == std::string::npos ? b : c + 1;
            const std::size_t v0 = line.find_first_not_of(" \t", c);
            const std::size_t v1 = line.find_last_not_of(" \t");
            *value = v0 == std::string::npos ? std::string() : line.substr(v0, v1 - v0 + 1);
        }
        return key;
    }

    // Coordinates of a TSPLIB file, index = id - 1 (ids outside 1..DIMENSION
    // are ignored, missing ids stay at the origin).
    static void read_tsplib_coords(const std::string& filename, std::string* name, std::vector<std::pair<T, T>>* xy)
    {
        bool coords = false;
        for (const std::string& line : FileUtils::read_lines(filename)) {
            std::string value;
            const std::string key = tsplib_key(line, &value);
            if (key.empty()) continue;
            if (key == "EOF") break;
            if (!coords) {
                if (key == "NAME") *name = value;
                else if (key == "DIMENSION") xy->assign(std::stoi(value), std::pair<T, T>(T(0), T(0)));
                else if (key == "NODE_COORD_SECTION") coords = true;
                continue;
            }
            std::istringstream in(line);
            int id;
            T x, y;
            if (in >> id >> x >> y && id > 0 && id <= (int)xy->size()) (*xy)[id - 1] = {x, y};
        }
    }

    int get_num_cities() const { return num_cities_; }
    const std::string& get_instance_name() const { return instance_name_; }
    bool is_gpu_available() const { return true; }
    // Coordinates of a coordinate instance (empty for a matrix instance).
    const std::vector<std::pair<T, T>>& get_coordinates() const { return coords_; }

    void print_instance_info() const
    {
        std::cout << "=== TSP Instance Info ===\nName: " << instance_name_ << "\nCities: " << num_cities_
                  << "\nGPU Support: Yes\n=========================" << std::endl;
    }

    int native_kind() const override { return coords_.empty() ? BRKGA_TSP_MATRIX : BRKGA_TSP_COORDS; }

protected:
    brkga_problem_t create_native_problem(brkga_engine_t engine) override
    {
        brkga_problem_t p = nullptr;
        if (!coords_.empty()) {
            std::vector<float> xy;
            for (const auto& c : coords_) {
                xy.push_back(static_cast<float>(c.first));
                xy.push_back(static_cast<float>(c.second));
            }
            brkga_facade::check(brkga_problem_create_tsp_coords(engine, xy.data(), num_cities_, &p),
                                "brkga_problem_create_tsp_coords");
        } else {
            std::vector<float> m;
            m.reserve((size_t)num_cities_ * num_cities_);
            for (const auto& row : distance_matrix_)
                for (T v : row) m.push_back(static_cast<float>(v));
            brkga_facade::check(brkga_problem_create_tsp_matrix(engine, m.data(), num_cities_, &p),
                                "brkga_problem_create_tsp_matrix");
        }
        return p;
    }

private:
    TSPConfig(int n, const std::string& name)
        : BRKGAConfig<T>(std::vector<int>{n}), num_cities_(n), instance_name_(name)
    {
        setup();
    }

    void setup()
    {
        this->fitness_function = [this](const Individual<T>& ind) { return calculate_tsp_fitness(ind); };
        this->decoder = [this](const Individual<T>& ind) { return decode_to_solution(ind); };
        this->comparator = [](T a, T b) { return a < b; };
        this->threads_per_block = 256;
        this->update_cuda_grid_size();
    }

    std::vector<std::vector<T>> distance_matrix_;
    std::vector<std::pair<T, T>> coords_;
    int num_cities_;
    std::string instance_name_;
};
