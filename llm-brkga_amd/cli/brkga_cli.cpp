// brkga_cli.cpp - prebuilt command-line solver on the C++ facade.
//
// Replaces the reference's driver (main.cu:46-305), which writes a main
// file per config and compiles it with nvcc at run time: the same flags,
// the same configs, the same stdout lines (the llm_solver layer parses
// "Best fitness:", "Generations:", "Pareto front size:") and the same
// exports (solutions/<config>_solution.txt, <config>_pareto_front.txt).
// QUICK_TEST=<n> caps the generation count, as the generated mains of
// llm_solver/core/execution_manager.py:341-349 do.
#include <cstdio>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <memory>
#include <string>

#include "brkga/configs/knapsack_config.hpp"
#include "brkga/configs/multi_objective_vrp_config.hpp"
#include "brkga/configs/multi_tsp_config.hpp"
#include "brkga/configs/tsp_config.hpp"
#include "brkga/configs/tspj_config.hpp"
#include "brkga/core/solver.hpp"

namespace {

void print_usage(const char* prog)
{
    std::cout << "Usage: " << prog << " --data <file> --config <config_name> [options]\n"
              << "\nRequired:\n"
              << "  --data <file>        Input data file\n"
              << "  --config <name>      Config name\n"
              << "\nOptional:\n"
              << "  --verbose            Enable verbose output\n"
              << "  --benchmark          Time the generation kernels before the run\n"
              << "  --pop <size>         Override population size (elite 1/5, mutants 1/10)\n"
              << "  --gen <count>        Override generation count\n"
              << "  --seed <n>           Philox seed (default 0x5EED)\n"
              << "  --islands <k> <interval> <migrants>  Island model on one GPU\n"
              << "  --device <id>        GPU to run on (default 0)\n"
              << "  --no-graph           Launch generations one by one instead of replaying a graph\n"
              << "\nAvailable configs:\n"
              << "  Single-Objective (BRKGA):\n"
              << "    tsp_config        - Traveling Salesman Problem (TSPLIB NODE_COORD_SECTION)\n"
              << "    tspj_config       - TSP with Job assignment (*_TT.csv / *_JT.csv pair)\n"
              << "    knapsack_config   - 0/1 Knapsack Problem (text or CSV)\n"
              << "\n  Multi-Objective (NSGA-II):\n"
              << "    multi_tsp_config  - Multi-objective TSP (distance + time)\n"
              << "    multi_objective_vrp_config - Multi-objective VRP (distance + max route)\n"
              << "\nExamples:\n"
              << "  " << prog << " --data data/berlin52.tsp --config tsp_config\n"
              << "  " << prog << " --data data/berlin52.tsp --config multi_tsp_config --pop 512\n"
              << "  " << prog << " --data data/A-n32-k5.vrp --config multi_objective_vrp_config\n"
              << "  " << prog << " --data data/knapsack.txt --config knapsack_config --gen 500\n";
}

struct Options {
    std::string data, config;
    bool verbose = false, benchmark = false, graph = true;
    int pop = -1, gen = -1, device = 0;
    int islands = 1, interval = 0, migrants = 0;
    uint64_t seed = 0x5EED;
};

std::unique_ptr<BRKGAConfig<float>> load(const Options& o)
{
    if (o.config == "tsp_config") {
        auto c = TSPConfig<float>::load_from_file(o.data);
        c->print_instance_info();
        TSPConfig<float>::configure_for_size(c.get(), c->get_num_cities());
        return c;
    }
    if (o.config == "tspj_config") {
        auto c = TSPJConfig<float>::load_from_file(o.data);
        c->print_instance_info();
        TSPJConfig<float>::configure_for_size(c.get(), c->get_num_cities());
        return c;
    }
    if (o.config == "knapsack_config") {
        auto c = KnapsackConfig<float>::load_from_file(o.data);
        c->print_instance_info();
        KnapsackConfig<float>::configure_for_size(c.get(), (int)c->get_weights().size());
        return c;
    }
    if (o.config == "multi_tsp_config") {
        auto c = MultiTSPConfig<float>::load_from_file(o.data);
        c->print_instance_info();
        MultiTSPConfig<float>::configure_for_size(c.get(), c->get_num_cities());
        return c;
    }
    if (o.config == "multi_objective_vrp_config") {
        auto c = MultiObjectiveVRPConfig<float>::load_from_file(o.data);
        c->print_instance_info();
        MultiObjectiveVRPConfig<float>::configure_for_size(c.get(), c->get_num_customers());
        return c;
    }
    throw std::runtime_error("Unknown config type: " + o.config);
}

int run(const Options& o)
{
    std::cout << "Loading data file: " << o.data << std::endl;
    auto config = load(o);
    if (o.pop > 0) {
        config->population_size = o.pop;
        config->elite_size = o.pop / 5;
        config->mutant_size = o.pop / 10;
        config->update_cuda_grid_size();
        std::cout << "Population size overridden to: " << o.pop << std::endl;
    }
    if (o.gen > 0) {
        config->max_generations = o.gen;
        std::cout << "Max generations overridden to: " << o.gen << std::endl;
    }
    if (const char* q = std::getenv("QUICK_TEST")) {
        config->max_generations = std::atoi(q);
        std::cout << "Quick test mode: limiting to " << config->max_generations << " generations" << std::endl;
    }
    config->set_device(o.device);
    const bool multi = config->is_multi_objective();
    std::cout << "Configuration loaded successfully" << std::endl;
    config->print_config();

    Solver<float> solver(std::move(config), true, 50);
    solver.set_seed(o.seed);
    solver.set_device(o.device);
    solver.set_use_graph(o.graph);
    if (o.islands > 1) solver.set_islands(o.islands, o.interval, o.migrants);
    if (o.benchmark) {
        std::cout << "\n=== Running Benchmark Mode ===" << std::endl;
        solver.initialize();
        double step_us = 0.0, close_us = 0.0;
        brkga_facade::check(brkga_population_profile(solver.get_population().handle(), 20, &step_us, &close_us,
                                                     nullptr),
                            "brkga_population_profile");
        std::cout << "Generation kernel: " << std::fixed << std::setprecision(2) << step_us
                  << " us, closing kernels: " << close_us << " us" << std::endl;
        std::cout << "================================" << std::endl;
    }
    solver.run();

    if (multi) {
        const std::string front_file = o.config + "_pareto_front.txt";
        solver.export_pareto_front(front_file);
        std::cout << "\n✓ Pareto front exported to: " << front_file << std::endl;
        const auto pareto = solver.get_pareto_front();
        if (o.config == "multi_objective_vrp_config" && !pareto.empty()) {
            auto* vrp = dynamic_cast<MultiObjectiveVRPConfig<float>*>(solver.get_config());
            if (vrp) {
                vrp->export_all_pareto_solutions(pareto, "pareto_solutions");
                std::cout << "\n✓ All " << pareto.size() << " solutions exported to: pareto_solutions/" << std::endl;
            }
        }
        std::cout << "\n=== Pareto Front Analysis ===\nSolutions in Pareto front: " << pareto.size() << std::endl;
        if (!pareto.empty()) {
            std::cout << "\nFirst 5 solutions:\n  Obj1      Obj2\n  ----------------------" << std::endl;
            for (std::size_t i = 0; i < pareto.size() && i < 5; ++i)
                std::cout << "  " << std::fixed << std::setprecision(4) << pareto[i].objectives[0] << "  "
                          << pareto[i].objectives[1] << std::endl;
        }
    } else {
        const auto& best = solver.get_best_individual();
        std::cout << "\n=== Best Solution ===\nFitness: " << std::fixed << std::setprecision(6) << best.fitness
                  << std::endl;
        FileUtils::ensure_directory("solutions");
        const std::string solution_file = "solutions/" + o.config + "_solution.txt";
        solver.get_config()->export_solution(best, solution_file);
        std::cout << "\n✓ Solution exported to: " << solution_file << std::endl;
    }
    std::cout << "\n✓ Optimization completed successfully!" << std::endl;
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    Options o;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&](const char* what) -> std::string {
            if (i + 1 >= argc) throw std::invalid_argument(std::string("missing value for ") + what);
            return argv[++i];
        };
        try {
            if (a == "--data") o.data = next("--data");
            else if (a == "--config") o.config = next("--config");
            else if (a == "--verbose") o.verbose = true;
            else if (a == "--benchmark") o.benchmark = true;
            else if (a == "--pop") o.pop = std::stoi(next("--pop"));
            else if (a == "--gen") o.gen = std::stoi(next("--gen"));
            else if (a == "--seed") o.seed = std::stoull(next("--seed"), nullptr, 0);
            else if (a == "--device") o.device = std::stoi(next("--device"));
            else if (a == "--no-graph") o.graph = false;
            else if (a == "--islands") {
                o.islands = std::stoi(next("--islands"));
                o.interval = std::stoi(next("--islands"));
                o.migrants = std::stoi(next("--islands"));
            } else if (a == "--help" || a == "-h") {
                print_usage(argv[0]);
                return 0;
            }
        } catch (const std::exception& e) {
            std::cerr << "Error: " << e.what() << std::endl;
            return 1;
        }
    }
    if (o.data.empty() || o.config.empty()) {
        std::cerr << "Error: Missing required arguments" << std::endl;
        print_usage(argv[0]);
        return 1;
    }
    if (!FileUtils::file_exists(o.data)) {
        std::cerr << "Error: Data file not found: " << o.data << std::endl;
        return 1;
    }
    try {
        return run(o);
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << std::endl;
        return 1;
    }
}
