// brkga/core/individual.hpp - host-side individual of the C++ facade.
//
// Same interface as the reference's Individual<T> (core/individual.hpp:10-226):
// a list of chromosome components, fitness / evaluated flag, objectives,
// NSGA-II rank and crowding distance.  On the engine the individual is one
// flattened row [component0 | component1 | ...] of the device key matrix;
// flatten()/unflatten() convert between the two views.
#pragma once

#include <cstddef>
#include <random>
#include <vector>

template <typename T>
class Individual {
public:
    std::vector<std::vector<T>> chromosomes;
    T fitness;
    bool evaluated;
    std::vector<T> objectives;
    int rank;
    T crowding_distance;
    std::vector<int> dominated_solutions;
    int domination_count;

    explicit Individual(int length)
        : chromosomes(1, std::vector<T>(length)), fitness(T(0)), evaluated(false), rank(-1),
          crowding_distance(T(0)), domination_count(0)
    {
    }

    explicit Individual(const std::vector<int>& component_lengths)
        : fitness(T(0)), evaluated(false), rank(-1), crowding_distance(T(0)), domination_count(0)
    {
        for (int len : component_lengths) chromosomes.emplace_back(len);
    }

    Individual(const std::vector<int>& component_lengths, int num_objectives) : Individual(component_lengths)
    {
        objectives.assign(num_objectives, T(0));
    }

    // Fills every gene from U[0,1) (individual.hpp:79-87); the engine itself
    // draws its keys from Philox on the device.
    void randomize(std::mt19937& rng)
    {
        std::uniform_real_distribution<T> u(T(0), T(1));
        for (auto& c : chromosomes)
            for (auto& g : c) g = u(rng);
        evaluated = false;
    }

    void reset_evaluation() { evaluated = false; }
    bool is_multi_objective() const { return !objectives.empty(); }

    // Minimisation dominance: no objective worse, at least one better.
    bool dominates(const Individual<T>& other) const
    {
        if (objectives.empty() || objectives.size() != other.objectives.size()) return false;
        bool strictly = false;
        for (std::size_t i = 0; i < objectives.size(); ++i) {
            if (other.objectives[i] < objectives[i]) return false;
            strictly = strictly || objectives[i] < other.objectives[i];
        }
        return strictly;
    }

    void set_fitness(T f)
    {
        fitness = f;
        evaluated = true;
    }

    void set_objectives(const std::vector<T>& objs)
    {
        objectives = objs;
        if (!objs.empty()) fitness = objs[0];
        evaluated = true;
    }

    int total_genes() const
    {
        int n = 0;
        for (const auto& c : chromosomes) n += (int)c.size();
        return n;
    }
    int get_total_chromosome_length() const { return total_genes(); }
    int num_components() const { return (int)chromosomes.size(); }
    bool is_evaluated() const { return evaluated; }
    T get_fitness() const { return fitness; }

    std::vector<T>& get_chromosome() { return chromosomes[0]; }
    const std::vector<T>& get_chromosome() const { return chromosomes[0]; }
    std::vector<T>& chromosome() { return chromosomes[0]; }
    const std::vector<T>& chromosome() const { return chromosomes[0]; }
    std::vector<T>& get_component(int i) { return chromosomes[i]; }
    const std::vector<T>& get_component(int i) const { return chromosomes[i]; }
    int size() const { return chromosomes.empty() ? 0 : (int)chromosomes[0].size(); }
    T& operator[](int i) { return chromosomes[0][i]; }
    const T& operator[](int i) const { return chromosomes[0][i]; }

    std::vector<T> flatten() const
    {
        std::vector<T> row;
        row.reserve(total_genes());
        for (const auto& c : chromosomes) row.insert(row.end(), c.begin(), c.end());
        return row;
    }

    void unflatten(const std::vector<T>& row)
    {
        std::size_t k = 0;
        for (auto& c : chromosomes)
            for (auto& g : c)
                if (k < row.size()) g = row[k++];
        evaluated = false;
    }

    void reset_nsga2_data()
    {
        rank = -1;
        crowding_distance = T(0);
        dominated_solutions.clear();
        domination_count = 0;
    }
};
