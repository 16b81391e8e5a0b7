// loader_test.cpp - parses the fixture files in tests/golden/data with the
// facade's config loaders and prints what they hold as JSON (no GPU work).
//   loader_test <data dir> <keys file>
// The keys file holds rows of 7 keys for the bi-objective TSP instance;
// their objectives are summed in the order of configs/multi_tsp_config.hpp:54-82.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "brkga/configs/knapsack_config.hpp"
#include "brkga/configs/multi_tsp_config.hpp"
#include "brkga/configs/tsp_config.hpp"
#include "brkga/configs/tspj_config.hpp"

static void print_vec(const char* key, const std::vector<float>& v, bool comma = true)
{
    std::printf("\"%s\": [", key);
    for (std::size_t i = 0; i < v.size(); ++i) std::printf("%.9g%s", v[i], i + 1 < v.size() ? ", " : "");
    std::printf("]%s\n", comma ? "," : "");
}

static void print_mat(const char* key, const std::vector<std::vector<float>>& m)
{
    std::printf("\"%s\": [", key);
    for (std::size_t i = 0; i < m.size(); ++i) {
        std::printf("[");
        for (std::size_t j = 0; j < m[i].size(); ++j) std::printf("%.9g%s", m[i][j], j + 1 < m[i].size() ? ", " : "");
        std::printf("]%s", i + 1 < m.size() ? ", " : "");
    }
    std::printf("],\n");
}

int main(int argc, char** argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: loader_test <data dir> <keys file>\n");
        return 2;
    }
    const std::string dir = argv[1];
    std::printf("{\n");
    auto tsp = TSPConfig<float>::load_from_file(dir + "/tiny8.tsp");
    std::printf("\"tsp_name\": \"%s\", \"tsp_n\": %d, \"tsp_kind\": %d,\n", tsp->get_instance_name().c_str(),
                tsp->get_num_cities(), tsp->native_kind());
    std::vector<float> xy;
    for (const auto& c : tsp->get_coordinates()) {
        xy.push_back(c.first);
        xy.push_back(c.second);
    }
    print_vec("tsp_xy", xy);
    for (const char* f : {"knap_small.txt", "knap_cap.csv", "knap_nocap.csv"}) {
        std::string key(f);
        for (auto& ch : key)
            if (ch == '.') ch = '_';
        auto k = KnapsackConfig<float>::load_from_file(dir + "/" + f);
        print_vec((key + "_w").c_str(), k->get_weights());
        print_vec((key + "_v").c_str(), k->get_values());
        print_vec((key + "_cap").c_str(), {k->get_capacity()});
        std::printf("\"%s_name\": \"%s\",\n", key.c_str(), k->get_instance_name().c_str());
    }
    for (const char* f : {"mini_TT.csv", "mini_JT.csv"}) {
        auto t = TSPJConfig<float>::load_from_file(dir + "/" + f);
        std::string key = std::string("tspj_from_") + (f[5] == 'T' ? "tt" : "jt");
        print_mat((key + "_travel").c_str(), t->get_travel_times());
        print_mat((key + "_jobs").c_str(), t->get_job_durations());
        std::printf("\"%s_name\": \"%s\", \"%s_n\": %d,\n", key.c_str(), t->get_instance_name().c_str(), key.c_str(),
                    t->get_num_cities());
    }
    const auto dist = MultiTSPConfig<float>::load_tsplib(dir + "/tiny8.tsp").first;
    const auto time = MultiTSPConfig<float>::generate_time_matrix(dist, 0.7f, 1.8f);
    auto mt = MultiTSPConfig<float>::load_from_file(dir + "/tiny8.tsp");
    std::printf("\"multi_n\": %d,\n", mt->get_num_cities());
    std::ifstream kf(argv[2]);
    std::vector<float> obj;
    std::string line;
    while (std::getline(kf, line)) {
        std::istringstream in(line);
        std::vector<std::pair<float, int>> keyed;
        float k;
        while (in >> k) keyed.push_back({k, (int)keyed.size()});
        if (keyed.empty()) continue;
        std::sort(keyed.begin(), keyed.end());
        float d = 0.0f, t = 0.0f;
        int cur = 0;
        for (const auto& kc : keyed) {
            d += dist[cur][kc.second + 1];
            t += time[cur][kc.second + 1];
            cur = kc.second + 1;
        }
        d += dist[cur][0];
        t += time[cur][0];
        obj.push_back(d);
        obj.push_back(t);
    }
    print_vec("multi_objectives", obj, false);
    std::printf("}\n");
    return 0;
}
