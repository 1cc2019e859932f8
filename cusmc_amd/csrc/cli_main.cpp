// cli_main.cpp — the `cusmc` command: cuSMC's command line over the C ABI.
//
// Same options, defaults, output blocks and return codes as cuSMC's main.cu
// (main.cu:10-110) and smc() (smc.cu:283-491):
//   -m/--model <xml>      model file (default ../xml_files/LargeModels/aloha_50.xml)
//   -s/--seed <n>         random seed (default 1235)
//   -q/--query <sims> <timeBound> <varThreshold> <varId> <isMax 0/1>
//              <isEstimate 0/1> <locQuery>
//                         varId -1 selects the location query "Process.location"
//   -h/--help             usage; returns 0 like the reference
// The timing line and the ../results5.csv matrix follow
// run_statistical_model_checking (smc.cu:234-257) and writeTimingToCSV
// (simulation/write_csv.cu:15-99).  Additions: --json prints one JSON record
// with the estimate, its Chernoff-Hoeffding half-width and throughput;
// SMC_RESULTS_CSV overrides the CSV path; SMC_DEVICE selects the GPU.
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "smc_capi.h"

namespace {

struct StatConf {
    int simulations = 16384;
    int timeBound = 100;
    int variable_threshhold = -1;
    int variable_id = 1;
    bool isMax = true;
    bool isEstimate = true;
    std::string loc_query;
};

bool handle_args(int argc, char **argv, std::string *filename, int *seed, StatConf *st, bool *json) {
    for (int i = 1; i < argc; i++) {
        std::string arg = argv[i];
        if (arg == "-m" || arg == "--model") {
            if (i + 1 < argc) { *filename = argv[++i]; }
            else { std::cerr << "Error: -i option requires a value" << std::endl; return false; }
        } else if (arg == "-s" || arg == "--seed") {
            if (i + 1 < argc) {
                std::string s = argv[++i];
                try { *seed = std::stoi(s); }
                catch (const std::invalid_argument &e) { std::cerr << "Error: invalid argument: " << e.what() << std::endl; return false; }
                catch (const std::out_of_range &e) { std::cerr << "Error: number too big: " << e.what() << std::endl; return false; }
            } else { std::cerr << "Error: -s option requires a value" << std::endl; return false; }
        } else if (arg == "-q" || arg == "--query") {
            if (i + 7 < argc) {
                std::string a[6];
                for (auto &x : a) x = argv[++i];
                st->loc_query = argv[++i];
                try {
                    st->simulations = std::stoi(a[0]);
                    st->timeBound = std::stoi(a[1]);
                    st->variable_threshhold = std::stoi(a[2]);
                    st->variable_id = std::stoi(a[3]);
                    st->isMax = std::stoi(a[4]) != 0;
                    st->isEstimate = std::stoi(a[5]) != 0;
                } catch (const std::invalid_argument &e) { std::cerr << "Error: invalid argument: " << e.what() << std::endl; return false; }
                catch (const std::out_of_range &e) { std::cerr << "Error: number too big: " << e.what() << std::endl; return false; }
            } else { std::cerr << "Error: -r option requires a value" << std::endl; return false; }
        } else if (arg == "-h" || arg == "--help") {
            std::cout << "Use -m or --model, followed by a path, for inputting a path the the model XML file." << std::endl;
            std::cout << "Use -s or --seed, followed by a number, to initialize curand with a constant seed. (0 = random seed)" << std::endl;
            std::cout << "Use -q or --query, followed by : number of simulation, time bound, a variable threshold, a variable name (id), Max=1 or Min=0, Estimate=1 or Probability=0." << std::endl;
            return false;
        } else if (arg == "--json") {
            *json = true;
        } else {
            std::cerr << "Error: unknown option: " << arg << std::endl;
            return false;
        }
    }
    return true;
}

void write_timing_csv(const std::string &xml_path, int MC, int simulations, double timeBound, double ms) {
    namespace fs = std::filesystem;
    const char *env = std::getenv("SMC_RESULTS_CSV");
    const std::string csv = env && *env ? std::string(env) : (fs::current_path().parent_path() / "results5.csv").string();
    const std::string filename = fs::path(xml_path).filename().string();
    std::stringstream ss;
    ss << std::fixed << std::setprecision(2) << timeBound;
    const std::string config = "MC" + std::to_string(MC) + "_Sim" + std::to_string(simulations) + "_Time" + ss.str();
    std::vector<std::vector<std::string>> rows;
    if (fs::exists(csv)) {
        std::ifstream f(csv);
        std::string line;
        while (std::getline(f, line)) {
            std::vector<std::string> r;
            std::stringstream ls(line);
            std::string cell;
            while (std::getline(ls, cell, ';')) r.push_back(cell);
            rows.push_back(r);
        }
    }
    if (rows.empty()) rows.push_back({""});
    if (rows[0].empty()) rows[0].push_back("");
    size_t col = 0;
    for (size_t c = 1; c < rows[0].size(); ++c) if (rows[0][c] == config) { col = c; break; }
    if (!col) { col = rows[0].size(); rows[0].push_back(config); }
    size_t row = 0;
    for (size_t r = 1; r < rows.size(); ++r) if (!rows[r].empty() && rows[r][0] == filename) { row = r; break; }
    if (!row) { row = rows.size(); rows.push_back({filename}); }
    for (auto &r : rows) while (r.size() <= col) r.push_back("");
    rows[row][col] = std::to_string(ms);
    std::ofstream out(csv);
    if (!out.is_open()) { std::cerr << "Failed to open file for writing" << std::endl; return; }
    for (const auto &r : rows) {
        out << r[0];
        for (size_t i = 1; i < r.size(); ++i) out << ";" << r[i];
        out << "\n";
    }
}

}  // namespace

int main(int argc, char **argv) {
    std::string filename = "../xml_files/LargeModels/aloha_50.xml";
    int seed = 1235;
    StatConf st;
    bool json = false;
    if (!handle_args(argc, argv, &filename, &seed, &st, &json)) return 0;

    const bool var_query = st.variable_id != -1;
    smc_parsed *parsed = nullptr;
    if (smc_parse_uppaal(filename.c_str(), st.loc_query.c_str(), &parsed) != SMC_OK) {
        std::cerr << "Error: " << smc_last_error() << std::endl;
        return 2;
    }
    const smc_network *net = smc_parsed_network(parsed);
    printf("Set number of variables to: %d\n", net->n_vars);
    fflush(stdout);

    smc_model *model = nullptr;
    if (smc_model_create(net, std::getenv("SMC_DEVICE") ? std::atoi(std::getenv("SMC_DEVICE")) : 0, &model) != SMC_OK) {
        std::cerr << "Error: " << smc_last_error() << std::endl;
        smc_parsed_destroy(parsed);
        return 2;
    }
    std::cout << "=================\n\n";
    std::cout << "Running SMC with the following settings:" << std::endl;
    std::cout << "- Number of simulations: " << st.simulations << std::endl;
    std::cout << "- Model: " << filename << std::endl;
    std::cout << "- Random seed: " << seed << std::endl;
    std::cout << "- Time bound: " << st.timeBound << std::endl;
    std::cout << "- Num components: " << net->n_components << std::endl;
    if (var_query) {
        std::cout << "- Logging type: variable query. Finding " << (st.isMax ? "max" : "min")
                  << " of variable with ID " << st.variable_id << std::endl;
    } else {
        std::cout << "- Logging type: location query on query: " << st.loc_query << std::endl;
    }
    std::cout << "=================\n\n";

    smc_query q{};
    q.kind = var_query ? SMC_QUERY_VAR : SMC_QUERY_REACH;
    q.var = var_query ? st.variable_id : -1;
    q.is_max = st.isMax;
    q.is_estimate = st.isEstimate;
    q.threshold = st.variable_threshhold;
    q.time_bound = st.timeBound;
    smc_result res{};
    if (st.simulations <= 0 || smc_run(model, &q, (uint64_t)(uint32_t)seed, 0, (uint64_t)st.simulations, nullptr, nullptr, &res) != SMC_OK) {
        std::cerr << "Error: " << (st.simulations <= 0 ? "number of simulations must be positive" : smc_last_error()) << std::endl;
        smc_model_destroy(model);
        smc_parsed_destroy(parsed);
        return 2;
    }
    const double ms = res.kernel_ms;
    const long long ns = (long long)(ms * 1e6);
    std::cout << "Time taken: " << ms << " ms on file " << filename << std::endl;
    std::cout << "Time taken: " << ns << " on file " << filename << std::endl;
    std::cout << "Time taken: " << ns << " ns (" << (ns / 1000000.0) << " ms)" << std::endl;
    write_timing_csv(filename, net->n_components, st.simulations, st.timeBound, ms);

    if (var_query) {
        std::cout << "Result: " << res.estimate << std::endl;
    } else {
        std::cout << "Number of simulations: " << res.n_runs << std::endl;
        std::cout << "Number of hits: " << res.hits << std::endl;
        std::cout << "The answer to " << st.loc_query << " is " << (float)((double)res.hits / (double)res.n_runs) << std::endl;
    }
    if (json) {
        printf("{\"n_runs\": %llu, \"hits\": %llu, \"estimate\": %.17g, \"ch_eps\": %.17g, \"clt_half\": %.17g, "
               "\"kernel_ms\": %.6f, \"runs_per_sec\": %.6g, \"transitions_per_sec\": %.6g, \"steps\": %llu, "
               "\"transitions\": %llu, \"truncated\": %llu}\n",
               (unsigned long long)res.n_runs, (unsigned long long)res.hits, res.estimate, res.ch_eps, res.clt_half, ms,
               res.n_runs / (ms / 1e3), res.transitions / (ms / 1e3), (unsigned long long)res.steps,
               (unsigned long long)res.transitions, (unsigned long long)res.truncated);
    }
    smc_model_destroy(model);
    smc_parsed_destroy(parsed);
    return 1;
}
