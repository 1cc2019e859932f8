// cli_main.cpp — the `cusmc` command: cuSMC's command line over the C ABI.
//
// The command-line contract of the reference binary (main.cu:10-110,
// smc.cu:283-491):
//   -m/--model <xml>      model file (default ../xml_files/LargeModels/aloha_50.xml)
//   -s/--seed <n>         random seed (default 1235)
//   -q/--query <sims> <timeBound> <varThreshold> <varId> <isMax 0/1>
//              <isEstimate 0/1> <locQuery>
//                         varId -1 selects the location query "Process.location"
//   -h/--help             usage
// A run exits with status 1, help and argument errors with 0 (main.cu:27,
// :38).  stdout carries, in order: the variable count, the settings block,
// three `Time taken:` lines, then `Result: <r>` (variable queries,
// smc.cu:401) or the hits block (location queries, statistics.cuh:165-169).
// The timing matrix `../results5.csv` (rows: model file names, columns:
// `MC<n>_Sim<n>_Time<t.2f>`, `;`-separated, cells std::to_string of the
// milliseconds) is updated in place (simulation/write_csv.cu:15-99).
// Additions: --json prints one JSON record with the estimate, its
// Chernoff-Hoeffding half-width and throughput; SMC_RESULTS_CSV overrides
// the CSV path; SMC_DEVICE selects the GPU.
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

// What the command line sets (main.cu:12-23 defaults).
struct Settings {
    std::string model = "../xml_files/LargeModels/aloha_50.xml";
    int seed = 1235;
    int simulations = 16384;
    int time_bound = 100;
    int threshold = -1;
    int var_id = 1;
    bool is_max = true;
    bool is_estimate = true;
    std::string loc_query;
    bool json = false;
};

// One option: its spellings, how many values follow, and the option name
// its "requires a value" message uses (the reference names -m as -i and -q
// as -r in these messages; kept, they are part of the observable output).
struct OptionSpec {
    const char *shortname, *longname;
    int values;
    const char *missing_name;
};

const OptionSpec kOptions[] = {
    {"-m", "--model", 1, "-i"},
    {"-s", "--seed", 1, "-s"},
    {"-q", "--query", 7, "-r"},
    {"-h", "--help", 0, nullptr},
    {"--json", "--json", 0, nullptr},
};

void print_usage() {
    std::cout << "Use -m or --model, followed by a path, for inputting a path the the model XML file.\n"
                 "Use -s or --seed, followed by a number, to initialize curand with a constant seed. (0 = random seed)\n"
                 "Use -q or --query, followed by : number of simulation, time bound, a variable threshold, a variable "
                 "name (id), Max=1 or Min=0, Estimate=1 or Probability=0."
              << std::endl;
}

// std::stoi with the reference's two error messages; false on failure.
bool to_int(const std::string &s, int *out) {
    try {
        *out = std::stoi(s);
        return true;
    } catch (const std::invalid_argument &e) {
        std::cerr << "Error: invalid argument: " << e.what() << std::endl;
    } catch (const std::out_of_range &e) {
        std::cerr << "Error: number too big: " << e.what() << std::endl;
    }
    return false;
}

// Fills `st` from the arguments; false means "stop without running" (help
// or an error, both exit 0 like the reference).
bool parse_command_line(int argc, char **argv, Settings *st) {
    int i = 1;
    while (i < argc) {
        const std::string arg = argv[i];
        const OptionSpec *spec = nullptr;
        for (const OptionSpec &o : kOptions)
            if (arg == o.shortname || arg == o.longname) spec = &o;
        if (!spec) {
            std::cerr << "Error: unknown option: " << arg << std::endl;
            return false;
        }
        if (i + spec->values >= argc) {
            std::cerr << "Error: " << spec->missing_name << " option requires a value" << std::endl;
            return false;
        }
        std::vector<std::string> v(argv + i + 1, argv + i + 1 + spec->values);
        i += 1 + spec->values;
        const std::string name = spec->shortname;
        if (name == "-h") {
            print_usage();
            return false;
        }
        if (name == "--json") {
            st->json = true;
        } else if (name == "-m") {
            st->model = v[0];
        } else if (name == "-s") {
            if (!to_int(v[0], &st->seed)) return false;
        } else {  // -q: six integers then the location query
            st->loc_query = v[6];
            int n[6];
            for (int k = 0; k < 6; ++k)
                if (!to_int(v[k], &n[k])) return false;
            st->simulations = n[0];
            st->time_bound = n[1];
            st->threshold = n[2];
            st->var_id = n[3];
            st->is_max = n[4] != 0;
            st->is_estimate = n[5] != 0;
        }
    }
    return true;
}

// The timing matrix: a grid of `;`-separated cells, corner cell empty,
// configurations along the first row, model files down the first column.
class TimingMatrix {
public:
    explicit TimingMatrix(const std::string &path) : path_(path) {
        std::ifstream f(path);
        for (std::string line; f && std::getline(f, line);) {
            std::vector<std::string> cells;
            std::stringstream ls(line);
            for (std::string c; std::getline(ls, c, ';');) cells.push_back(c);
            grid_.push_back(cells);
        }
        if (grid_.empty()) grid_.push_back({""});
        if (grid_[0].empty()) grid_[0].push_back("");
    }

    void set(const std::string &file, const std::string &config, const std::string &value) {
        const size_t col = column_of(config), row = row_of(file);
        for (auto &r : grid_)
            if (r.size() <= col) r.resize(col + 1);
        grid_[row][col] = value;
    }

    bool save() const {
        std::ofstream out(path_);
        if (!out.is_open()) return false;
        for (const auto &r : grid_) {
            for (size_t k = 0; k < r.size(); ++k) out << (k ? ";" : "") << r[k];
            out << "\n";
        }
        return true;
    }

private:
    std::string path_;
    std::vector<std::vector<std::string>> grid_;

    size_t column_of(const std::string &config) {
        auto &head = grid_[0];
        for (size_t c = 1; c < head.size(); ++c)
            if (head[c] == config) return c;
        head.push_back(config);
        return head.size() - 1;
    }
    size_t row_of(const std::string &file) {
        for (size_t r = 1; r < grid_.size(); ++r)
            if (!grid_[r].empty() && grid_[r][0] == file) return r;
        grid_.push_back({file});
        return grid_.size() - 1;
    }
};

void record_timing(const Settings &st, int components, double ms) {
    namespace fs = std::filesystem;
    const char *env = std::getenv("SMC_RESULTS_CSV");
    const std::string path = env && *env ? env : (fs::current_path().parent_path() / "results5.csv").string();
    std::ostringstream config;
    config << "MC" << components << "_Sim" << st.simulations << "_Time" << std::fixed << std::setprecision(2)
           << (double)st.time_bound;
    TimingMatrix m(path);
    m.set(fs::path(st.model).filename().string(), config.str(), std::to_string(ms));
    if (!m.save()) std::cerr << "Failed to open file for writing" << std::endl;
}

void print_settings(const Settings &st, int components, bool var_query) {
    std::cout << "=================\n\n"
              << "Running SMC with the following settings:\n"
              << "- Number of simulations: " << st.simulations << "\n"
              << "- Model: " << st.model << "\n"
              << "- Random seed: " << st.seed << "\n"
              << "- Time bound: " << st.time_bound << "\n"
              << "- Num components: " << components << "\n";
    if (var_query)
        std::cout << "- Logging type: variable query. Finding " << (st.is_max ? "max" : "min")
                  << " of variable with ID " << st.var_id << "\n";
    else
        std::cout << "- Logging type: location query on query: " << st.loc_query << "\n";
    std::cout << "=================\n" << std::endl;
}

int fail_run(const char *what, smc_model *model, smc_parsed *parsed) {
    std::cerr << "Error: " << what << std::endl;
    if (model) smc_model_destroy(model);
    if (parsed) smc_parsed_destroy(parsed);
    return 2;
}

}  // namespace

int main(int argc, char **argv) {
    Settings st;
    if (!parse_command_line(argc, argv, &st)) return 0;
    const bool var_query = st.var_id != -1;

    smc_parsed *parsed = nullptr;
    if (smc_parse_uppaal(st.model.c_str(), st.loc_query.c_str(), &parsed) != SMC_OK)
        return fail_run(smc_last_error(), nullptr, nullptr);
    const smc_network *net = smc_parsed_network(parsed);
    std::printf("Set number of variables to: %d\n", net->n_vars);
    std::fflush(stdout);

    smc_model *model = nullptr;
    const char *dev = std::getenv("SMC_DEVICE");
    if (smc_model_create(net, dev ? std::atoi(dev) : 0, &model) != SMC_OK)
        return fail_run(smc_last_error(), nullptr, parsed);
    print_settings(st, net->n_components, var_query);
    if (st.simulations <= 0) return fail_run("number of simulations must be positive", model, parsed);

    smc_query q{};
    q.kind = var_query ? SMC_QUERY_VAR : SMC_QUERY_REACH;
    q.var = var_query ? st.var_id : -1;
    q.is_max = st.is_max;
    q.is_estimate = st.is_estimate;
    q.threshold = st.threshold;
    q.time_bound = st.time_bound;
    smc_result res{};
    if (smc_run(model, &q, (uint64_t)(uint32_t)st.seed, 0, (uint64_t)st.simulations, nullptr, nullptr, &res) != SMC_OK)
        return fail_run(smc_last_error(), model, parsed);

    const double ms = res.kernel_ms;
    const long long ns = (long long)(ms * 1e6);
    std::cout << "Time taken: " << ms << " ms on file " << st.model << "\n"
              << "Time taken: " << ns << " on file " << st.model << "\n"
              << "Time taken: " << ns << " ns (" << (ns / 1000000.0) << " ms)" << std::endl;
    record_timing(st, net->n_components, ms);

    if (var_query) {
        std::cout << "Result: " << res.estimate << std::endl;
    } else {
        // the reference divides in float (statistics.cuh:148-149)
        const float p = (float)res.hits / (float)res.n_runs;
        std::cout << "Number of simulations: " << res.n_runs << "\n"
                  << "Number of hits: " << res.hits << "\n"
                  << "The answer to " << st.loc_query << " is " << p << std::endl;
    }
    if (st.json)
        std::printf("{\"n_runs\": %llu, \"hits\": %llu, \"estimate\": %.17g, \"ch_eps\": %.17g, \"clt_half\": %.17g, "
                    "\"kernel_ms\": %.6f, \"runs_per_sec\": %.6g, \"transitions_per_sec\": %.6g, \"steps\": %llu, "
                    "\"transitions\": %llu, \"truncated\": %llu}\n",
                    (unsigned long long)res.n_runs, (unsigned long long)res.hits, res.estimate, res.ch_eps,
                    res.clt_half, ms, res.n_runs / (ms / 1e3), res.transitions / (ms / 1e3),
                    (unsigned long long)res.steps, (unsigned long long)res.transitions,
                    (unsigned long long)res.truncated);
    smc_model_destroy(model);
    smc_parsed_destroy(parsed);
    return 1;
}
