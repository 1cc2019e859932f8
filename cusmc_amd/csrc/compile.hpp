// compile.hpp — smc_network -> device model tables and register bytecode.
#pragma once

#include <string>
#include <vector>
#include "devmodel.h"
#include "smc_network.h"

namespace smc {

struct ModelImage {
    std::vector<DevLoc> locs;
    std::vector<DevInv> invs;
    std::vector<DevEdge> edges;
    std::vector<DevGuard> guards;
    std::vector<DevUpd> upds;
    std::vector<double> consts;
    std::vector<DevInsn> code;
    std::vector<double> var_init;
    std::vector<int32_t> comp_init_loc;
    std::vector<int32_t> comp_first_loc;
    std::vector<uint8_t> var_clock;     // per variable: 1 for a clock
    std::vector<int32_t> var_kind;
    std::vector<int32_t> loc_edge0;     // per location: the network's index of its first edge
    int n_comp = 0, n_vars = 0, n_leaves = 1;
    bool freeze = true;   // no location's rate reads a clock
    bool guarded = false; // some location is blocked or not by the variables
};

// Throws std::runtime_error (message prefixed "limit:" for size limits).
ModelImage compile_model(const smc_network &net);

// Host evaluation of a compiled expression reference (tests, debugging).
double eval_ref_host(const ModelImage &m, uint32_t ref, const double *vars);

}  // namespace smc
