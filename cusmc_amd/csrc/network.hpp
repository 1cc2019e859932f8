// network.hpp — owning C++ form of smc_network (include/smc_network.h).
#pragma once

#include <string>
#include <vector>
#include "smc_network.h"

namespace smc {

struct Network {
    std::vector<int32_t> var_kind;
    std::vector<double> var_init;
    std::vector<int32_t> comp_first_loc{0};
    std::vector<int32_t> comp_init_loc;
    std::vector<smc_location> locations;
    std::vector<smc_edge> edges;
    std::vector<smc_constraint> constraints;
    std::vector<smc_update> updates;
    std::vector<smc_expr_node> exprs;

    // names for diagnostics and queries
    std::vector<std::string> comp_names, loc_names, var_names;

    int add_expr(int op, int a = -1, int b = -1, int c = -1, int var = 0, double value = 0.0) {
        smc_expr_node n{};
        n.op = op; n.a = a; n.b = b; n.c = c; n.var = var; n.value = value;
        exprs.push_back(n);
        return (int)exprs.size() - 1;
    }
    int lit(double v) { return add_expr(SMC_E_LIT, -1, -1, -1, 0, v); }
    int var(int id) { return add_expr(SMC_E_VAR, -1, -1, -1, id, 0.0); }

    smc_network view() const {
        smc_network n{};
        n.n_components = (int32_t)comp_init_loc.size();
        n.n_vars = (int32_t)var_kind.size();
        n.n_locations = (int32_t)locations.size();
        n.n_edges = (int32_t)edges.size();
        n.n_constraints = (int32_t)constraints.size();
        n.n_updates = (int32_t)updates.size();
        n.n_exprs = (int32_t)exprs.size();
        n.var_kind = var_kind.data();
        n.var_init = var_init.data();
        n.comp_first_loc = comp_first_loc.data();
        n.comp_init_loc = comp_init_loc.data();
        n.locations = locations.data();
        n.edges = edges.data();
        n.constraints = constraints.data();
        n.updates = updates.data();
        n.exprs = exprs.data();
        return n;
    }
};

// Structural validation shared by every consumer; returns "" when valid.
std::string validate(const smc_network &n);

// Evaluate a constant expression tree with the run-time semantics; used for
// constant folding.  Returns false if the tree references a variable.
bool fold_constant(const smc_network &n, int e, double *out);

}  // namespace smc
