// network.cpp — validation and constant folding of smc_network.
#include "network.hpp"

#include <cstdio>
#include <string>
#include <vector>
#include "expr_ops.h"

namespace smc {

namespace {

std::string fmt(const char *f, long a, long b = 0) {
    char buf[256];
    snprintf(buf, sizeof buf, f, a, b);
    return buf;
}

// Checks expression e (and its subtree) without recursion depth surprises.
std::string check_expr(const smc_network &n, int e, int depth = 0) {
    if (e < 0 || e >= n.n_exprs) return fmt("expression index %ld out of range", e);
    if (depth > 256) return "expression nesting deeper than 256";
    const smc_expr_node &x = n.exprs[e];
    if (x.op < SMC_E_LIT || x.op > SMC_E_COND) return fmt("expression %ld has unknown operator %ld", e, x.op);
    if (x.op == SMC_E_VAR && (x.var < 0 || x.var >= n.n_vars))
        return fmt("expression %ld reads variable %ld which does not exist", e, x.var);
    int ar = smc_op_arity(x.op);
    const int kids[3] = {x.a, x.b, x.c};
    for (int i = 0; i < ar; ++i) {
        if (kids[i] == e) return fmt("expression %ld is its own operand", e);
        std::string s = check_expr(n, kids[i], depth + 1);
        if (!s.empty()) return s;
    }
    return "";
}

}  // namespace

std::string validate(const smc_network &n) {
    if (n.n_components <= 0) return "network has no components";
    if (n.n_vars < 0 || n.n_locations <= 0 || n.n_edges < 0 || n.n_constraints < 0 ||
        n.n_updates < 0 || n.n_exprs < 0)
        return "negative table size";
    if (n.n_vars > 0 && (!n.var_kind || !n.var_init)) return "variable arrays missing";
    if (!n.comp_first_loc || !n.comp_init_loc || !n.locations) return "component or location arrays missing";
    if (n.n_edges > 0 && !n.edges) return "edge array missing";
    if (n.n_constraints > 0 && !n.constraints) return "constraint array missing";
    if (n.n_updates > 0 && !n.updates) return "update array missing";
    if (n.n_exprs > 0 && !n.exprs) return "expression array missing";
    for (int v = 0; v < n.n_vars; ++v)
        if (n.var_kind[v] != SMC_VAR_INT && n.var_kind[v] != SMC_VAR_CLOCK)
            return fmt("variable %ld has unknown kind %ld", v, n.var_kind[v]);
    if (n.comp_first_loc[0] != 0) return "comp_first_loc[0] must be 0";
    for (int c = 0; c < n.n_components; ++c) {
        if (n.comp_first_loc[c + 1] <= n.comp_first_loc[c])
            return fmt("component %ld has no locations", c);
        if (n.comp_init_loc[c] < n.comp_first_loc[c] || n.comp_init_loc[c] >= n.comp_first_loc[c + 1])
            return fmt("initial location of component %ld is not one of its locations", c);
    }
    if (n.comp_first_loc[n.n_components] != n.n_locations)
        return "comp_first_loc does not cover every location";
    for (int c = 0; c < n.n_components; ++c) {
        for (int l = n.comp_first_loc[c]; l < n.comp_first_loc[c + 1]; ++l) {
            const smc_location &L = n.locations[l];
            if (L.type < SMC_LOC_NORMAL || L.type > SMC_LOC_COMMITTED) return fmt("location %ld has unknown type %ld", l, L.type);
            if (L.rate >= 0) { std::string s = check_expr(n, L.rate); if (!s.empty()) return s; }
            if (L.inv_count < 0 || L.inv_first < 0 || L.inv_first + L.inv_count > n.n_constraints)
                return fmt("location %ld invariant range out of bounds", l);
            if (L.edge_count < 0 || L.edge_first < 0 || L.edge_first + L.edge_count > n.n_edges)
                return fmt("location %ld edge range out of bounds", l);
            for (int k = 0; k < L.edge_count; ++k) {
                const smc_edge &E = n.edges[L.edge_first + k];
                if (E.src != l) return fmt("edge %ld is listed under location %ld but has another source", L.edge_first + k, l);
                if (E.dst < n.comp_first_loc[c] || E.dst >= n.comp_first_loc[c + 1])
                    return fmt("edge %ld leaves its component (destination %ld)", L.edge_first + k, E.dst);
            }
        }
    }
    for (int e = 0; e < n.n_edges; ++e) {
        const smc_edge &E = n.edges[e];
        if (E.src < 0 || E.src >= n.n_locations) return fmt("edge %ld source out of range", e);
        if (E.weight >= 0) { std::string s = check_expr(n, E.weight); if (!s.empty()) return s; }
        if (E.guard_count < 0 || E.guard_first < 0 || E.guard_first + E.guard_count > n.n_constraints)
            return fmt("edge %ld guard range out of bounds", e);
        if (E.upd_count < 0 || E.upd_first < 0 || E.upd_first + E.upd_count > n.n_updates)
            return fmt("edge %ld update range out of bounds", e);
    }
    for (int k = 0; k < n.n_constraints; ++k) {
        const smc_constraint &C = n.constraints[k];
        if (C.op < SMC_C_LE || C.op > SMC_C_NE) return fmt("constraint %ld has unknown operator %ld", k, C.op);
        if (C.uses_variable) {
            if (C.var < 0 || C.var >= n.n_vars) return fmt("constraint %ld reads variable %ld which does not exist", k, C.var);
        } else if (C.lhs >= 0) {
            std::string s = check_expr(n, C.lhs); if (!s.empty()) return s;
        }
        std::string s = check_expr(n, C.rhs); if (!s.empty()) return s;
    }
    for (int u = 0; u < n.n_updates; ++u) {
        const smc_update &U = n.updates[u];
        if (U.var < 0 || U.var >= n.n_vars) return fmt("update %ld assigns variable %ld which does not exist", u, U.var);
        std::string s = check_expr(n, U.expr); if (!s.empty()) return s;
    }
    return "";
}

bool fold_constant(const smc_network &n, int e, double *out) {
    const smc_expr_node &x = n.exprs[e];
    if (x.op == SMC_E_LIT) { *out = x.value; return true; }
    if (x.op == SMC_E_VAR) return false;
    int ar = smc_op_arity(x.op);
    double v[3] = {0, 0, 0};
    const int kids[3] = {x.a, x.b, x.c};
    if (x.op == SMC_E_RANDOM) { *out = 1.0; return true; }
    for (int i = 0; i < ar; ++i)
        if (!fold_constant(n, kids[i], &v[i])) return false;
    *out = smc_apply_op(x.op, v[0], v[1], v[2]);
    return true;
}

}  // namespace smc
