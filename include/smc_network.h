/*
 * smc_network.h — plain-C description of a network of stochastic timed
 * automata, the input format of the simulation path.
 *
 * This is the data that cuSMC's init_shared_model_state() consumes
 * (simulation/state/shared_model_state.cuh:219-225, built from the parser's
 * node_edge_map / node_map / variable registry), expressed as flat arrays of
 * plain structs with integer indices instead of pointers so that it crosses a
 * C ABI, a ctypes binding or a file without translation.
 *
 * Numbering that is part of the contract:
 *   - expression operators use cuSMC's expr::operators values
 *     (include/engine/domain.h:39-75) so a tree built by either front end
 *     means the same thing;
 *   - constraint operators use constraint::operators (domain.h:100-109);
 *   - location types use node::node_types (domain.h:140-147):
 *     0 location, 1 goal, 2 branchpoint, 3 urgent, 4 committed;
 *   - channels follow the sign convention of domain.h:162-166 and
 *     uppaal_xml_parser.cpp:600: 0 = no synchronisation, c > 0 = broadcast
 *     send on channel c, c < 0 = receive on channel -c;
 *   - variable ids are the parser's declaration order (globals first, then
 *     each template's locals; variable_usage_visitor.h:153-185); they are the
 *     ids the `-q ... <varId>` command line refers to.
 *
 * Every array is owned by the caller; consumers copy what they keep.
 */
#ifndef SMC_NETWORK_H
#define SMC_NETWORK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Expression operators: identical values to cuSMC expr::operators. */
enum smc_expr_op {
    SMC_E_LIT = 0,   /* value                                  */
    SMC_E_VAR = 1,   /* variables[var]                          */
    SMC_E_RANDOM = 2,/* random(a): evaluates to 1.0 (expressions.cu:18-20) */
    SMC_E_ADD = 3, SMC_E_SUB = 4, SMC_E_MUL = 5, SMC_E_DIV = 6,
    SMC_E_POW = 7, SMC_E_NEG = 8, SMC_E_SQRT = 9, SMC_E_MOD = 10,
    SMC_E_AND = 11, SMC_E_OR = 12,
    SMC_E_LE = 13, SMC_E_GE = 14, SMC_E_LT = 15, SMC_E_GT = 16,
    SMC_E_EQ = 17, SMC_E_NE = 18, SMC_E_NOT = 19,
    SMC_E_COND = 20  /* a ? b : c                               */
};

/* Constraint operators: identical values to cuSMC constraint::operators. */
enum smc_cons_op {
    SMC_C_LE = 0, SMC_C_LT = 1, SMC_C_GE = 2, SMC_C_GT = 3,
    SMC_C_EQ = 4, SMC_C_NE = 5
};

enum smc_loc_type {
    SMC_LOC_NORMAL = 0, SMC_LOC_GOAL = 1, SMC_LOC_BRANCH = 2,
    SMC_LOC_URGENT = 3, SMC_LOC_COMMITTED = 4
};

enum smc_var_kind { SMC_VAR_INT = 0, SMC_VAR_CLOCK = 1 };

/* One expression tree node.  Children are indices into smc_network.exprs,
 * -1 when absent.  For SMC_E_COND: a = condition, b = value when true,
 * c = value when false (uppaal_xml_parser.cpp:515-521).  Unary operators use
 * a only. */
typedef struct smc_expr_node {
    int32_t op;
    int32_t a, b, c;
    int32_t var;        /* SMC_E_VAR */
    int32_t reserved;
    double value;       /* SMC_E_LIT */
} smc_expr_node;

/* A guard or invariant conjunct `lhs op rhs`.  uses_variable != 0 means the
 * left side is the variable `var` (cuSMC GuardInfo with var_info,
 * shared_model_state.cu:355-425); otherwise the left side is the expression
 * `lhs`, and the simulator ignores the conjunct (simulation.cu:209,464). */
typedef struct smc_constraint {
    int32_t op;
    int32_t uses_variable;
    int32_t var;
    int32_t lhs;        /* expression index, used when uses_variable == 0 */
    int32_t rhs;        /* expression index                                */
} smc_constraint;

/* A location.  Its outgoing edges are edges[edge_first .. edge_first+edge_count)
 * in the order the transitions appear in the model file; its invariant
 * conjuncts are constraints[inv_first .. inv_first+inv_count).
 * rate = expression index of the exponential rate, -1 for the default rate 1
 * (uppaal_xml_parser.cpp:385-387). */
typedef struct smc_location {
    int32_t type;
    int32_t rate;
    int32_t inv_first, inv_count;
    int32_t edge_first, edge_count;
} smc_location;

/* An edge.  weight = expression index of the branch weight, -1 for 1.0
 * (uppaal_xml_parser.cpp:472-477). */
typedef struct smc_edge {
    int32_t src, dst;
    int32_t channel;
    int32_t weight;
    int32_t guard_first, guard_count;
    int32_t upd_first, upd_count;
} smc_edge;

/* An assignment `variables[var] := expr`, applied in order. */
typedef struct smc_update {
    int32_t var;
    int32_t expr;
} smc_update;

typedef struct smc_network {
    int32_t n_components;
    int32_t n_vars;
    int32_t n_locations;
    int32_t n_edges;
    int32_t n_constraints;
    int32_t n_updates;
    int32_t n_exprs;
    int32_t reserved;
    const int32_t *var_kind;        /* [n_vars]  smc_var_kind                  */
    const double *var_init;         /* [n_vars]  initial value                 */
    const int32_t *comp_first_loc;  /* [n_components+1] locations of component
                                       c are [comp_first_loc[c], comp_first_loc[c+1]) */
    const int32_t *comp_init_loc;   /* [n_components] initial location index   */
    const smc_location *locations;  /* [n_locations]                           */
    const smc_edge *edges;          /* [n_edges]                               */
    const smc_constraint *constraints; /* [n_constraints] guards and invariants */
    const smc_update *updates;      /* [n_updates]                             */
    const smc_expr_node *exprs;     /* [n_exprs]                               */
} smc_network;

/* Query kinds of the `-q` command line (main.cu:69-97, smc.cu:345-436). */
enum smc_query_kind {
    SMC_QUERY_REACH = 0,   /* Pr[<=T](<> goal location)                    */
    SMC_QUERY_VAR = 1      /* max/min of one variable over each run        */
};

typedef struct smc_query {
    int32_t kind;
    int32_t var;           /* SMC_QUERY_VAR: variable id                   */
    int32_t is_max;        /* SMC_QUERY_VAR: 1 = max, 0 = min              */
    int32_t is_estimate;   /* SMC_QUERY_VAR: 1 = mean of the per-run value,
                              0 = fraction of runs beyond threshold         */
    double threshold;      /* SMC_QUERY_VAR probability mode               */
    double time_bound;
    uint32_t max_steps;    /* race rounds after which a run is cut off (a
                              Zeno cycle of zero delays never reaches the time
                              bound); 0 = SMC_DEFAULT_MAX_STEPS              */
    uint32_t reserved;
} smc_query;

#define SMC_DEFAULT_MAX_STEPS (1u << 24)

#ifdef __cplusplus
}
#endif

#endif /* SMC_NETWORK_H */
