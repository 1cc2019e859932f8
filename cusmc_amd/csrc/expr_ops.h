// expr_ops.h — value semantics of one expression operator, shared by the
// host-side constant folder and the run kernel's bytecode interpreter.
//
// Follows cuSMC's stack machine evaluate_expression_node_coalesced
// (simulation/expressions.cu:10-95) with its tolerance DBL_EPSILON redefined
// to 1e-5 (include/common_macros.h:72):
//   %  : int(a) % int(b + 1e-5)            (expressions.cu:47-50)
//   && : |a| > 1e-5 and |b| > 1e-5          (:51-54)     || likewise (:55-58)
//   == : |a - b| <= 1e-5, != : > 1e-5       (:75-82)
//   !  : |a| < 1e-5                         (:83-85)
//   ?: : |a| > 1e-5 selects b               (pn_compile_visitor.cpp:44-63 with
//        the skip test of expressions.cu:105-111)
//   random(a) : 1.0                          (:18-20)
// Deliberate, documented differences: integer conversions saturate instead
// of being undefined, `%` by zero gives 0, and `^` uses the reproducible
// smc_pow (include/smc_stream.h) instead of the platform pow().
#pragma once

#include <math.h>
#include <stdint.h>
#include "smc_network.h"
#include "smc_stream.h"

#define SMC_TOL 1e-5

SMC_HD int smc_to_int(double x) {
    if (!(x == x)) return 0;
    if (x >= 2147483647.0) return 2147483647;
    if (x <= -2147483648.0) return (-2147483647 - 1);
    return (int)x;
}

SMC_HD double smc_apply_op(int op, double a, double b, double c) {
    switch (op) {
    case SMC_E_RANDOM: return 1.0;
    case SMC_E_ADD: return a + b;
    case SMC_E_SUB: return a - b;
    case SMC_E_MUL: return a * b;
    case SMC_E_DIV: return a / b;
    case SMC_E_POW: return smc_pow(a, b);
    case SMC_E_NEG: return -a;
    case SMC_E_SQRT: return sqrt(a);
    case SMC_E_MOD: {
        int ia = smc_to_int(a), ib = smc_to_int(b + SMC_TOL);
        if (ib == 0 || (ia == (-2147483647 - 1) && ib == -1)) return 0.0;
        return (double)(ia % ib);
    }
    case SMC_E_AND: return (fabs(a) > SMC_TOL && fabs(b) > SMC_TOL) ? 1.0 : 0.0;
    case SMC_E_OR: return (fabs(a) > SMC_TOL || fabs(b) > SMC_TOL) ? 1.0 : 0.0;
    case SMC_E_LE: return a <= b ? 1.0 : 0.0;
    case SMC_E_GE: return a >= b ? 1.0 : 0.0;
    case SMC_E_LT: return a < b ? 1.0 : 0.0;
    case SMC_E_GT: return a > b ? 1.0 : 0.0;
    case SMC_E_EQ: return fabs(a - b) <= SMC_TOL ? 1.0 : 0.0;
    case SMC_E_NE: return fabs(a - b) > SMC_TOL ? 1.0 : 0.0;
    case SMC_E_NOT: return fabs(a) < SMC_TOL ? 1.0 : 0.0;
    case SMC_E_COND: return fabs(a) > SMC_TOL ? b : c;
    default: return 0.0;
    }
}

// Number of operands an operator reads.
SMC_HD int smc_op_arity(int op) {
    switch (op) {
    case SMC_E_LIT: case SMC_E_VAR: return 0;
    case SMC_E_RANDOM: case SMC_E_NEG: case SMC_E_SQRT: case SMC_E_NOT: return 1;
    case SMC_E_COND: return 3;
    default: return 2;
    }
}
