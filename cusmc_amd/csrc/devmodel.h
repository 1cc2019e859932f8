// devmodel.h — device-side model tables of the run kernel (host and device).
//
// cuSMC keeps a level-major NodeInfo table with 40-byte nodes, 152-byte
// GuardInfo records and one device allocation per expression node
// (simulation/state/shared_model_state.cuh:45-217, shared_model_state.cu:
// 15-161).  Here the model is one arena of small fixed-size records indexed
// by 32-bit offsets, and every expression is either an inline constant, an
// inline variable reference or a short run of register bytecode:
//
//   expression reference (uint32):  bits 31..30 kind
//       SMC_REF_CONST  index into consts[]
//       SMC_REF_VAR    variable id (an integer variable)
//       SMC_REF_CODE   bits 0..23 first instruction, bits 24..29 count
//       SMC_REF_CLOCK  variable id of a clock
//
//   instruction (8 bytes): op, dst register, up to three operands; an
//   operand (uint16) is a register r0..r3, a variable, a constant or a
//   clock: bits 15..14 kind (0 reg, 1 var, 2 const, 3 clock), bits 13..0
//   index.  The value of an expression is register r0 after its last
//   instruction.
//
// Clocks are stored as reset stamps: a run's variable array holds, for a
// clock x, the time s at which x would have been 0, and its value is t - s
// with t the run's clock time, kept just before the variables (vars[-1]).
// Time then advances with one store, however many clocks the network has;
// assigning x := v stores s = t - v.
#pragma once

#include <stdint.h>
#include <type_traits>

#define SMC_REF_CONST 0u
#define SMC_REF_VAR 1u
#define SMC_REF_CODE 2u
#define SMC_REF_CLOCK 3u
#define SMC_REF_KIND(r) ((r) >> 30)
#define SMC_REF_MAKE(kind, payload) (((uint32_t)(kind) << 30) | (uint32_t)(payload))
#define SMC_REF_CODE_MAKE(first, count) (((uint32_t)SMC_REF_CODE << 30) | ((uint32_t)(count) << 24) | (uint32_t)(first))

#define SMC_OPND_REG 0u
#define SMC_OPND_VAR 1u
#define SMC_OPND_CONST 2u
#define SMC_OPND_CLOCK 3u
#define SMC_OPND(kind, idx) ((uint16_t)(((kind) << 14) | (idx)))

#define SMC_NREGS 4
#define SMC_MAX_CODE_LEN 63
#define SMC_MAX_OPND_INDEX 16383
#define SMC_MAX_EDGES_PER_LOC 64

// Delay classes of a location (simulation.cu:437-540).
enum smc_delay_class {
    SMC_K_ZERO = 0,     // branchpoint, urgent, committed: delay 0
    SMC_K_EXP = 1,      // exponential at the location's rate
    SMC_K_BOUNDED = 2   // uniform up to the nearest clock upper bound, else exponential
};

// Location flags.
enum {
    SMC_LF_TIME_FREE = 1,   // rate and spontaneous guards read no clock: enabling and rate
                            // change only with transitions
    SMC_LF_DEAD = 2,        // no spontaneous edge can ever fire (each has a `!=` conjunct, or
                            // there is none): always blocked
    SMC_LF_GUARDED = 4,     // blocked or not by the variables: every spontaneous edge without
                            // `!=` has a conjunct that reads no clock (DevLoc.block_first)
    SMC_LF_RECV_FAST = 8,   // every receive edge commutes (see DevUpd)
    SMC_LF_INL_INV = 16,    // one invariant with a constant bound, inline in DevLoc.inl_*
    SMC_LF_INL_BLK = 32     // the blocking test is one conjunct `var op constant`, inline
};

struct DevInsn {
    uint8_t op;
    uint8_t dst;
    uint16_t a, b, c;
};

struct DevLoc {             // 48 bytes: 32 of location, 16 of inline data
    uint8_t klass;
    uint8_t type;
    uint16_t n_inv;         // clock upper-bound invariants
    uint32_t rate;          // expression reference
    uint32_t inv_first;
    uint32_t edge_first;    // edges [edge_first, edge_first + edge_n)
    double rate_value;      // the rate when it is a constant (rate_const)
    uint16_t edge_n;
    uint8_t rate_const;
    uint8_t flags;          // SMC_LF_*
    uint32_t block_first;   // SMC_LF_GUARDED: the blocking test, guard records from here on
    // The commonest invariant (one, constant bound) or blocking test (one
    // conjunct `var op constant`) as plain values, read with the location
    // from the LDS-staged tables (one dependent trip less in the race's slot
    // loop).  Kernels reading the tables from global memory load the first
    // 32 bytes only and take the general records (load_loc).
    double inl_value;       // SMC_LF_INL_INV: the bound; SMC_LF_INL_BLK: the constant
    uint16_t inl_var;       // the clock bounded, or the variable compared
    uint8_t inl_op;         // SMC_LF_INL_INV: 1 if strict; SMC_LF_INL_BLK: smc_cons_op
    uint8_t inl_pad0;
    uint32_t inl_pad1;
};

struct DevInv {             // x <= bound  (strict: x < bound), 16 bytes
    uint16_t var;
    uint8_t strict;
    uint8_t bound_const;    // bound is the constant bound_value
    uint32_t bound;         // expression reference
    double bound_value;
};

struct DevEdge {            // 32 bytes
    int32_t dst;            // destination relative to the source location
    int32_t channel;
    uint32_t weight;
    uint32_t guard_first;
    uint32_t upd_first;
    uint16_t guard_n;
    uint16_t upd_n;
    uint8_t dst_goal;       // destination is a goal location
    uint8_t never;          // has a `!=` guard: never enabled (simulation.cu:249-251)
    uint16_t pad0;
    uint32_t pad1;
};

// DevGuard.flags
enum {
    SMC_GF_CLOCK_FREE = 1,  // neither var nor rhs reads a clock: fixed between transitions
    SMC_GF_CLOCK = 2,       // var is a clock
    SMC_GF_END_EDGE = 4,    // blocking test: last conjunct of its edge
    SMC_GF_END_LOC = 8      // blocking test: last record of the location
};

// A guard-blocked location's blocking test is a run of guard records (from
// DevLoc.block_first): the clock-free conjuncts of each spontaneous edge
// without `!=`, edge after edge.  The location is blocked when every edge
// has a false conjunct among them.

struct DevGuard {           // variables[var] op rhs
    uint16_t var;
    uint8_t op;             // smc_cons_op, never SMC_C_NE
    uint8_t flags;          // SMC_GF_*
    uint32_t rhs;
};

// Update kinds: a commuting receive edge has only SMC_U_ADD updates
// (integer variable += integer constant) and SMC_U_SET updates (a variable
// written by one component only := constant), so the receivers of one
// broadcast can fire in any order with the same result.
enum { SMC_U_EXPR = 0, SMC_U_ADD = 1, SMC_U_SET = 2 };

struct DevUpd {             // 16 bytes
    uint16_t var;
    uint8_t kind;           // SMC_U_*
    uint8_t clock;          // var is a clock (stored as a reset stamp)
    uint32_t expr;
    double value;           // the added constant (ADD) or the value (SET)
};

// The tables live in device global memory.  Device code sees their pointers
// in the global address space, so reads compile to global loads (their own
// wait counter) rather than flat loads, which the compiler must also order
// against LDS traffic; host code sees plain pointers of the same size.  A
// run kernel may stage the tables in LDS and read them through a second
// instance of the same record (DevModelT<SmcLdsTables>).
#if defined(__HIP_DEVICE_COMPILE__)
#define SMC_GLOBAL __attribute__((address_space(1)))
#define SMC_LDS __attribute__((address_space(3)))
#else
#define SMC_GLOBAL
#define SMC_LDS
#endif

struct SmcGlobalTables {
    template <class T> using ptr = const SMC_GLOBAL T *;
};
struct SmcLdsTables {
    template <class T> using ptr = const SMC_LDS T *;
};
// Location records from global memory, every other hot table from the LDS
// copy: for networks whose many location records would crowd LDS while
// their edge, guard, update and expression records (shared by template
// instances) are few.
struct SmcMixedTables {
    template <class T>
    using ptr = typename std::conditional<std::is_same<T, DevLoc>::value, const SMC_GLOBAL T *,
                                          const SMC_LDS T *>::type;
};

// Everything the run kernel needs to know about one model; passed by value.
template <class A>
struct DevModelT {
    typename A::template ptr<DevLoc> locs;
    typename A::template ptr<DevInv> invs;
    typename A::template ptr<DevEdge> edges;
    typename A::template ptr<DevGuard> guards;
    typename A::template ptr<DevUpd> upds;
    typename A::template ptr<double> consts;
    typename A::template ptr<DevInsn> code;
    const SMC_GLOBAL double *var_init;     // stored form at time 0: initial value, -initial value for a clock
    const SMC_GLOBAL int32_t *comp_init_loc;
    const SMC_GLOBAL uint8_t *var_clock;   // [n_vars] 1: the variable is a clock
    const SMC_GLOBAL int32_t *comp_first_loc;  // [n_comp + 1] location range of each component
    const SMC_GLOBAL int32_t *loc_edge0;       // [n_locs] the network's index of each location's first edge
                                               // (edge records are shared: digests name network edges)
    int32_t n_comp;
    int32_t n_vars;
    int32_t n_clocks;
    int32_t n_leaves;           // next power of two >= n_comp
    int32_t n_locs;
    int32_t n_edges;
    int32_t freeze;             // no rate reads a clock: blocked components may leave the race
    int32_t guarded;            // some location is flagged SMC_LF_GUARDED
};
using DevModel = DevModelT<SmcGlobalTables>;
