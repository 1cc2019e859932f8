// device_common.h — device helpers shared by the run kernel and the
// importance-sampling kernel: expression evaluation, edge enabling and
// choice, transitions, segment exchanges.
#pragma once

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "devmodel.h"
#include "expr_ops.h"
#include "smc_stream.h"

namespace {

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ int lowest_lane(uint64_t m) { return __ffsll((unsigned long long)m) - 1; }

// ------------------------------------------------------------ segment exchange
//
// Level k of a butterfly over an aligned segment: each lane receives the
// value held by the sibling block of 2^k lanes.  Levels 0-1 are DPP quad
// permutes (lane xor 1, xor 2), levels 2-3 DPP mirrors within 8 and 16
// lanes, level 4 a swizzle xor 16, level 5 a cross-half permute.  The
// mirrors reach a lane of the sibling block rather than the xor partner,
// which gives the same value whenever every lane of a block holds the same
// value — true at every level of a butterfly reduction.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
    int lo = __double2loint(x), hi = __double2hiint(x);
    lo = __builtin_amdgcn_update_dpp(0, lo, CTRL, 0xF, 0xF, false);
    hi = __builtin_amdgcn_update_dpp(0, hi, CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double seg_xchg(double x, int k) {
    switch (k) {
    case 0: return dpp_f64<0xB1>(x);   // quad_perm [1,0,3,2]
    case 1: return dpp_f64<0x4E>(x);   // quad_perm [2,3,0,1]
    case 2: return dpp_f64<0x141>(x);  // row_half_mirror
    case 3: return dpp_f64<0x140>(x);  // row_mirror
    case 4: {
        int lo = __double2loint(x), hi = __double2hiint(x);
        lo = __builtin_amdgcn_ds_swizzle(lo, 0x401F);  // lane xor 16 within 32
        hi = __builtin_amdgcn_ds_swizzle(hi, 0x401F);
        return __hiloint2double(hi, lo);
    }
    default: return __shfl_xor(x, 32);
    }
}


// Pairwise tree sum over the segment, as in the run kernel (in-lane levels,
// then the butterfly); every lane gets the total.
template <int C>
__device__ __forceinline__ double tree_total(const double (&leaf)[C], int S) {
    double tr[2 * C];
#pragma unroll
    for (int j = 0; j < C; ++j) tr[C + j] = leaf[j];
#pragma unroll
    for (int i = C - 1; i >= 1; --i) tr[i] = tr[2 * i] + tr[2 * i + 1];
    double s = (C == 1) ? leaf[0] : tr[1];
#pragma unroll
    for (int k = 0; k < 6; ++k)
        if ((1 << k) < S) s = s + seg_xchg(s, k);
    return s;
}

// ------------------------------------------------------------ table records

// A whole table record in 16-byte (8-byte) vector loads.  Copying a record
// field by field costs one narrow load per field used (seven for a
// location), each one a separate trip to the cache; the empty asm keeps the
// optimiser from narrowing the vector loads back to the fields used.
template <class T>
__device__ __forceinline__ T load_rec(const SMC_GLOBAL T *p) {
    static_assert(sizeof(T) % 8 == 0, "record size");
    T r;
    if constexpr (sizeof(T) % 16 == 0) {
        const SMC_GLOBAL uint4 *q = (const SMC_GLOBAL uint4 *)p;
        uint4 v[sizeof(T) / 16];
#pragma unroll
        for (int i = 0; i < (int)(sizeof(T) / 16); ++i) {
            v[i] = q[i];
            asm("" : "+v"(v[i].x), "+v"(v[i].y), "+v"(v[i].z), "+v"(v[i].w));
        }
        __builtin_memcpy(&r, v, sizeof(T));
    } else {
        const SMC_GLOBAL uint2 *q = (const SMC_GLOBAL uint2 *)p;
        uint2 v[sizeof(T) / 8];
#pragma unroll
        for (int i = 0; i < (int)(sizeof(T) / 8); ++i) {
            v[i] = q[i];
            asm("" : "+v"(v[i].x), "+v"(v[i].y));
        }
        __builtin_memcpy(&r, v, sizeof(T));
    }
    return r;
}

#if defined(__HIP_DEVICE_COMPILE__)
// The same from LDS (records 16-byte aligned there too).
template <class T>
__device__ __forceinline__ T load_rec(const SMC_LDS T *p) {
    static_assert(sizeof(T) % 8 == 0, "record size");
    T r;
    if constexpr (sizeof(T) % 16 == 0) {
        const SMC_LDS uint4 *q = (const SMC_LDS uint4 *)p;
        uint4 v[sizeof(T) / 16];
#pragma unroll
        for (int i = 0; i < (int)(sizeof(T) / 16); ++i) {
            v[i] = q[i];
            asm("" : "+v"(v[i].x), "+v"(v[i].y), "+v"(v[i].z), "+v"(v[i].w));
        }
        __builtin_memcpy(&r, v, sizeof(T));
    } else {
        const SMC_LDS uint2 *q = (const SMC_LDS uint2 *)p;
        uint2 v[sizeof(T) / 8];
#pragma unroll
        for (int i = 0; i < (int)(sizeof(T) / 8); ++i) {
            v[i] = q[i];
            asm("" : "+v"(v[i].x), "+v"(v[i].y));
        }
        __builtin_memcpy(&r, v, sizeof(T));
    }
    return r;
}
#endif

// A location record.  From the LDS-staged tables: all 48 bytes.  From global
// memory: the first 32 (one pair of 16-byte loads, as many cache lines as
// before the inline part existed), with the inline flags cleared, so the
// general invariant and guard records are used.
__device__ __forceinline__ DevLoc load_loc(const SMC_GLOBAL DevLoc *p) {
    const SMC_GLOBAL uint4 *q = (const SMC_GLOBAL uint4 *)p;
    uint4 v[3];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        v[i] = q[i];
        asm("" : "+v"(v[i].x), "+v"(v[i].y), "+v"(v[i].z), "+v"(v[i].w));
    }
    v[2] = uint4{0u, 0u, 0u, 0u};
    DevLoc r;
    __builtin_memcpy(&r, v, sizeof(DevLoc));
    r.flags &= (uint8_t)~(SMC_LF_INL_INV | SMC_LF_INL_BLK);
    return r;
}
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ DevLoc load_loc(const SMC_LDS DevLoc *p) { return load_rec(p); }
#endif

// Table pointer of the global arena rebased onto its LDS copy.
template <class T>
__device__ __forceinline__ const SMC_LDS T *in_lds(const SMC_GLOBAL T *p, const SMC_GLOBAL uint8_t *base,
                                                   const SMC_LDS uint8_t *copy) {
    return (const SMC_LDS T *)(copy + ((const SMC_GLOBAL uint8_t *)p - base));
}

// Workgroup-wide copy of n 16-byte words into LDS, eight loads in flight
// per thread before their stores (a plain loop waits on every load).
__device__ __forceinline__ void copy_to_lds(const SMC_GLOBAL uint4 *src, SMC_LDS uint4 *dst, uint32_t n) {
    const uint32_t step = blockDim.x;
    for (uint32_t i = threadIdx.x; i < n; i += 8 * step) {
        uint4 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            v[u] = uint4{0u, 0u, 0u, 0u};
            if (i + u * step < n) v[u] = src[i + u * step];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (i + u * step < n) dst[i + u * step] = v[u];
    }
}

// Copy the arena prefix [0, bytes) into the start of the workgroup's LDS
// (every thread of the workgroup takes part; ends with a barrier).
__device__ __forceinline__ void stage_tables(const SMC_GLOBAL uint8_t *src, uint32_t bytes, double *lds) {
    const SMC_GLOBAL uint4 *s4 = (const SMC_GLOBAL uint4 *)src;
    SMC_LDS uint4 *d4 = (SMC_LDS uint4 *)(SMC_LDS double *)lds;
    copy_to_lds(s4, d4, bytes / 16);
    __syncthreads();
}

// The model with every hot table but the location records read from the
// staged copy, which then starts at the invariants (base: their global
// address).
__device__ __forceinline__ DevModelT<SmcMixedTables> lds_model_mixed(const DevModel &mg,
                                                                     const SMC_GLOBAL uint8_t *base, double *lds) {
    const SMC_LDS uint8_t *tab = (const SMC_LDS uint8_t *)(SMC_LDS double *)lds;
    DevModelT<SmcMixedTables> m;
    m.locs = mg.locs;
    m.invs = in_lds(mg.invs, base, tab);
    m.edges = in_lds(mg.edges, base, tab);
    m.guards = in_lds(mg.guards, base, tab);
    m.upds = in_lds(mg.upds, base, tab);
    m.consts = in_lds(mg.consts, base, tab);
    m.code = in_lds(mg.code, base, tab);
    m.var_init = mg.var_init;
    m.comp_init_loc = mg.comp_init_loc;
    m.var_clock = mg.var_clock;
    m.comp_first_loc = mg.comp_first_loc;
    m.loc_edge0 = mg.loc_edge0;
    m.n_comp = mg.n_comp;
    m.n_vars = mg.n_vars;
    m.n_clocks = mg.n_clocks;
    m.n_leaves = mg.n_leaves;
    m.n_locs = mg.n_locs;
    m.n_edges = mg.n_edges;
    m.freeze = mg.freeze;
    m.guarded = mg.guarded;
    return m;
}

// The model with its hot tables read from the staged copy.
__device__ __forceinline__ DevModelT<SmcLdsTables> lds_model(const DevModel &mg, const SMC_GLOBAL uint8_t *base,
                                                             double *lds) {
    const SMC_LDS uint8_t *tab = (const SMC_LDS uint8_t *)(SMC_LDS double *)lds;
    DevModelT<SmcLdsTables> m;
    m.locs = in_lds(mg.locs, base, tab);
    m.invs = in_lds(mg.invs, base, tab);
    m.edges = in_lds(mg.edges, base, tab);
    m.guards = in_lds(mg.guards, base, tab);
    m.upds = in_lds(mg.upds, base, tab);
    m.consts = in_lds(mg.consts, base, tab);
    m.code = in_lds(mg.code, base, tab);
    m.var_init = mg.var_init;
    m.comp_init_loc = mg.comp_init_loc;
    m.var_clock = mg.var_clock;
    m.comp_first_loc = mg.comp_first_loc;
    m.loc_edge0 = mg.loc_edge0;
    m.n_comp = mg.n_comp;
    m.n_vars = mg.n_vars;
    m.n_clocks = mg.n_clocks;
    m.n_leaves = mg.n_leaves;
    m.n_locs = mg.n_locs;
    m.n_edges = mg.n_edges;
    m.freeze = mg.freeze;
    m.guarded = mg.guarded;
    return m;
}

// ------------------------------------------------------------ expressions

// Register-bytecode interpreter (expressions.cu:10-119).  Inlined: an
// out-of-line call costs the whole kernel registers and time even on models
// that never reach it.
template <class PC, class PI>
__device__ __forceinline__ double eval_code(uint32_t ref, const double *vars, PC consts, PI code) {
    const uint32_t first = ref & 0xFFFFFFu, count = (ref >> 24) & 0x3Fu;
    double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0;
    for (uint32_t k = 0; k < count; ++k) {
        const DevInsn in = load_rec(code + first + k);
        double o[3];
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            const uint16_t x = s == 0 ? in.a : s == 1 ? in.b : in.c;
            const uint32_t kind = x >> 14, i = x & 0x3FFFu;
            o[s] = kind == SMC_OPND_REG ? (i == 0 ? r0 : i == 1 ? r1 : i == 2 ? r2 : r3)
                 : kind == SMC_OPND_VAR ? vars[i] : kind == SMC_OPND_CLOCK ? vars[-1] - vars[i] : consts[i];
        }
        const double v = smc_apply_op(in.op, o[0], o[1], o[2]);
        r0 = in.dst == 0 ? v : r0;
        r1 = in.dst == 1 ? v : r1;
        r2 = in.dst == 2 ? v : r2;
        r3 = in.dst == 3 ? v : r3;
    }
    return r0;
}

// Value of an expression reference: constants and variables inline.
template <class M>
__device__ __forceinline__ double eval_ref(uint32_t ref, const double *vars, const M &m) {
    const uint32_t kind = ref >> 30;
    if (kind == SMC_REF_CONST) return m.consts[ref & 0x3FFFFFFFu];
    if (kind == SMC_REF_VAR) return vars[ref & 0x3FFFFFFFu];
    if (kind == SMC_REF_CLOCK) return vars[-1] - vars[ref & 0x3FFFFFFFu];
    return eval_code(ref, vars, m.consts, m.code);
}

// A location's rate within one race, remembering the last expression this
// lane evaluated: identical expressions share one reference (the compiler
// deduplicates bytecode), and variables do not change during a race, so a
// repeated reference has the same value.  Components of one template in the
// same location (every susceptible agent of an epidemic, say) then cost one
// evaluation per lane instead of one per slot.
struct RateMemo {
    uint32_t ref = 0xFFFFFFFFu;
    double value = 0.0;
};

template <class M>
__device__ __forceinline__ double location_rate(const DevLoc &L, const double *vars, const M &m, RateMemo &memo) {
    if (L.rate_const) return L.rate_value;
    if (L.rate != memo.ref) {
        memo.value = eval_ref(L.rate, vars, m);
        memo.ref = L.rate;
    }
    return memo.value;
}

// Rates of the slots in `pend` whose rate is a bytecode expression (refs
// in ref[]): each pass evaluates, on every lane at once, the expression of
// the lane's first pending slot and hands the value to all of its pending
// slots with the same expression.  Evaluating inside the per-slot loop would
// run the interpreter once per slot index at which any lane of the wave
// meets a new expression (up to C times per race); here it runs once per
// distinct expression a lane holds.
template <int C, class M>
__device__ __forceinline__ void pending_rates(uint32_t pend, const uint32_t (&ref)[C], double (&leaf)[C],
                                              const double *vars, const M &m) {
    while (__ballot(pend != 0)) {
        uint32_t r = 0xFFFFFFFFu;
        double v = 0.0;
        if (pend) {
            const int j0 = __ffs(pend) - 1;
            r = ref[0];
#pragma unroll
            for (int k = 1; k < C; ++k) r = (j0 == k) ? ref[k] : r;
            v = eval_ref(r, vars, m);
            v = v > 0.0 ? v : 0.0;
        }
#pragma unroll
        for (int j = 0; j < C; ++j)
            if (((pend >> j) & 1u) && ref[j] == r) {
                leaf[j] = v;
                pend &= ~(1u << j);
            }
    }
}

// ------------------------------------------------------------ edges

__device__ __forceinline__ bool guard_holds(int op, double x, double r) {
    switch (op) {
    case SMC_C_LE: return x <= r;
    case SMC_C_LT: return x < r;
    case SMC_C_GE: return x >= r;
    case SMC_C_GT: return x > r;
    case SMC_C_EQ: return fabs(x - r) <= SMC_TOL;
    default: return false;
    }
}

// Enabled edges of location `loc` as a bit mask over its edge list:
// spontaneous edges (recv_chan == 0) or receive edges on channel recv_chan
// (check_enabled_edges simulation.cu:268-305, SyncInSerial :363-377).
template <class M>
__device__ __forceinline__ uint64_t enabled_mask(int loc, int recv_chan, const double *vars, const M &m) {
    const DevLoc L = load_loc(m.locs + loc);
    const int ne = L.edge_n;
    uint64_t mask = 0;
    for (int i = 0; i < ne; ++i) {
        const DevEdge E = load_rec(m.edges + L.edge_first + i);
        if (E.never) continue;
        if (recv_chan ? (E.channel != -recv_chan) : (E.channel < 0)) continue;
        bool ok = true;
        for (int g = 0; g < E.guard_n && ok; ++g) {
            const DevGuard G = load_rec(m.guards + E.guard_first + g);
            const double x = (G.flags & SMC_GF_CLOCK) ? vars[-1] - vars[G.var] : vars[G.var];
            ok = guard_holds(G.op, x, eval_ref(G.rhs, vars, m));
        }
        if (ok) mask |= 1ull << i;
    }
    return mask;
}

template <class M>
__device__ __forceinline__ long long edge_weight(int e, const double *vars, const M &m) {
    const int w = smc_to_int(eval_ref(m.edges[e].weight, vars, m));
    return w > 0 ? w : 0;
}

// Weighted choice among the edges in mask (two or more) for the uniform u
// (take_transition simulation.cu:43-79 and get_sum_edge_weight :9-24, with
// the weights of the enabled edges).
template <class M>
__device__ __forceinline__ int pick_weighted(uint64_t mask, int edge_first, double u, const double *vars,
                                             const M &m) {
    long long W = 0;
    for (uint64_t r = mask; r; r &= r - 1) W += edge_weight(edge_first + lowest_lane(r), vars, m);
    const long long rnd = (long long)((double)W * u);
    long long temp = W;
    for (uint64_t r = mask; r;) {
        const int i = 63 - __clzll((long long)r);
        r &= ~(1ull << i);
        const long long w2 = edge_weight(edge_first + i, vars, m);
        if (temp - w2 <= rnd) return i;
        temp -= w2;
    }
    return lowest_lane(mask);
}

// Weighted choice among the enabled edges in mask; the draw is made only
// when there is a choice.
template <class M>
__device__ __forceinline__ int pick_edge(uint64_t mask, int edge_first, uint64_t seed, uint64_t run,
                                         uint32_t step, uint32_t purpose, uint32_t comp,
                                         const double *vars, const M &m) {
    if ((mask & (mask - 1)) == 0) return lowest_lane(mask);
    const smc_u32x4 w = smc_draw(seed, run, step, purpose, comp);
    return pick_weighted(mask, edge_first, smc_u01(w.v[0], w.v[1]), vars, m);
}

// The same for component c as a broadcast receiver: its uniform is word
// c & 3 of the receivers' shared draw c >> 2.
template <class M>
__device__ __forceinline__ int pick_receiver(uint64_t mask, int edge_first, uint64_t seed, uint64_t run,
                                             uint32_t step, uint32_t c, const double *vars, const M &m) {
    if ((mask & (mask - 1)) == 0) return lowest_lane(mask);
    const smc_u32x4 w = smc_draw(seed, run, step, SMC_P_RECEIVER, smc_receiver_draw(c));
    return pick_weighted(mask, edge_first, smc_receiver_u(w, c), vars, m);
}

// Rate sum, on the race tree, of the components that cannot fire before
// the next transition (frozen: blocked, exponential now) in the slots of
// deadmask.  They leave the race when every racer is exponential; the
// transition that crosses the time bound still competes with their first
// null event after the bound, which is drawn from this sum (memoryless
// races).
template <int C, class M>
__device__ __forceinline__ double dead_total(uint32_t deadmask, const int (&loc)[C], const double *vars,
                                             const M &m, int S) {
    double dl[C];
    RateMemo memo;
#pragma unroll
    for (int j = 0; j < C; ++j) {
        dl[j] = 0.0;
        if ((deadmask >> j) & 1u) {
            const DevLoc L = load_loc(m.locs + loc[j]);
            const double r = location_rate(L, vars, m, memo);
            dl[j] = r > 0.0 ? r : 0.0;
        }
    }
    return tree_total<C>(dl, S);
}

struct LaneQuery {
    int qvar;
    double qmax, qmin;
};

// Fire edge `e` of location `src`: updates in order, query variable
// tracking, new location (take_transition simulation.cu:88-180).  Returns
// the destination (edge records hold it relative to the source).
template <class M>
__device__ __forceinline__ int fire_edge(int e, int src, double *vars, const M &m, LaneQuery &q,
                                         bool &goal, int &chan) {
    const DevEdge E = load_rec(m.edges + e);
    for (int k = 0; k < E.upd_n; ++k) {
        const DevUpd U = load_rec(m.upds + E.upd_first + k);
        const double v = eval_ref(U.expr, vars, m);
        if ((int)U.var == q.qvar) {
            if (v > q.qmax) q.qmax = v;
            if (v < q.qmin) q.qmin = v;
        }
        vars[U.var] = U.clock ? vars[-1] - v : v;
    }
    if (E.dst_goal) goal = true;
    if (E.channel > 0) chan = E.channel;
    return src + E.dst;
}

// Slots (bit j) whose component is blocked by the variables: in a
// location flagged SMC_LF_GUARDED (not zero-delay), every spontaneous edge
// has a `!=` conjunct or a false conjunct that reads no clock.  A blocked
// component cannot fire until a transition changes the variables; the
// kernels recompute the mask after every change of state, on models with
// such locations (DevModel.guarded).  Locations flagged SMC_LF_DEAD are
// always blocked and are tested by their flag.
// The blocking test of a location flagged SMC_LF_GUARDED: every edge has a
// false clock-free conjunct (one pass over its blocking records).
template <class M>
__device__ __forceinline__ bool location_blocked(const DevLoc &L, const double *vars, const M &m) {
    if (L.flags & SMC_LF_INL_BLK) return !guard_holds(L.inl_op, vars[L.inl_var], L.inl_value);
    bool e_blk = false;
    for (uint32_t k = L.block_first;; ++k) {
        const DevGuard G = load_rec(m.guards + k);
        if (!e_blk) e_blk = !guard_holds(G.op, vars[G.var], eval_ref(G.rhs, vars, m));
        if (G.flags & SMC_GF_END_EDGE) {
            if (!e_blk) return false;
            e_blk = false;
        }
        if (G.flags & SMC_GF_END_LOC) return true;
    }
}

template <int C, class M>
__device__ __forceinline__ uint32_t blocked_slots(const int (&loc)[C], const double *vars, const M &m, int sl,
                                                  int n) {
    uint32_t b = 0;
#pragma unroll
    for (int j = 0; j < C; ++j) {
        const int c = sl * C + j;
        if (c >= n) continue;
        const DevLoc L = load_loc(m.locs + loc[j]);
        if ((L.flags & SMC_LF_GUARDED) && L.klass != SMC_K_ZERO && location_blocked(L, vars, m)) b |= 1u << j;
    }
    return b;
}

// Whether no component of the segment's run can fire again before the time
// bound: each one is blocked (bit of `blk`), or is in a location whose rate
// and spontaneous guards read no clock and has no enabled spontaneous edge
// or races at rate 0 (an exponential location).  Zero-delay and
// clock-bounded locations with an enabled edge can fire.  Nothing but time
// changes such a state, so the run's outcome is decided; the kernels check
// it after every change of state.
template <int C, class M>
__device__ __forceinline__ bool absorbed(const int (&loc)[C], const double *vars, const M &m, int sl, int n,
                                         uint64_t seg_mask, uint32_t blk) {
    bool can = false;
    RateMemo memo;
#pragma unroll
    for (int j = 0; j < C; ++j) {
        const int c = sl * C + j;
        if (can || c >= n) continue;
        const DevLoc L = load_loc(m.locs + loc[j]);
        if (L.klass == SMC_K_ZERO) { can = true; continue; }
        if ((L.flags & SMC_LF_DEAD) || ((blk >> j) & 1u)) continue;
        if (!(L.flags & SMC_LF_TIME_FREE)) { can = true; continue; }
        if (enabled_mask(loc[j], 0, vars, m) == 0) continue;
        if (L.klass != SMC_K_EXP) { can = true; continue; }
        const double r = location_rate(L, vars, m, memo);
        if (r > 0.0) can = true;
    }
    return (__ballot(can) & seg_mask) == 0;
}

template <int C, class T>
__device__ __forceinline__ T slot_get(const T (&a)[C], int j);
template <int C, class T>
__device__ __forceinline__ void slot_set(T (&a)[C], int j, T v);

// Receivers of one broadcast fired in parallel (SyncInSerial
// simulation.cu:378-404 fires them one by one in component order).  Valid
// when every receive edge of the segment's receivers commutes (SMC_LF_RECV_FAST):
// constant weights make each receiver's pick independent of the others'
// updates, and integer additions and private constants give the same
// variables in any order.  The query variable's running max/min needs the
// values in component order: they are v0 plus the prefix sums of the
// additions, an exclusive scan over the segment (blocked layout: lane order
// then slot order is component order).  TRACE mixes the fired edges into
// the digest in component order.  Returns false, having changed nothing,
// when some enabled edge does not commute.
template <int C, class M>
__device__ __forceinline__ bool fire_receivers_fast(const uint64_t (&mk)[C], int (&loc)[C], double *vars,
                                                    const M &m, LaneQuery &q, bool &goal, uint64_t seed,
                                                    uint64_t run, uint32_t step, int sl, int S, int seg_base,
                                                    uint64_t seg_mask, uint32_t &trans, bool trace,
                                                    uint64_t &digest) {
    bool slow = false;
#pragma unroll
    for (int j = 0; j < C; ++j)
        if (mk[j] && !(m.locs[loc[j]].flags & SMC_LF_RECV_FAST)) slow = true;
    if (__ballot(slow) & seg_mask) return false;

    int ej[C], eg[C];  // fired edge per slot: record index, and the network's index (digests)
    int fired = 0;
    bool any_fired = false;
    double qsum = 0.0, qhi = -DBL_MAX, qlo = DBL_MAX;  // prefix sums of this lane's query additions
    bool qany = false;
    // Receivers 4g .. 4g+3 share draw g: the lane's slots of one draw are
    // G consecutive slots (C is a power of two), drawn once at the first of
    // them when any of them has a choice to make.
    constexpr int G = C < 4 ? C : 4;
    smc_u32x4 rw = {{0u, 0u, 0u, 0u}};
#pragma unroll
    for (int j = 0; j < C; ++j) {
        ej[j] = -1;
        eg[j] = -1;
        const uint32_t c = (uint32_t)(sl * C + j);
        if (j % G == 0) {
            bool choice = false;
#pragma unroll
            for (int k = j; k < j + G; ++k) choice |= (mk[k] & (mk[k] - 1)) != 0;
            if (choice) rw = smc_draw(seed, run, step, SMC_P_RECEIVER, smc_receiver_draw(c));
        }
        if (mk[j]) {
            const int first = (int)m.locs[loc[j]].edge_first;
            const int i = (mk[j] & (mk[j] - 1)) ? pick_weighted(mk[j], first, smc_receiver_u(rw, c), vars, m)
                                                : lowest_lane(mk[j]);
            const int e = first + i;
            ej[j] = e;
            if (trace) eg[j] = m.loc_edge0[loc[j]] + i;
            ++fired;
            any_fired = true;
            const DevEdge E = load_rec(m.edges + e);
            for (int k = 0; k < E.upd_n; ++k) {
                const DevUpd U = load_rec(m.upds + E.upd_first + k);
                if ((int)U.var != q.qvar) continue;
                if (U.kind == SMC_U_ADD) {
                    qsum += U.value;
                    qhi = fmax(qhi, qsum);
                    qlo = fmin(qlo, qsum);
                    qany = true;
                } else {
                    q.qmax = fmax(q.qmax, U.value);
                    q.qmin = fmin(q.qmin, U.value);
                }
            }
        }
    }
    if (q.qvar >= 0 && (__ballot(qany) & seg_mask)) {
        double incl = qsum;
        for (int k = 1; k < S; k <<= 1) {
            const double y = __shfl_up(incl, k, S);
            if (sl >= k) incl += y;
        }
        const double v0 = vars[q.qvar] + (incl - qsum);  // value before this lane's additions
        if (qany) {
            q.qmax = fmax(q.qmax, v0 + qhi);
            q.qmin = fmin(q.qmin, v0 + qlo);
        }
    }
    wave_sync();  // every lane has read the variables before they change
#pragma unroll
    for (int j = 0; j < C; ++j) {
        if (ej[j] >= 0) {
            const DevEdge &E = m.edges[ej[j]];
            for (int k = 0; k < E.upd_n; ++k) {
                const DevUpd U = load_rec(m.upds + E.upd_first + k);
                if (U.kind == SMC_U_ADD) atomicAdd(&vars[U.var], U.value);
                else vars[U.var] = U.clock ? vars[-1] - U.value : U.value;
            }
            if (E.dst_goal) goal = true;
            loc[j] = loc[j] + E.dst;
        }
    }
    wave_sync();
    // the segment's total, bit by bit (fired <= C <= 16): independent ballots
    // instead of a chain of shuffles
    uint32_t total = 0;
#pragma unroll
    for (int b = 0; b < 5; ++b) total += (uint32_t)__popcll(__ballot((fired >> b) & 1) & seg_mask) << b;
    trans += total;
    if (trace) {
        for (uint64_t lanes_with = __ballot(any_fired) & seg_mask; lanes_with; lanes_with &= lanes_with - 1) {
            const int L = lowest_lane(lanes_with);
#pragma unroll
            for (int j = 0; j < C; ++j) {
                const int e = __shfl(eg[j], L);
                if (e >= 0)
                    digest = smc_mix(digest, ((uint64_t)step << 40) | ((uint64_t)((L - seg_base) * C + j) << 20) |
                                                 (uint64_t)e);
            }
        }
    }
    return true;
}

// Rotate slot arrays by one (slot 1 to the front): a rolled loop over the
// slots then touches slot 0 only, a constant index, so the arrays stay in
// registers rather than being spilled for dynamic indexing.
template <int C, class T>
__device__ __forceinline__ void rotate_slots(T (&a)[C]) {
    const T f = a[0];
#pragma unroll
    for (int k = 0; k < C - 1; ++k) a[k] = a[k + 1];
    a[C - 1] = f;
}

// Receivers of one broadcast fired one at a time in component order
// (SyncInSerial simulation.cu:378-404): the lanes holding enabled receivers
// in lane order, their slots in slot order.  Enabled sets mk were computed
// before any of them fires; each picks with its own stream.
template <int C, class M>
__device__ __forceinline__ void fire_receivers_serial(uint64_t (&mk)[C], int (&loc)[C], bool any, double *vars,
                                                      const M &m, LaneQuery &q, bool &goal, uint64_t seed,
                                                      uint64_t run, uint32_t step, int lane, int seg_base,
                                                      uint64_t seg_mask, uint32_t &trans, bool trace,
                                                      uint64_t &digest) {
    int dummy_chan = 0;
    for (uint64_t lanes_with = __ballot(any) & seg_mask; lanes_with; lanes_with &= lanes_with - 1) {
        const int L = lowest_lane(lanes_with);
        // slot j is at the front of the rotated slot arrays
#pragma clang loop unroll(disable)
        for (int j = 0; j < C; ++j) {
            if ((__ballot(mk[0] != 0) >> L) & 1ull) {
                const int c = (L - seg_base) * C + j;
                int e_fired = -1;  // the network's index of the fired edge (digests)
                if (lane == L) {
                    const int first = (int)m.locs[loc[0]].edge_first;
                    const int i = pick_receiver(mk[0], first, seed, run, step, (uint32_t)c, vars, m);
                    if (trace) e_fired = m.loc_edge0[loc[0]] + i;
                    loc[0] = fire_edge(first + i, loc[0], vars, m, q, goal, dummy_chan);
                }
                wave_sync();
                ++trans;
                if (trace) {
                    const int e = __shfl(e_fired, L);
                    digest = smc_mix(digest, ((uint64_t)step << 40) | ((uint64_t)c << 20) | (uint64_t)e);
                }
            }
            rotate_slots<C>(mk);
            rotate_slots<C>(loc);
        }
    }
}

// Register-resident slot arrays read at a run-time slot index.  Each element
// passes through an empty asm so the select chain stays a chain of register
// selects: without it the optimiser folds select(load a[k], load a[j]) into
// a load through a selected address, which pins the array in scratch.
template <class T>
__device__ __forceinline__ T in_register(T v) {
    asm("" : "+v"(v));
    return v;
}
template <int C, class T>
__device__ __forceinline__ T slot_get(const T (&a)[C], int j) {
    T r = in_register(a[0]);
#pragma unroll
    for (int k = 1; k < C; ++k) r = (j == k) ? in_register(a[k]) : r;
    return r;
}
template <int C, class T>
__device__ __forceinline__ void slot_set(T (&a)[C], int j, T v) {
#pragma unroll
    for (int k = 0; k < C; ++k) a[k] = (j == k) ? v : a[k];
}

}  // namespace
