// oracle.cpp — CPU restatement of cuSMC's simulation semantics.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).  Written independently of the
// GPU kernel and of the product's bytecode: expressions are evaluated by
// walking the trees of the smc_network, each run is simulated sequentially.
//
// The step, following cuSMC (file:line in the reference):
//   loop while t < bound                              simulation.cu:835
//     location query and goal flag set -> hit, stop    simulation.cu:839-849
//     pending broadcast: every component except the sender collects its
//       enabled receive edges on the current state, then they fire one
//       after the other in component order              simulation.cu:357-407
//     delays: branch/urgent/committed 0                 simulation.cu:438-446
//       clock upper bounds x<=U (x<U: U-1e-6) with U-x>=0 bound the delay,
//       uniform on (0, min(U-x)]                        simulation.cu:460-521
//       otherwise exponential, delay -log2(u)/rate      simulation.cu:522-540
//     minimum delay, lowest component index on ties     simulation.cu:593-625
//     clocks advance by it                              simulation.cu:632-651
//       (stored as reset stamps: a clock's value is the clock time minus
//       its stamp, the representation of the kernels; assigning x := v
//       stores stamp = time - v)
//     the winner fires one enabled non-receive edge     simulation.cu:654-667
//       weighted by integer-truncated weights           simulation.cu:9-79
//       updates in order, tracking the query variable   simulation.cu:101-134
//       a goal destination sets the goal flag           simulation.cu:147-168
//     t += delay                                        simulation.cu:876-882
//   a run is cut off after max_steps race rounds (no counterpart: cuSMC
//     spins forever on a cycle of zero delays)
//   CANONICAL only: a run ends when no component can fire again before the
//     bound (absorption: every component is blocked — each spontaneous edge
//     has a false conjunct that reads no clock — or has no enabled
//     spontaneous edge or races at rate 0, in a location whose rate and
//     guards read no clock); cuSMC races null steps to the bound there, with
//     the same outcome
//   guards: only conjuncts on a variable; `==` within 1e-5; `!=` disables
//     the edge                                          simulation.cu:184-266
//
// Two ways to sample the race (oracle_mode):
//   LITERAL   — as cuSMC: every component draws its own delay every step.
//   CANONICAL — the GPU kernel's sampling of the same law: zero delays win
//     first (lowest index); the exponential components form one race, total
//     rate on a fixed pairwise tree, delay -log2(u1)/total, winner by
//     descending the tree with u2*total; bounded components draw uniforms.
//     Components that can only make null steps until the next transition
//     (frozen: exponential, every spontaneous edge blocked by a clock-free
//     conjunct) leave the race; the step that crosses the bound draws their
//     first event after the bound.
//     Same Philox stream, same operations, same bits as the GPU.
// Resolutions of undefined behaviour (by intent): all variables and all
// clocks are tracked (cuSMC: 3), a receiver picks with its own stream,
// weights are those of the enabled edges, negative or NaN rates and weights
// count as 0, `%` by zero gives 0.  FP64 time and clocks (cuSMC: FP32).

#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "smc_stream.h"

namespace {

thread_local std::string g_err;
const double TOL = 1e-5;  // cuSMC's DBL_EPSILON (include/common_macros.h:72)

int to_int_sat(double x) {
    if (std::isnan(x)) return 0;
    if (x >= 2147483647.0) return 2147483647;
    if (x <= -2147483648.0) return INT32_MIN;
    return (int)x;
}

// Algorithmic cost model (lane-operations of one run's work, counted by
// the canonical simulation when enabled): the floor a SIMT implementation
// must issue, independent of how it maps runs to lanes.  Operand fetches are
// free; the constants below are the operation counts of the published
// methods the simulator uses.
thread_local double g_cost = 0.0;
thread_local bool g_count = false;
const double COST_OP = 1.0;       // add, mul, compare, select, logic
const double COST_DIV = 10.0;     // FP64 reciprocal + 2 Newton steps + correction
const double COST_SQRT = 10.0;
const double COST_POW = 60.0;     // log2, multiply, exp2
const double COST_PHILOX = 100.0; // 10 rounds x (2 mul_lo + 2 mul_hi + 4 xor + 2 key adds)
const double COST_U01 = 4.0;      // 53-bit combine, convert, scale
const double COST_LOG2 = 30.0;    // frexp + atanh series
const double COST_EXP2 = 20.0;
inline void cost(double c) {
    if (g_count) g_cost += c;
}

// Tree evaluation with the operator meanings of expressions.cu:10-95.
double eval(const smc_network &n, int e, const double *vars) {
    const smc_expr_node &x = n.exprs[e];
    if (g_count && x.op != SMC_E_LIT && x.op != SMC_E_VAR)
        g_cost += x.op == SMC_E_DIV ? COST_DIV : x.op == SMC_E_SQRT ? COST_SQRT : x.op == SMC_E_POW ? COST_POW : COST_OP;
    switch (x.op) {
    case SMC_E_LIT: return x.value;
    case SMC_E_VAR: return vars[x.var];
    case SMC_E_RANDOM: return 1.0;
    case SMC_E_NEG: return -eval(n, x.a, vars);
    case SMC_E_SQRT: return std::sqrt(eval(n, x.a, vars));
    case SMC_E_NOT: return std::fabs(eval(n, x.a, vars)) < TOL ? 1.0 : 0.0;
    case SMC_E_COND: {
        double c = eval(n, x.a, vars);
        return std::fabs(c) > TOL ? eval(n, x.b, vars) : eval(n, x.c, vars);
    }
    default: break;
    }
    const double a = eval(n, x.a, vars), b = eval(n, x.b, vars);
    switch (x.op) {
    case SMC_E_ADD: return a + b;
    case SMC_E_SUB: return a - b;
    case SMC_E_MUL: return a * b;
    case SMC_E_DIV: return a / b;
    case SMC_E_POW: return smc_pow(a, b);
    case SMC_E_MOD: {
        int ia = to_int_sat(a), ib = to_int_sat(b + TOL);
        if (ib == 0 || (ia == INT32_MIN && ib == -1)) return 0.0;
        return (double)(ia % ib);
    }
    case SMC_E_AND: return (std::fabs(a) > TOL && std::fabs(b) > TOL) ? 1.0 : 0.0;
    case SMC_E_OR: return (std::fabs(a) > TOL || std::fabs(b) > TOL) ? 1.0 : 0.0;
    case SMC_E_LE: return a <= b ? 1.0 : 0.0;
    case SMC_E_GE: return a >= b ? 1.0 : 0.0;
    case SMC_E_LT: return a < b ? 1.0 : 0.0;
    case SMC_E_GT: return a > b ? 1.0 : 0.0;
    case SMC_E_EQ: return std::fabs(a - b) <= TOL ? 1.0 : 0.0;
    case SMC_E_NE: return std::fabs(a - b) > TOL ? 1.0 : 0.0;
    default: return 0.0;
    }
}

// Canonical pairwise tree: heap t[1..2P), leaves t[P + c].
struct RaceTree {
    std::vector<double> t;
    int P = 1;
    void build(const std::vector<double> &leaf, int n) {
        P = 1;
        while (P < n) P <<= 1;
        t.assign(2 * P, 0.0);
        for (int c = 0; c < n; ++c) t[P + c] = leaf[c];
        for (int i = P - 1; i >= 1; --i) t[i] = t[2 * i] + t[2 * i + 1];
    }
    double total() const { return P == 1 ? t[1] : t[1]; }
    int descend(double v) const {
        int i = 1;
        while (i < P) {
            const double left = t[2 * i], right = t[2 * i + 1];
            if (v < left || right == 0.0) i = 2 * i;
            else { v = v - left; i = 2 * i + 1; }
        }
        return i - P;
    }
};

struct RunOut {
    bool flag;
    double value;
    double log2lr;
    uint32_t steps, transitions;
    uint64_t digest;
    double time;
};

// Importance-sampling options of a run: per-location log2 tilts of the
// exponential rates, early stop of decided var-probability runs, and the
// recording of macro-steps (features at the start of each, wins and
// ln2 * dt * base rate per location) used by cross-entropy training.
struct IsOpts {
    const float *theta = nullptr;
    bool early_stop = false;
    const float *var_scale = nullptr;
    int fpad = 0, lpad = 0;
    std::vector<uint16_t> *X = nullptr;
    std::vector<float> *A = nullptr, *B = nullptr;
};

uint16_t bf16_rne(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)((u >> 16) | ((u & 0xFFFF) ? 0x40 : 0));
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

class Simulator {
public:
    Simulator(const smc_network &n, const smc_query &q, uint64_t seed, int mode)
        : n_(n), q_(q), seed_(seed), mode_(mode),
          max_steps_(q.max_steps ? q.max_steps : SMC_DEFAULT_MAX_STEPS) {
        leaf_.resize(n.n_components);
        um_.resize(n.n_components);
        loc_.resize(n.n_components);
        vars_.resize(n.n_vars > 0 ? n.n_vars : 1);
        stamp_.assign(n.n_vars > 0 ? n.n_vars : 1, 0.0);
        recv_.resize(n.n_components);
        for (int v = 0; v < n.n_vars; ++v)
            if (n.var_kind[v] == SMC_VAR_CLOCK) clocks_.push_back(v);
        // per location: rate and spontaneous guards free of clocks; clock
        // upper bounds present (the conditions of absorption below)
        time_free_.assign(n.n_locations, 1);
        clock_bounded_.assign(n.n_locations, 0);
        blocked_.assign(n.n_components, 0);
        null_.assign(n.n_components, 0);
        blk_kind_.assign(n.n_locations, 0);
        for (int l = 0; l < n.n_locations; ++l) {
            const smc_location &L = n.locations[l];
            if (L.rate >= 0 && reads_clock(L.rate)) time_free_[l] = 0;
            for (int k = 0; k < L.edge_count; ++k) {
                const smc_edge &E = n.edges[L.edge_first + k];
                if (E.channel < 0) continue;
                for (int g = 0; g < E.guard_count; ++g) {
                    const smc_constraint &G = n.constraints[E.guard_first + g];
                    if (G.uses_variable && (n.var_kind[G.var] == SMC_VAR_CLOCK || reads_clock(G.rhs)))
                        time_free_[l] = 0;
                }
            }
            for (int k = 0; k < L.inv_count; ++k) {
                const smc_constraint &I = n.constraints[L.inv_first + k];
                if (I.uses_variable && n.var_kind[I.var] == SMC_VAR_CLOCK && (I.op == SMC_C_LE || I.op == SMC_C_LT))
                    clock_bounded_[l] = 1;
            }
            // blocking: 1 always (every spontaneous edge has a `!=` conjunct,
            // or there is none), 2 depends on the variables (a spontaneous
            // edge without `!=` has a conjunct that reads no clock), 0 never
            int kind = 1;
            for (int k = 0; k < L.edge_count && kind != 0; ++k) {
                const smc_edge &E = n.edges[L.edge_first + k];
                if (E.channel < 0) continue;
                bool never = false, clock_free = false;
                for (int g = 0; g < E.guard_count; ++g) {
                    const smc_constraint &G = n.constraints[E.guard_first + g];
                    if (!G.uses_variable) continue;
                    if (G.op == SMC_C_NE) never = true;
                    else if (n.var_kind[G.var] != SMC_VAR_CLOCK && !reads_clock(G.rhs)) clock_free = true;
                }
                if (never) continue;
                kind = clock_free ? 2 : 0;
            }
            blk_kind_[l] = (uint8_t)kind;
        }
        // Leaving null racers out of an all-exponential race keeps the law
        // only while every rate is constant between transitions: a rate that
        // reads a clock is re-evaluated at each of cuSMC's null steps
        // (simulation.cu:522-530), so then every component races.
        for (int l = 0; l < n.n_locations; ++l)
            if (n.locations[l].rate >= 0 && reads_clock(n.locations[l].rate)) freeze_ = false;
    }

    // events: optional log of (step, component, edge) per fired edge
    RunOut run(uint64_t run, std::vector<int32_t> *events = nullptr, std::vector<double> *times = nullptr,
               const IsOpts *is = nullptr) {
        is_ = is;
        hazards_ = nullptr;
        log2lr_ = 0.0;
        const bool record = is && is->X;
        bool new_macro = true, open = false;
        std::vector<int> loc0;
        std::vector<double> hz;
        std::vector<uint32_t> wins;
        const int N = n_.n_components;
        // clocks are kept as reset stamps (value = clock time - stamp), the
        // representation the kernels use; vars_ holds every value
        tclk_ = 0.0;
        for (int v = 0; v < n_.n_vars; ++v) {
            vars_[v] = n_.var_init[v];
            if (n_.var_kind[v] == SMC_VAR_CLOCK) {
                stamp_[v] = -n_.var_init[v];
                vars_[v] = tclk_ - stamp_[v];
            }
        }
        for (int c = 0; c < N; ++c) loc_[c] = n_.comp_init_loc[c];
        const bool reach = q_.kind == SMC_QUERY_REACH;
        const int qvar = reach ? -1 : q_.var;
        const double qinit = (qvar >= 0 && n_.var_kind[qvar] == SMC_VAR_INT) ? n_.var_init[qvar] : 0.0;
        qmax_ = qmin_ = qinit;
        qvar_ = qvar;
        double t = 0.0;
        uint32_t step = 0, trans = 0;
        int chan = 0, sender = -1;
        bool goal = false, hit = false;
        bool changed = true;  // state changed since the last absorption check
        uint64_t digest = smc_mix(0, run);
        auto log = [&](int c, int e) {
            ++trans;
            digest = smc_mix(digest, ((uint64_t)step << 40) | ((uint64_t)c << 20) | (uint64_t)e);
            if (events) { events->push_back((int32_t)step); events->push_back(c); events->push_back(e); events->push_back(0); }
            if (times) times->push_back(t);
        };

        auto close_macro = [&]() {
            if (!open) return;
            const size_t base = is->A->size();
            is->A->resize(base + is->lpad, 0.f);
            is->B->resize(base + is->lpad, 0.f);
            for (int c = 0; c < N; ++c) {
                const int l = loc0[c];
                if (n_.locations[l].type > SMC_LOC_GOAL) continue;  // not tiltable
                (*is->A)[base + l] = (float)wins[c];
                (*is->B)[base + l] = (float)hz[c];
            }
            open = false;
        };

        for (;;) {
            if (!(t < q_.time_bound)) break;
            if (reach && goal) { hit = true; break; }
            if (is && is->early_stop && !reach && (q_.is_max ? qmax_ > q_.threshold : qmin_ < q_.threshold)) {
                hit = true;
                break;
            }
            if (step >= max_steps_) break;

            if (chan > 0) {
                for (int c = 0; c < N; ++c) {
                    recv_[c].clear();
                    if (c != sender) enabled(loc_[c], chan, &recv_[c]);
                }
                int group = -1;  // receivers 4g .. 4g+3 share one draw
                for (int c = 0; c < N; ++c) {
                    if (recv_[c].empty()) continue;
                    const uint32_t g = smc_receiver_draw((uint32_t)c);
                    if (recv_[c].size() > 1 && (int)g != group) {
                        cost(COST_PHILOX);
                        group = (int)g;
                    }
                    smc_u32x4 w = smc_draw(seed_, run, step, SMC_P_RECEIVER, g);
                    int e = pick(recv_[c], smc_receiver_u(w, (uint32_t)c));
                    int dummy = 0;
                    fire(c, e, &goal, &dummy);
                    log(c, e);
                }
                chan = 0;
                sender = -1;
                changed = true;
            }

            if (record && new_macro) {
                close_macro();
                // features: one-hot locations, scaled variables, t / bound
                const size_t base = is->X->size();
                is->X->resize(base + is->fpad, 0);
                uint16_t *f = is->X->data() + base;
                for (int c = 0; c < N; ++c) f[loc_[c]] = 0x3F80;
                const int nl = n_.n_locations;
                for (int v = 0; v < n_.n_vars; ++v)
                    f[nl + v] = bf16_rne((float)(vars_[v] * (double)is->var_scale[v]));
                f[nl + n_.n_vars] = bf16_rne((float)(t / q_.time_bound));
                loc0.assign(loc_.begin(), loc_.end());
                hz.assign(N, 0.0);
                wins.assign(N, 0);
                open = true;
                hazards_ = &hz;
            }
            new_macro = false;
            // A goal entered by a broadcast receiver (this iteration's
            // receivers; a goal entered by a winner has ended the run at the
            // loop top) counts only if this iteration's step, raced by every
            // component, keeps t < bound (simulation.cu:409-415 fire the
            // receivers inside the iteration, :835-849 test the goal at the
            // next loop top).  That step is raced in full: no absorption
            // exit and no left-out null racers.
            const bool goal_pending = reach && goal;
            // absorption (canonical sampling; cuSMC keeps racing null
            // steps to the bound, with the same outcome).  With frozen
            // racers (freeze_) the race below finds such a state by itself:
            // blocked components leave it, the rest race at rate 0, so it
            // has no event and ends the run.
            if (mode_ == ORACLE_CANONICAL && changed) {
                // blocked components: unchanged until the next transition
                for (int c = 0; c < N; ++c)
                    blocked_[c] = n_.locations[loc_[c]].type <= SMC_LOC_GOAL && blocked(c);
                if (!goal_pending) {
                    changed = false;
                    if (!freeze_ && absorbed()) break;
                }
            }

            double d;
            int w;
            bool won = false;
            if (mode_ == ORACLE_LITERAL) race_literal(run, step, &d, &w);
            else race_canonical(run, step, t, goal_pending, &d, &w, &won);
            if (record && won) ++wins[w];
            if (!(d < HUGE_VAL)) break;  // no event: the run ends at its clock

            // clocks advance: the clock time moves, every clock's value is
            // the clock time minus its stamp
            tclk_ = t + d;
            for (int v : clocks_) vars_[v] = tclk_ - stamp_[v];
            cost(2 * COST_OP);  // clock time, time
            std::vector<int> en;
            if (w >= 0) enabled(loc_[w], 0, &en);
            if (!en.empty()) {
                smc_u32x4 r = smc_draw(seed_, run, step, SMC_P_WINNER, (uint32_t)w);
                if (en.size() > 1) cost(COST_PHILOX);
                int e = pick(en, smc_u01(r.v[0], r.v[1]));
                int nc = 0;
                fire(w, e, &goal, &nc);
                log(w, e);
                if (nc > 0) { chan = nc; sender = w; }
                new_macro = true;
                changed = true;
            }
            t = t + d;
            ++step;
        }
        if (record) close_macro();

        RunOut o;
        o.value = q_.is_max ? qmax_ : qmin_;
        o.flag = reach ? hit : (q_.is_max ? o.value > q_.threshold : o.value < q_.threshold);
        o.steps = step;
        o.transitions = trans;
        o.time = t;
        digest = smc_mix(digest, smc_digest_bits(t));
        for (int v = 0; v < n_.n_vars; ++v) digest = smc_mix(digest, smc_digest_bits(vars_[v]));
        for (int c = 0; c < N; ++c) digest = smc_mix(digest, (uint64_t)(uint32_t)loc_[c]);
        o.digest = digest;
        o.log2lr = log2lr_;
        return o;
    }

private:
    const smc_network &n_;
    smc_query q_;
    uint64_t seed_;
    int mode_;
    uint32_t max_steps_;
    std::vector<double> leaf_, vars_;
    std::vector<double> stamp_;  // clocks: reset stamps
    double tclk_ = 0.0;          // clock time: the time of the last advance
    std::vector<int> loc_, clocks_;
    std::vector<std::vector<int>> recv_;
    RaceTree tree_, base_tree_, dead_tree_;
    std::vector<double> dleaf_;
    std::vector<double> um_;   // per component: its uniform window M this race, 0 if not a uniform racer
    std::vector<double> tilted_;
    const IsOpts *is_ = nullptr;
    std::vector<double> *hazards_ = nullptr;  // recording: base cumulative hazard of each component
    double log2lr_ = 0.0;
    int qvar_ = -1;
    double qmax_ = 0, qmin_ = 0;

    std::vector<uint8_t> time_free_, clock_bounded_;
    std::vector<uint8_t> blocked_;  // per component, at the last change of state
    std::vector<uint8_t> blk_kind_; // per location: blocked never (0), always (1), by the variables (2)
    std::vector<uint8_t> null_;     // per component: left out of this race
    bool freeze_ = true;            // no rate reads a clock

    bool reads_clock(int x) const {
        if (x < 0) return false;
        const smc_expr_node &nd = n_.exprs[x];
        if (nd.op == SMC_E_VAR) return n_.var_kind[nd.var] == SMC_VAR_CLOCK;
        if (nd.op == SMC_E_LIT) return false;
        return reads_clock(nd.a) || reads_clock(nd.b) || reads_clock(nd.c);
    }

    // Every spontaneous edge of component c has a false conjunct that reads
    // no clock (or a `!=` conjunct, never enabled): c cannot fire until a
    // transition changes the variables.
    bool blocked(int c) const {
        const int kind = blk_kind_[loc_[c]];
        if (kind != 2) return kind == 1;
        const smc_location &L = n_.locations[loc_[c]];
        for (int k = 0; k < L.edge_count; ++k) {
            const smc_edge &E = n_.edges[L.edge_first + k];
            cost(2 * COST_OP);
            if (E.channel < 0) continue;
            bool b = false;
            for (int g = 0; g < E.guard_count && !b; ++g) {
                const smc_constraint &G = n_.constraints[E.guard_first + g];
                if (!G.uses_variable) continue;
                cost(2 * COST_OP);
                if (G.op == SMC_C_NE) { b = true; break; }
                if (n_.var_kind[G.var] == SMC_VAR_CLOCK || reads_clock(G.rhs)) continue;
                const double x = vars_[G.var];
                const double r = eval(n_, G.rhs, vars_.data());
                bool ok;
                switch (G.op) {
                case SMC_C_LT: ok = x < r; break;
                case SMC_C_LE: ok = x <= r; break;
                case SMC_C_GT: ok = x > r; break;
                case SMC_C_GE: ok = x >= r; break;
                default: ok = std::fabs(x - r) <= TOL; break;  // SMC_C_EQ
                }
                b = !ok;
            }
            if (!b) return false;
        }
        return true;
    }

    // No component can fire before the time bound: each is blocked, or sits
    // in a location whose rate and spontaneous guards read no clock, with no
    // enabled spontaneous edge or an exponential rate of 0.  Only time
    // changes such a state, so the outcome is decided.
    bool absorbed() const {
        for (int c = 0; c < n_.n_components; ++c) {
            const int l = loc_[c];
            if (n_.locations[l].type > SMC_LOC_GOAL) return false;
            if (blocked_[c]) continue;
            if (!time_free_[l]) return false;
            std::vector<int> en;
            enabled(l, 0, &en);
            if (en.empty()) continue;
            if (clock_bounded_[l]) return false;
            if (rate_of(c) > 0.0) return false;
        }
        return true;
    }

    bool guards_hold(const smc_edge &E) const {
        for (int g = 0; g < E.guard_count; ++g) {
            const smc_constraint &G = n_.constraints[E.guard_first + g];
            if (!G.uses_variable) continue;                  // simulation.cu:209
            const double x = vars_[G.var];
            const double r = eval(n_, G.rhs, vars_.data());
            bool ok;
            switch (G.op) {
            case SMC_C_LT: ok = x < r; break;
            case SMC_C_LE: ok = x <= r; break;
            case SMC_C_GT: ok = x > r; break;
            case SMC_C_GE: ok = x >= r; break;
            case SMC_C_EQ: ok = std::fabs(x - r) <= TOL; break;
            default: ok = false; break;                      // `!=`: simulation.cu:249-251
            }
            if (!ok) return false;
        }
        return true;
    }

    // Enabled edges of location l (global edge indices, file order).
    void enabled(int l, int recv_chan, std::vector<int> *out) const {
        const smc_location &L = n_.locations[l];
        for (int k = 0; k < L.edge_count; ++k) {
            const int e = L.edge_first + k;
            const smc_edge &E = n_.edges[e];
            cost(2 * COST_OP);
            if (recv_chan ? (E.channel != -recv_chan) : (E.channel < 0)) continue;
            cost(E.guard_count * 2 * COST_OP);
            if (guards_hold(E)) out->push_back(e);
        }
    }

    long long weight(int e) const {
        const smc_edge &E = n_.edges[e];
        const double w = E.weight >= 0 ? eval(n_, E.weight, vars_.data()) : 1.0;
        const int wi = to_int_sat(w);
        return wi > 0 ? wi : 0;
    }

    int pick(const std::vector<int> &en, double u) const {
        if (en.size() == 1) return en[0];
        cost(COST_U01 + 4 * COST_OP * (double)en.size());  // the caller charges the draw
        long long W = 0;
        for (int e : en) W += weight(e);
        const long long rnd = (long long)((double)W * u);
        long long temp = W;
        for (int i = (int)en.size() - 1; i >= 0; --i) {
            const long long w = weight(en[i]);
            if (temp - w <= rnd) return en[i];
            temp -= w;
        }
        return en[0];
    }

    void fire(int c, int e, bool *goal, int *chan) {
        const smc_edge &E = n_.edges[e];
        cost(2 * COST_OP + 2 * COST_OP * E.upd_count);
        for (int k = 0; k < E.upd_count; ++k) {
            const smc_update &U = n_.updates[E.upd_first + k];
            const double v = eval(n_, U.expr, vars_.data());
            if (U.var == qvar_) {
                if (v > qmax_) qmax_ = v;
                if (v < qmin_) qmin_ = v;
            }
            if (n_.var_kind[U.var] == SMC_VAR_CLOCK) {
                stamp_[U.var] = tclk_ - v;
                vars_[U.var] = tclk_ - stamp_[U.var];
            } else {
                vars_[U.var] = v;
            }
        }
        loc_[c] = E.dst;
        if (n_.locations[E.dst].type == SMC_LOC_GOAL) *goal = true;
        if (E.channel > 0) *chan = E.channel;
    }

    // Delay bound of component c from its clock upper-bound invariants;
    // returns false when nothing bounds it (simulation.cu:460-511).
    bool bound_of(int c, double *M) const {
        const smc_location &L = n_.locations[loc_[c]];
        bool bounded = false;
        double m = 1.7976931348623157e308;
        for (int k = 0; k < L.inv_count; ++k) {
            const smc_constraint &I = n_.constraints[L.inv_first + k];
            if (!I.uses_variable || n_.var_kind[I.var] != SMC_VAR_CLOCK) continue;
            if (I.op != SMC_C_LE && I.op != SMC_C_LT) continue;
            cost(4 * COST_OP);
            double ttb = eval(n_, I.rhs, vars_.data()) - vars_[I.var];
            if (I.op == SMC_C_LT) ttb -= 1e-6;
            if (ttb >= 0.0) { m = ttb < m ? ttb : m; bounded = true; }
        }
        *M = m;
        return bounded;
    }

    double rate_of(int c) const {
        const smc_location &L = n_.locations[loc_[c]];
        return L.rate >= 0 ? eval(n_, L.rate, vars_.data()) : 1.0;
    }

    // t: the time before the step (for the crossing of the bound); full:
    // every component races, the null racers included
    //
    // Importance sampling (is_->theta): an exponential component's rate is
    // r 2^theta; a uniform (clock-bounded) component's hazard 1/(M - d) is
    // multiplied by a = 2^theta, so its delay has survival (1 - d/M)^a on
    // the same window [0, M] and is drawn as M (1 - (1 - u)^(1/a)).  Both
    // keep the log-likelihood of a macro-step linear in theta and 2^theta:
    // ln q = ln2 theta (wins) - 2^theta (base cumulative hazard) + const.
    // won: the winner is an exponential or uniform racer (its own draw).
    void race_canonical(uint64_t run, uint32_t step, double t, bool full, double *d_out, int *w_out,
                        bool *won_out) {
        const int N = n_.n_components;
        const float *theta = is_ ? is_->theta : nullptr;
        int zero = -1, wu = -1, ugroup = -1;
        double du = HUGE_VAL;
        cost(2 * COST_OP * N);  // delay class and leaf of every component
        bool uniform = false;
        for (int c = 0; c < N; ++c) {
            leaf_[c] = 0.0;
            null_[c] = 0;
            um_[c] = 0.0;
            const int type = n_.locations[loc_[c]].type;
            if (type > SMC_LOC_GOAL) { if (zero < 0) zero = c; continue; }
            double M;
            if (bound_of(c, &M)) {
                if (M == 0.0) { if (zero < 0) zero = c; continue; }
                const uint32_t g = smc_uniform_draw((uint32_t)c);  // components 2g, 2g+1 share a draw
                if ((int)g != ugroup) {
                    cost(COST_PHILOX);
                    ugroup = (int)g;
                }
                cost(COST_U01 + 3 * COST_OP);
                smc_u32x4 w = smc_draw(seed_, run, step, SMC_P_UNIFORM, g);
                const double u = smc_uniform_u(w, (uint32_t)c);
                const double th = theta ? (double)theta[loc_[c]] : 0.0;
                const double d = th == 0.0 ? M * u : M * (1.0 - smc_exp2(smc_log2(1.0 - u) / smc_exp2(th)));
                um_[c] = M;
                if (d < du) { du = d; wu = c; }
                uniform = true;
            } else if (freeze_ && blocked_[c]) {
                null_[c] = 1;  // frozen
            } else {
                const double r = rate_of(c);
                leaf_[c] = r > 0.0 ? r : 0.0;
            }
        }
        // A component that cannot fire only makes null steps.  Leaving it
        // out keeps the law of every other event: exponential racers are
        // memoryless, and a uniform racer redrawn on (0, W - tau] at every
        // step has hazard 1/(W - tau), a function of the absolute time tau
        // alone (W: the absolute end of its window), so its firing time
        // does not depend on how many null steps come between.  It races as
        // in cuSMC when the step itself decides the outcome (full).
        if (full)
            for (int c = 0; c < N; ++c)
                if (null_[c]) {
                    const double r = rate_of(c);
                    leaf_[c] = r > 0.0 ? r : 0.0;
                }
        if (zero >= 0) { *d_out = 0.0; *w_out = zero; return; }
        {
            int P = 1, lv = 0;
            while (P < N) { P <<= 1; ++lv; }
            // tree sum, descent, exponential draw, winner against the uniforms
            cost(COST_OP * (P - 1) + 3 * COST_OP * lv + COST_PHILOX + 2 * COST_U01 + COST_LOG2 + COST_DIV +
                 4 * COST_OP);
        }
        if (theta) {
            // tilted proposal: rate r 2^theta[location] for exponential components
            tilted_.assign(leaf_.begin(), leaf_.end());
            for (int c = 0; c < N; ++c)
                if (leaf_[c] > 0.0) tilted_[c] = leaf_[c] * smc_exp2((double)theta[loc_[c]]);
            tree_.build(tilted_, N);
            base_tree_.build(leaf_, N);
        } else {
            tree_.build(leaf_, N);
        }
        const double Rt = tree_.t[1];
        const double R = theta ? base_tree_.t[1] : Rt;
        double dexp = HUGE_VAL;
        int wexp = 0x7FFFFFFF;
        if (Rt > 0.0) {
            smc_u32x4 w = smc_draw(seed_, run, step, SMC_P_RACE, 0);
            dexp = -smc_log2(smc_u01(w.v[0], w.v[1])) / Rt;
            wexp = tree_.descend(smc_u01(w.v[2], w.v[3]) * Rt);
        }
        if (wu < 0) wu = 0x7FFFFFFF;
        bool won = false, uwon = false;
        if (du < dexp || (du == dexp && wu < wexp)) { *d_out = du; *w_out = wu; uwon = true; }
        else if (dexp < HUGE_VAL) { *d_out = dexp; *w_out = wexp; won = true; }
        else { *d_out = HUGE_VAL; *w_out = -1; }  // no racer: no event
        // the step that crosses the bound: the left-out null racers' first
        // event after the bound (base rates) may come first; then the step
        // is a null step (w = -1)
        if (!full && *d_out < HUGE_VAL && !(t + *d_out < q_.time_bound)) {
            bool any = false;
            dleaf_.assign(N, 0.0);
            for (int c = 0; c < N; ++c)
                if (null_[c]) {
                    const double r = rate_of(c);
                    dleaf_[c] = r > 0.0 ? r : 0.0;
                    any = true;
                }
            if (any) {
                dead_tree_.build(dleaf_, N);
                const double Rd = dead_tree_.t[1];
                if (Rd > 0.0) {
                    smc_u32x4 w = smc_draw(seed_, run, step, SMC_P_DEAD, 0);
                    const double td = q_.time_bound + (-smc_log2(smc_u01(w.v[0], w.v[1])) / Rd);
                    if (td < t + *d_out) { *d_out = td - t; *w_out = -1; won = false; uwon = false; }
                }
            }
        }
        if (won_out) *won_out = won || uwon;
        if (theta && *d_out < HUGE_VAL) {
            // log2 of the likelihood ratio of this race: -theta_w (the
            // winner's own draw) + (R' - R) d for the exponential racers
            // + (1 - a_c) log2(1 - d/M_c) for each tilted uniform racer c
            // (the uniform terms summed on the race's pairwise tree)
            const double th_w = (won || uwon) ? (double)theta[loc_[*w_out]] : 0.0;
            double lr = log2lr_ - th_w + (Rt - R) * *d_out;
            if (uniform) {
                dleaf_.assign(N, 0.0);
                for (int c = 0; c < N; ++c)
                    if (um_[c] > 0.0 && theta[loc_[c]] != 0.0)
                        dleaf_[c] = (1.0 - smc_exp2((double)theta[loc_[c]])) * smc_log2(1.0 - *d_out / um_[c]);
                dead_tree_.build(dleaf_, N);
                lr = lr + dead_tree_.t[1];
            }
            log2lr_ = lr;
        }
        if (hazards_ && *d_out < HUGE_VAL) {
            // recording: each racer's base cumulative hazard over this step
            // (the step's whole delay, as in its likelihood ratio): r ln2 d
            // (exponential), -ln(1 - d/M) (uniform)
            const double LN2 = 0.69314718055994530942;
            const double dd = *d_out;
            for (int c = 0; c < N; ++c) {
                if (leaf_[c] > 0.0) (*hazards_)[c] += LN2 * leaf_[c] * dd;
                else if (um_[c] > 0.0) (*hazards_)[c] += -LN2 * smc_log2(1.0 - dd / um_[c]);
            }
        }
    }

    // cuSMC's own sampling: each component draws (simulation.cu:513-540),
    // then the minimum with the lowest index on ties (:593-625).
    void race_literal(uint64_t run, uint32_t step, double *d_out, int *w_out) {
        const int N = n_.n_components;
        double best = HUGE_VAL;
        int win = -1;
        for (int c = 0; c < N; ++c) {
            double d;
            const int type = n_.locations[loc_[c]].type;
            if (type > SMC_LOC_GOAL) d = 0.0;
            else {
                double M;
                smc_u32x4 w = smc_draw(seed_, run, step, SMC_P_UNIFORM, (uint32_t)c);
                if (bound_of(c, &M)) d = M * smc_u01(w.v[0], w.v[1]);
                else {
                    const double r = rate_of(c);
                    d = r > 0.0 ? -smc_log2(smc_u01(w.v[0], w.v[1])) / r : HUGE_VAL;
                }
            }
            if (d < best) { best = d; win = c; }
        }
        *d_out = best;
        *w_out = win < 0 ? 0 : win;
    }
};

}  // namespace

extern "C" {

const char *oracle_last_error(void) { return g_err.c_str(); }

double oracle_tree_pick(const double *rates, int n, double v, int *chosen) {
    RaceTree t;
    std::vector<double> leaf(rates, rates + n);
    t.build(leaf, n);
    if (chosen) *chosen = t.t[1] > 0.0 ? t.descend(v * t.t[1]) : -1;
    return t.t[1];
}

int oracle_run(const smc_network *net, const smc_query *q, uint64_t seed, uint64_t run_begin,
               uint64_t n_runs, int mode, int threads, oracle_outputs *out, oracle_result *res) {
    if (!net || !q || !res) { g_err = "null argument"; return -1; }
    if (q->kind == SMC_QUERY_VAR && (q->var < 0 || q->var >= net->n_vars)) { g_err = "query variable out of range"; return -1; }
    if (mode != ORACLE_CANONICAL && mode != ORACLE_LITERAL) { g_err = "unknown mode"; return -1; }
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    if ((uint64_t)threads > n_runs) threads = (int)std::max<uint64_t>(1, n_runs);
    std::vector<RunOut> outs(n_runs);
    auto t0 = std::chrono::steady_clock::now();
    auto work = [&](int k) {
        Simulator sim(*net, *q, seed, mode);
        const uint64_t b = n_runs * k / threads, e = n_runs * (k + 1) / threads;
        for (uint64_t i = b; i < e; ++i) outs[i] = sim.run(run_begin + i);
    };
    if (threads == 1) work(0);
    else {
        std::vector<std::thread> pool;
        for (int k = 0; k < threads; ++k) pool.emplace_back(work, k);
        for (auto &th : pool) th.join();
    }
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    memset(res, 0, sizeof *res);
    res->n_runs = n_runs;
    for (uint64_t i = 0; i < n_runs; ++i) {
        const RunOut &o = outs[i];
        res->hits += o.flag;
        res->sum += o.value;
        res->sumsq += o.value * o.value;
        res->steps += o.steps;
        res->transitions += o.transitions;
        if (out) {
            if (out->flag) out->flag[i] = o.flag;
            if (out->value) out->value[i] = o.value;
            if (out->steps) out->steps[i] = o.steps;
            if (out->transitions) out->transitions[i] = o.transitions;
            if (out->digest) out->digest[i] = o.digest;
            if (out->time) out->time[i] = o.time;
        }
    }
    res->estimate = (q->kind == SMC_QUERY_VAR && q->is_estimate) ? res->sum / (double)n_runs
                                                                 : (double)res->hits / (double)n_runs;
    res->seconds = secs;
    res->threads = threads;
    return 0;
}

int oracle_is_run(const smc_network *net, const smc_query *q, uint64_t seed, uint64_t run_begin, uint64_t n_runs,
                  const float *theta, int early_stop, int threads, oracle_is_outputs *out) {
    if (!net || !q || !out) { g_err = "null argument"; return -1; }
    if (q->kind == SMC_QUERY_VAR && (q->var < 0 || q->var >= net->n_vars)) { g_err = "query variable out of range"; return -1; }
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    if ((uint64_t)threads > n_runs) threads = (int)std::max<uint64_t>(1, n_runs);
    IsOpts opts;
    opts.theta = theta;
    opts.early_stop = early_stop != 0;
    auto work = [&](int k) {
        Simulator sim(*net, *q, seed, ORACLE_CANONICAL);
        const uint64_t b = n_runs * k / threads, e = n_runs * (k + 1) / threads;
        for (uint64_t i = b; i < e; ++i) {
            RunOut o = sim.run(run_begin + i, nullptr, nullptr, &opts);
            if (out->flag) out->flag[i] = o.flag;
            if (out->log2lr) out->log2lr[i] = o.log2lr;
            if (out->digest) out->digest[i] = o.digest;
            if (out->steps) out->steps[i] = o.steps;
            if (out->transitions) out->transitions[i] = o.transitions;
            if (out->time) out->time[i] = o.time;
        }
    };
    if (threads == 1) work(0);
    else {
        std::vector<std::thread> pool;
        for (int k = 0; k < threads; ++k) pool.emplace_back(work, k);
        for (auto &th : pool) th.join();
    }
    return 0;
}

long oracle_is_record(const smc_network *net, const smc_query *q, uint64_t seed, uint64_t run, const float *theta,
                      int early_stop, const float *var_scale, int fpad, int lpad, uint16_t *X, float *A, float *B,
                      long cap) {
    if (!net || !q || !var_scale) { g_err = "null argument"; return -1; }
    if (fpad < net->n_locations + net->n_vars + 1 || lpad < net->n_locations) { g_err = "row length too small"; return -1; }
    std::vector<uint16_t> xs;
    std::vector<float> as, bs;
    IsOpts opts;
    opts.theta = theta;
    opts.early_stop = early_stop != 0;
    opts.var_scale = var_scale;
    opts.fpad = fpad;
    opts.lpad = lpad;
    opts.X = &xs;
    opts.A = &as;
    opts.B = &bs;
    Simulator sim(*net, *q, seed, ORACLE_CANONICAL);
    sim.run(run, nullptr, nullptr, &opts);
    const long n = (long)(xs.size() / (size_t)fpad);
    for (long k = 0; k < n && k < cap; ++k) {
        if (X) memcpy(X + (size_t)k * fpad, xs.data() + (size_t)k * fpad, (size_t)fpad * 2);
        if (A) memcpy(A + (size_t)k * lpad, as.data() + (size_t)k * lpad, (size_t)lpad * 4);
        if (B) memcpy(B + (size_t)k * lpad, bs.data() + (size_t)k * lpad, (size_t)lpad * 4);
    }
    return n;
}

int oracle_cost(const smc_network *net, const smc_query *q, uint64_t seed, uint64_t run_begin, uint64_t n_runs,
                int threads, uint64_t *steps, uint64_t *transitions, double *lane_ops) {
    if (!net || !q) { g_err = "null argument"; return -1; }
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    if ((uint64_t)threads > n_runs) threads = (int)std::max<uint64_t>(1, n_runs);
    std::vector<uint64_t> st(threads, 0), tr(threads, 0);
    std::vector<double> co(threads, 0.0);
    auto work = [&](int k) {
        Simulator sim(*net, *q, seed, ORACLE_CANONICAL);
        g_count = true;
        g_cost = 0.0;
        const uint64_t b = n_runs * k / threads, e = n_runs * (k + 1) / threads;
        for (uint64_t i = b; i < e; ++i) {
            RunOut o = sim.run(run_begin + i);
            st[k] += o.steps;
            tr[k] += o.transitions;
        }
        co[k] = g_cost;
        g_count = false;
    };
    if (threads == 1) work(0);
    else {
        std::vector<std::thread> pool;
        for (int k = 0; k < threads; ++k) pool.emplace_back(work, k);
        for (auto &th : pool) th.join();
    }
    if (steps) *steps = std::accumulate(st.begin(), st.end(), (uint64_t)0);
    if (transitions) *transitions = std::accumulate(tr.begin(), tr.end(), (uint64_t)0);
    if (lane_ops) *lane_ops = std::accumulate(co.begin(), co.end(), 0.0);
    return 0;
}

void oracle_philox(const uint32_t *ctr, const uint32_t *key, uint32_t *out) {
    smc_u32x4 c;
    for (int i = 0; i < 4; ++i) c.v[i] = ctr[i];
    smc_u32x4 o = smc_philox4x32_10(c, key[0], key[1]);
    for (int i = 0; i < 4; ++i) out[i] = o.v[i];
}
double oracle_u01(uint32_t hi, uint32_t lo) { return smc_u01(hi, lo); }
double oracle_shared_u(const uint32_t *w, uint32_t c, int receiver) {
    smc_u32x4 x;
    for (int i = 0; i < 4; ++i) x.v[i] = w[i];
    return receiver ? smc_receiver_u(x, c) : smc_uniform_u(x, c);
}
uint32_t oracle_shared_draw(uint32_t c, int receiver) {
    return receiver ? smc_receiver_draw(c) : smc_uniform_draw(c);
}
double oracle_log2(double x) { return smc_log2(x); }
double oracle_pow(double a, double b) { return smc_pow(a, b); }
double oracle_eval_expr(const smc_network *net, int e, const double *vars) { return eval(*net, e, vars); }

long oracle_trace(const smc_network *net, const smc_query *q, uint64_t seed, uint64_t run, int mode,
                  int32_t *events, double *times, long cap) {
    if (!net || !q) { g_err = "null argument"; return -1; }
    Simulator sim(*net, *q, seed, mode);
    std::vector<int32_t> ev;
    std::vector<double> tm;
    sim.run(run, &ev, &tm);
    long n = (long)tm.size();
    for (long k = 0; k < n && k < cap; ++k) {
        if (events) for (int j = 0; j < 4; ++j) events[4 * k + j] = ev[4 * k + j];
        if (times) times[k] = tm[k];
    }
    return n;
}

}  // extern "C"
