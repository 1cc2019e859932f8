// compile.cpp — lowers an smc_network to the run kernel's tables.
//
// What the reference does at this point (shared_model_state.cu:164-689,
// pn_compile_visitor.cpp) and what is done here instead:
//   * expressions: cuSMC keeps guard right-hand sides as pointer trees and
//     compiles the rest to a Polish-notation array evaluated with a
//     10-double stack in local memory (expressions.cu:97-119).  Here every
//     expression becomes an inline constant, an inline variable or a short
//     three-address program over four registers (Sethi-Ullman order), with
//     identical programs and constants shared across components.
//   * invariants: only conjuncts that can bound a delay are kept — a clock
//     on the left, `<` or `<=` (simulation.cu:464-509); the kernel never
//     reads any other invariant.
//   * guards: only conjuncts on a variable are kept (simulation.cu:209); an
//     edge with a `!=` conjunct can never be enabled (simulation.cu:249-251)
//     and is flagged instead.
//   * locations get a delay class so the kernel does not re-derive it.
#include "compile.hpp"

#include <cstring>
#include <functional>
#include <map>
#include <stdexcept>
#include <algorithm>
#include <cmath>
#include <string>
#include <unordered_map>
#include "expr_ops.h"
#include "network.hpp"

namespace smc {

namespace {

class Lowerer {
public:
    Lowerer(const smc_network &n, ModelImage &img) : n_(n), img_(img) {}

    uint32_t ref(int e, double dflt) {
        if (e < 0) return SMC_REF_MAKE(SMC_REF_CONST, const_index(dflt));
        const smc_expr_node &x = n_.exprs[e];
        if (x.op == SMC_E_LIT) return SMC_REF_MAKE(SMC_REF_CONST, const_index(x.value));
        if (x.op == SMC_E_RANDOM) return SMC_REF_MAKE(SMC_REF_CONST, const_index(1.0));
        if (x.op == SMC_E_VAR)
            return SMC_REF_MAKE(n_.var_kind[x.var] == SMC_VAR_CLOCK ? SMC_REF_CLOCK : SMC_REF_VAR, (uint32_t)x.var);
        std::vector<DevInsn> prog;
        int maxreg = 0;
        uint16_t r = emit(e, 0, prog, &maxreg);
        if (r != SMC_OPND(SMC_OPND_REG, 0)) throw std::runtime_error("internal: expression result not in r0");
        if ((int)prog.size() > SMC_MAX_CODE_LEN)
            throw std::runtime_error("limit: expression needs more than 63 instructions");
        std::string key((const char *)prog.data(), prog.size() * sizeof(DevInsn));
        auto it = code_cache_.find(key);
        uint32_t first;
        if (it != code_cache_.end()) first = it->second;
        else {
            first = (uint32_t)img_.code.size();
            if (first + prog.size() >= (1u << 24)) throw std::runtime_error("limit: bytecode larger than 16M instructions");
            img_.code.insert(img_.code.end(), prog.begin(), prog.end());
            code_cache_[key] = first;
        }
        return SMC_REF_CODE_MAKE(first, prog.size());
    }

private:
    const smc_network &n_;
    ModelImage &img_;
    std::unordered_map<uint64_t, uint32_t> const_cache_;
    std::unordered_map<std::string, uint32_t> code_cache_;

    uint32_t const_index(double v) {
        uint64_t b;
        memcpy(&b, &v, 8);
        auto it = const_cache_.find(b);
        if (it != const_cache_.end()) return it->second;
        uint32_t i = (uint32_t)img_.consts.size();
        if (i > SMC_MAX_OPND_INDEX) throw std::runtime_error("limit: more than 16384 distinct constants");
        img_.consts.push_back(v);
        const_cache_[b] = i;
        return i;
    }

    int need(int e) const {
        const smc_expr_node &x = n_.exprs[e];
        if (x.op == SMC_E_LIT || x.op == SMC_E_VAR || x.op == SMC_E_RANDOM) return 0;
        int ar = smc_op_arity(x.op);
        const int kids[3] = {x.a, x.b, x.c};
        int nk[3] = {0, 0, 0};
        for (int i = 0; i < ar; ++i) nk[i] = need(kids[i]);
        // evaluate neediest first; each evaluated non-leaf child holds one register
        int order[3] = {0, 1, 2};
        for (int i = 0; i < ar; ++i)
            for (int j = i + 1; j < ar; ++j)
                if (nk[order[j]] > nk[order[i]]) std::swap(order[i], order[j]);
        int held = 0, peak = 1;
        for (int i = 0; i < ar; ++i) {
            int k = order[i];
            if (nk[k] > 0) { peak = std::max(peak, held + nk[k]); ++held; }
        }
        return std::max(peak, 1);
    }

    uint16_t emit(int e, int base, std::vector<DevInsn> &prog, int *maxreg) {
        const smc_expr_node &x = n_.exprs[e];
        if (x.op == SMC_E_LIT) return SMC_OPND(SMC_OPND_CONST, const_index(x.value));
        if (x.op == SMC_E_RANDOM) return SMC_OPND(SMC_OPND_CONST, const_index(1.0));
        if (x.op == SMC_E_VAR) {
            if (x.var > SMC_MAX_OPND_INDEX) throw std::runtime_error("limit: variable index above 16383 in an expression");
            return SMC_OPND(n_.var_kind[x.var] == SMC_VAR_CLOCK ? SMC_OPND_CLOCK : SMC_OPND_VAR, (uint32_t)x.var);
        }
        int ar = smc_op_arity(x.op);
        const int kids[3] = {x.a, x.b, x.c};
        int nk[3] = {0, 0, 0};
        for (int i = 0; i < ar; ++i) nk[i] = need(kids[i]);
        int order[3] = {0, 1, 2};
        for (int i = 0; i < ar; ++i)
            for (int j = i + 1; j < ar; ++j)
                if (nk[order[j]] > nk[order[i]]) std::swap(order[i], order[j]);
        uint16_t opnd[3] = {0, 0, 0};
        int r = base;
        for (int i = 0; i < ar; ++i) {
            int k = order[i];
            opnd[k] = emit(kids[k], r, prog, maxreg);
            if ((opnd[k] >> 14) == SMC_OPND_REG) ++r;
        }
        if (base >= SMC_NREGS) throw std::runtime_error("limit: expression needs more than 4 registers");
        *maxreg = std::max(*maxreg, base);
        DevInsn in{};
        in.op = (uint8_t)x.op;
        in.dst = (uint8_t)base;
        in.a = opnd[0]; in.b = opnd[1]; in.c = opnd[2];
        prog.push_back(in);
        return SMC_OPND(SMC_OPND_REG, (uint32_t)base);
    }
};

// Instances of one template differ in their location numbers only where
// they read or write local variables; the rest of their records are equal.
// With edge destinations relative to the source location, equal runs of
// records (a location's edges, an edge's guards and updates, a location's
// invariants and blocking test) are stored once: SIR-350's 112 KB of
// tables shrink to its 34 KB of location records, which fit in LDS.  The
// network's edge numbering stays available for digests (loc_edge0).
void share_records(ModelImage &img) {
    std::vector<DevEdge> edges;
    std::vector<DevGuard> guards;
    std::vector<DevUpd> upds;
    std::vector<DevInv> invs;
    std::unordered_map<std::string, uint32_t> gmap, umap, emap, imap;
    auto intern = [](auto &out, auto &map, const auto *first, size_t n) -> uint32_t {
        if (n == 0) return 0;
        std::string key((const char *)first, n * sizeof(*first));
        auto it = map.find(key);
        if (it != map.end()) return it->second;
        const uint32_t at = (uint32_t)out.size();
        out.insert(out.end(), first, first + n);
        map.emplace(std::move(key), at);
        return at;
    };
    img.loc_edge0.assign(img.locs.size(), 0);
    for (size_t l = 0; l < img.locs.size(); ++l) {
        DevLoc &L = img.locs[l];
        std::vector<DevEdge> run(img.edges.begin() + L.edge_first, img.edges.begin() + L.edge_first + L.edge_n);
        for (DevEdge &d : run) {
            d.dst -= (int32_t)l;
            d.guard_first = intern(guards, gmap, img.guards.data() + d.guard_first, d.guard_n);
            d.upd_first = intern(upds, umap, img.upds.data() + d.upd_first, d.upd_n);
        }
        img.loc_edge0[l] = (int32_t)L.edge_first;
        L.edge_first = intern(edges, emap, run.data(), run.size());
        L.inv_first = intern(invs, imap, img.invs.data() + L.inv_first, L.n_inv);
        if (L.flags & SMC_LF_GUARDED) {
            size_t k = L.block_first;
            while (!(img.guards[k].flags & SMC_GF_END_LOC)) ++k;
            L.block_first = intern(guards, gmap, img.guards.data() + L.block_first, k + 1 - L.block_first);
        }
    }
    img.edges.swap(edges);
    img.guards.swap(guards);
    img.upds.swap(upds);
    img.invs.swap(invs);
}

}  // namespace

ModelImage compile_model(const smc_network &n) {
    std::string err = validate(n);
    if (!err.empty()) throw std::runtime_error(err);
    ModelImage img;
    img.n_comp = n.n_components;
    img.n_vars = n.n_vars;
    int leaves = 1;
    while (leaves < n.n_components) leaves <<= 1;
    img.n_leaves = leaves;
    Lowerer lw(n, img);

    img.var_init.assign(n.var_init, n.var_init + n.n_vars);
    img.var_kind.assign(n.var_kind, n.var_kind + n.n_vars);
    for (int v = 0; v < n.n_vars; ++v) img.var_clock.push_back(n.var_kind[v] == SMC_VAR_CLOCK ? 1 : 0);
    img.comp_init_loc.assign(n.comp_init_loc, n.comp_init_loc + n.n_components);
    img.comp_first_loc.assign(n.comp_first_loc, n.comp_first_loc + n.n_components + 1);

    for (int l = 0; l < n.n_locations; ++l) {
        const smc_location &L = n.locations[l];
        DevLoc d{};
        d.type = (uint8_t)L.type;
        d.rate = lw.ref(L.rate, 1.0);
        d.inv_first = (uint32_t)img.invs.size();
        for (int k = 0; k < L.inv_count; ++k) {
            const smc_constraint &C = n.constraints[L.inv_first + k];
            if (!C.uses_variable) continue;
            if (n.var_kind[C.var] != SMC_VAR_CLOCK) continue;
            if (C.op != SMC_C_LE && C.op != SMC_C_LT) continue;
            DevInv iv{};
            iv.var = (uint16_t)C.var;
            iv.strict = C.op == SMC_C_LT;
            iv.bound = lw.ref(C.rhs, 0.0);
            if (SMC_REF_KIND(iv.bound) == SMC_REF_CONST) {
                iv.bound_const = 1;
                iv.bound_value = img.consts[iv.bound & 0x3FFFFFFFu];
            }
            img.invs.push_back(iv);
        }
        size_t ninv = img.invs.size() - d.inv_first;
        if (ninv > 65535) throw std::runtime_error("limit: more than 65535 invariants on a location");
        d.n_inv = (uint16_t)ninv;
        if (ninv == 1 && img.invs[d.inv_first].bound_const) {
            d.flags |= SMC_LF_INL_INV;
            d.inl_value = img.invs[d.inv_first].bound_value;
            d.inl_var = img.invs[d.inv_first].var;
            d.inl_op = img.invs[d.inv_first].strict;
        }
        d.klass = L.type > SMC_LOC_GOAL ? SMC_K_ZERO : (ninv > 0 ? SMC_K_BOUNDED : SMC_K_EXP);
        if (L.edge_count > SMC_MAX_EDGES_PER_LOC)
            throw std::runtime_error("limit: a location has more than 64 outgoing edges");
        d.edge_first = (uint32_t)L.edge_first;
        d.edge_n = (uint16_t)L.edge_count;
        if (SMC_REF_KIND(d.rate) == SMC_REF_CONST) {
            d.rate_const = 1;
            d.rate_value = img.consts[d.rate & 0x3FFFFFFFu];
        }
        img.locs.push_back(d);
    }
    // Receive edges whose updates commute within one broadcast.  Forms of
    // an update: ADD (integer variable += integer constant), SET (:= a
    // constant).  Per variable over all receive edges: all ADD, or all SET
    // of one value, or written by a single component — then the writes of
    // one broadcast's receivers give the same result in any order.
    auto literal = [&](int x, double *v) {
        if (x < 0 || n.exprs[x].op != SMC_E_LIT) return false;
        *v = n.exprs[x].value;
        return true;
    };
    auto is_var = [&](int x, int var) { return x >= 0 && n.exprs[x].op == SMC_E_VAR && n.exprs[x].var == var; };
    enum { F_OTHER = 0, F_ADD = 1, F_SET = 2 };
    auto form = [&](const smc_update &U, double *c) -> int {
        const smc_expr_node *x = U.expr >= 0 ? &n.exprs[U.expr] : nullptr;
        if (!x) return F_OTHER;
        if (n.var_kind[U.var] == SMC_VAR_INT &&
            ((x->op == SMC_E_ADD && is_var(x->a, U.var) && literal(x->b, c)) ||
             (x->op == SMC_E_ADD && literal(x->a, c) && is_var(x->b, U.var)) ||
             (x->op == SMC_E_SUB && is_var(x->a, U.var) && literal(x->b, c) && ((*c = -*c), true))) &&
            *c == std::floor(*c) && std::fabs(*c) < 2147483648.0)
            return F_ADD;
        if (literal(U.expr, c)) return F_SET;
        return F_OTHER;
    };
    struct VarUse { int add = 0, set = 0, other = 0, writer = -1; bool one_value = true; double value = 0.0; };
    std::vector<VarUse> use(n.n_vars);
    for (int c = 0; c < n.n_components; ++c)
        for (int l = n.comp_first_loc[c]; l < n.comp_first_loc[c + 1]; ++l)
            for (int k = 0; k < n.locations[l].edge_count; ++k) {
                const smc_edge &E = n.edges[n.locations[l].edge_first + k];
                if (E.channel >= 0) continue;
                for (int u = 0; u < E.upd_count; ++u) {
                    const smc_update &U = n.updates[E.upd_first + u];
                    VarUse &vu = use[U.var];
                    double cv = 0.0;
                    const int f = form(U, &cv);
                    if (f == F_ADD) ++vu.add;
                    else if (f == F_SET) {
                        if (vu.set && cv != vu.value) vu.one_value = false;
                        vu.value = cv;
                        ++vu.set;
                    } else ++vu.other;
                    vu.writer = (vu.writer == -1 || vu.writer == c) ? c : -2;
                }
            }
    auto var_commutes = [&](int v) {
        const VarUse &vu = use[v];
        if (vu.other) return false;
        if (vu.writer >= 0) return true;
        return vu.set == 0 || (vu.add == 0 && vu.one_value);
    };
    std::vector<uint8_t> fast_edge(n.n_edges, 0);
    std::vector<int8_t> clock_memo(n.n_exprs, -1);
    std::function<bool(int)> reads_clock = [&](int x) -> bool {
        if (x < 0) return false;
        if (clock_memo[x] >= 0) return clock_memo[x] != 0;
        const smc_expr_node &nd = n.exprs[x];
        bool r = nd.op == SMC_E_VAR && n.var_kind[nd.var] == SMC_VAR_CLOCK;
        if (!r && nd.op != SMC_E_LIT && nd.op != SMC_E_VAR)
            r = reads_clock(nd.a) || reads_clock(nd.b) || reads_clock(nd.c);
        clock_memo[x] = r ? 1 : 0;
        return r;
    };

    for (int e = 0; e < n.n_edges; ++e) {
        const smc_edge &E = n.edges[e];
        DevEdge d{};
        d.dst = E.dst;
        d.channel = E.channel;
        d.weight = lw.ref(E.weight, 1.0);
        d.dst_goal = n.locations[E.dst].type == SMC_LOC_GOAL;
        d.guard_first = (uint32_t)img.guards.size();
        for (int k = 0; k < E.guard_count; ++k) {
            const smc_constraint &C = n.constraints[E.guard_first + k];
            if (!C.uses_variable) continue;
            if (C.op == SMC_C_NE) { d.never = 1; continue; }
            if (C.var > 65535) throw std::runtime_error("limit: guarded variable index above 65535");
            DevGuard g{};
            g.var = (uint16_t)C.var;
            g.op = (uint8_t)C.op;
            g.flags = (uint8_t)((n.var_kind[C.var] != SMC_VAR_CLOCK && !reads_clock(C.rhs) ? SMC_GF_CLOCK_FREE : 0) |
                                (n.var_kind[C.var] == SMC_VAR_CLOCK ? SMC_GF_CLOCK : 0));
            g.rhs = lw.ref(C.rhs, 0.0);
            img.guards.push_back(g);
        }
        size_t ng = img.guards.size() - d.guard_first;
        if (ng > 65535) throw std::runtime_error("limit: more than 65535 guards on an edge");
        d.guard_n = (uint16_t)ng;
        d.upd_first = (uint32_t)img.upds.size();
        bool commute = E.channel < 0 && SMC_REF_KIND(d.weight) == SMC_REF_CONST;
        std::vector<int> seen;
        for (int k = 0; k < E.upd_count; ++k) {
            const smc_update &U = n.updates[E.upd_first + k];
            if (U.var > 65535) throw std::runtime_error("limit: assigned variable index above 65535");
            DevUpd u{};
            u.var = (uint16_t)U.var;
            u.clock = n.var_kind[U.var] == SMC_VAR_CLOCK ? 1 : 0;
            u.expr = lw.ref(U.expr, 0.0);
            double cv = 0.0;
            const int f = E.channel < 0 ? form(U, &cv) : F_OTHER;
            if (f != F_OTHER && var_commutes(U.var)) {
                u.kind = f == F_ADD ? SMC_U_ADD : SMC_U_SET;
                u.value = cv;
            } else {
                commute = false;
            }
            if (std::find(seen.begin(), seen.end(), U.var) != seen.end()) commute = false;
            seen.push_back(U.var);
            img.upds.push_back(u);
        }
        fast_edge[e] = commute ? 1 : 0;
        if (E.upd_count > 65535) throw std::runtime_error("limit: more than 65535 updates on an edge");
        d.upd_n = (uint16_t)E.upd_count;
        img.edges.push_back(d);
    }
    // time_free: the location's rate and its spontaneous edges' guards read
    // no clock, so between transitions neither its enabled edges nor its
    // rate change
    // Null racers leave an all-exponential race only while every rate is
    // constant between transitions (a rate reading a clock changes at each
    // null step of the reference's race).
    bool clock_rate = false;
    for (int l = 0; l < n.n_locations; ++l)
        if (reads_clock(n.locations[l].rate)) clock_rate = true;
    for (int l = 0; l < n.n_locations; ++l) {
        const smc_location &L = n.locations[l];
        bool tf = !reads_clock(L.rate);
        for (int k = 0; k < L.edge_count && tf; ++k) {
            const smc_edge &E = n.edges[L.edge_first + k];
            if (E.channel < 0) continue;
            for (int g = 0; g < E.guard_count; ++g) {
                const smc_constraint &G = n.constraints[E.guard_first + g];
                if (!G.uses_variable) continue;
                if (n.var_kind[G.var] == SMC_VAR_CLOCK || reads_clock(G.rhs)) tf = false;
            }
        }
        // blocking: always (every spontaneous edge has `!=`, or none), by
        // the variables (every other spontaneous edge has a clock-free
        // conjunct that may be false), or never
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
        img.locs[l].flags |= (uint8_t)((tf ? SMC_LF_TIME_FREE : 0) | (kind == 1 ? SMC_LF_DEAD : 0) |
                                       (kind == 2 ? SMC_LF_GUARDED : 0));
        if (kind == 2 && img.locs[l].klass != SMC_K_ZERO) img.guarded = true;
        if (kind == 2) {
            // the blocking test: clock-free conjuncts, edge after edge
            img.locs[l].block_first = (uint32_t)img.guards.size();
            for (int k = 0; k < L.edge_count; ++k) {
                const smc_edge &E = n.edges[L.edge_first + k];
                if (E.channel < 0) continue;
                const DevEdge &DE = img.edges[L.edge_first + k];
                if (DE.never) continue;
                const size_t first = img.guards.size();
                for (int g = 0; g < DE.guard_n; ++g) {
                    const DevGuard G = img.guards[DE.guard_first + g];
                    if (G.flags & SMC_GF_CLOCK_FREE) img.guards.push_back(G);
                }
                if (img.guards.size() == first) throw std::runtime_error("internal: guarded edge without a clock-free conjunct");
                img.guards.back().flags |= SMC_GF_END_EDGE;
            }
            img.guards.back().flags |= SMC_GF_END_LOC;
            const DevGuard &G = img.guards.back();
            if (img.locs[l].block_first + 1 == img.guards.size() && SMC_REF_KIND(G.rhs) == SMC_REF_CONST &&
                !(img.locs[l].flags & SMC_LF_INL_INV)) {
                img.locs[l].flags |= SMC_LF_INL_BLK;
                img.locs[l].inl_value = img.consts[G.rhs & 0x3FFFFFFFu];
                img.locs[l].inl_var = G.var;
                img.locs[l].inl_op = G.op;
            }
        }
    }
    img.freeze = !clock_rate;
    for (int l = 0; l < n.n_locations; ++l) {
        const smc_location &L = n.locations[l];
        bool fast = true;
        for (int k = 0; k < L.edge_count; ++k)
            if (n.edges[L.edge_first + k].channel < 0 && !fast_edge[L.edge_first + k]) fast = false;
        if (fast) img.locs[l].flags |= SMC_LF_RECV_FAST;
    }
    share_records(img);
    // keep every table non-empty so device pointers are valid
    if (img.invs.empty()) img.invs.push_back(DevInv{});
    if (img.edges.empty()) img.edges.push_back(DevEdge{});
    if (img.guards.empty()) img.guards.push_back(DevGuard{});
    if (img.upds.empty()) img.upds.push_back(DevUpd{});
    if (img.consts.empty()) img.consts.push_back(0.0);
    if (img.code.empty()) img.code.push_back(DevInsn{});
    if (img.var_init.empty()) img.var_init.push_back(0.0);
    if (img.var_clock.empty()) img.var_clock.push_back(0);
    if (img.loc_edge0.empty()) img.loc_edge0.push_back(0);
    return img;
}

double eval_ref_host(const ModelImage &m, uint32_t ref, const double *vars) {
    // vars: values (clocks by value, not as stamps)
    uint32_t kind = SMC_REF_KIND(ref);
    if (kind == SMC_REF_CONST) return m.consts[ref & 0x3FFFFFFFu];
    if (kind == SMC_REF_VAR || kind == SMC_REF_CLOCK) return vars[ref & 0x3FFFFFFFu];
    uint32_t first = ref & 0xFFFFFFu, count = (ref >> 24) & 0x3Fu;
    double reg[SMC_NREGS] = {0, 0, 0, 0};
    auto fetch = [&](uint16_t o) -> double {
        uint32_t k = o >> 14, i = o & 0x3FFFu;
        return k == SMC_OPND_REG ? reg[i & 3] : (k == SMC_OPND_VAR || k == SMC_OPND_CLOCK) ? vars[i] : m.consts[i];
    };
    for (uint32_t k = 0; k < count; ++k) {
        const DevInsn &in = m.code[first + k];
        int ar = smc_op_arity(in.op);
        double a = ar > 0 ? fetch(in.a) : 0.0, b = ar > 1 ? fetch(in.b) : 0.0, c = ar > 2 ? fetch(in.c) : 0.0;
        reg[in.dst & 3] = smc_apply_op(in.op, a, b, c);
    }
    return reg[0];
}

}  // namespace smc
