// sim_body.h — the run kernel for gfx950 (compiled once per slot count and
// register policy by sim_variant.hip).
//
// Counterpart of cuSMC's simulation_kernel (simulation/simulation.cu:674-901)
// and its device functions (compute_possible_delay :409-541,
// find_minimum_delay :544-671, SyncInSerial :357-407, check_edge_enabled
// :184-266, take_transition :27-181), and of the expression evaluator
// (simulation/expressions.cu).  Same step semantics, different machine
// mapping:
//
//   * cuSMC runs one block per run and one thread per component, with
//     ~8 + log2(MC) __syncthreads per step and thread-0 serial sections.
//     Here a run is a segment of S = 2^k lanes of one wave64 (64/S runs per
//     wave), and each lane owns C = next_pow2(components)/S component slots
//     (blocked: lane l holds components l*C .. l*C+C-1).  A workgroup is one
//     wave, so no workgroup barrier is ever needed; lanes of a segment talk
//     through cross-lane shuffles, ballots and the segment's LDS block.
//   * Variables and clocks of each run live in LDS as doubles, clocks as
//     reset stamps beside the run's clock time (a clock's value is the time
//     minus its stamp), so advancing time is one store; time accumulates in
//     FP64.
//   * The race: cuSMC re-samples every component's delay every step
//     (simulation.cu:513-540) and takes the minimum.  Here the same law is
//     sampled with far fewer draws: zero-delay components (branchpoint,
//     urgent, committed, or a clock bound already reached) win first, lowest
//     index first; otherwise the exponential components are one
//     exponential race — minimum delay Exp(sum of rates) and winner chosen in
//     proportion to its rate — and only components under a clock upper bound
//     draw their own uniform delay.  The rate sum and the proportional
//     choice use one fixed pairwise tree over the next power of two of the
//     component count (in-lane levels, then a butterfly over the segment),
//     so the result does not depend on the launch shape and a sequential
//     host implementation reproduces it bit for bit.
//   * Broadcast receivers compute their enabled sets in parallel and then
//     fire one at a time in component order (SyncInSerial), each lane
//     executing its own component's transition.
//   * Runs are handed out by an atomic counter, so a segment that finishes a
//     short run immediately starts the next one.
#pragma once

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>
#include <type_traits>

#include "device_common.h"
#include "kernels.hpp"

namespace {

// ------------------------------------------------------------ run kernel


// One wave at a time takes a chunk of runs of one instance from the launch's
// queue, so the instance (its tables, query and outputs) is uniform across
// the wave and lives in scalar registers; the wave's segments then pull the
// chunk's runs one by one from a counter in LDS.  A workgroup holds
// blockDim/64 independent waves.  With TAB (single-instance launches) the
// workgroup first copies the model's tables into LDS and every wave reads
// them there: the race and the transitions are chains of dependent table
// reads, and LDS answers them several times sooner than the caches.
//
// LDS: [tables: lp.table_bytes, TAB only] then per wave [clock time and
// variables of the wave's 64/S runs: var_stride doubles each][race-tree levels: 2 log2(S)
// doubles per lane, level-major]; lp.wave_lds covers the largest of these
// over the launch's instances.
template <int C, bool TRACE, int TAB>
__device__ __forceinline__ void sim_body(const LaunchParams &lp) {
    extern __shared__ __attribute__((aligned(16))) double lds[];
    __shared__ unsigned int chunk_next_of[SMC_MAX_GROUP_WAVES];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    unsigned int &chunk_next = chunk_next_of[wv];
    double *wbase = lds + (TAB ? lp.table_bytes / 8 : 0) + (size_t)wv * (lp.wave_lds / 8);
    const double INF = __builtin_huge_val();
    const uint64_t round = lp.round ? *lp.round : 0;
    // TAB 1: every hot table staged in LDS; 2: all but the location records
    using Tables = typename std::conditional<
        TAB == 1, SmcLdsTables, typename std::conditional<TAB == 2, SmcMixedTables, SmcGlobalTables>::type>::type;
    if (TAB) stage_tables(lp.tables, lp.table_bytes, lds);

  for (;;) {
    unsigned long long g = 0;
    if (lane == 0) g = atomicAdd(lp.counter, 1ull);
    g = __shfl(g, 0);
    // the same on every lane: make it a scalar, so the instance lookup and
    // its tables and parameters live in scalar registers
    g = ((unsigned long long)__builtin_amdgcn_readfirstlane((uint32_t)(g >> 32)) << 32) |
        (unsigned long long)__builtin_amdgcn_readfirstlane((uint32_t)g);
    if (g >= lp.total_chunks) break;
    int lo = 0, hi = lp.n_inst - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (lp.inst[mid].chunk_first <= g) lo = mid;
        else hi = mid - 1;
    }
    const DevModel mg = lp.inst[lo].m;
    DevModelT<Tables> m;
    if constexpr (TAB == 1) m = lds_model(mg, lp.tables, lds);
    else if constexpr (TAB == 2) m = lds_model_mixed(mg, lp.tables, lds);
    else m = mg;
    const SimParams p = lp.inst[lo].p;
    // the instance's packing: S lanes per run, 64 / S runs per wave
    const int S = lp.inst[lo].lanes;
    const int sl = lane & (S - 1);
    const int seg_base = lane - sl;
    const uint64_t seg_mask = (S == 64) ? ~0ull : (((1ull << S) - 1ull) << seg_base);
    // a run's block: [clock time][variables] (clocks as reset stamps)
    double *vars = wbase + (size_t)(lane / S) * lp.var_stride + 1;
    // per lane: the race tree's partial sums at each butterfly level (own at
    // [64 (2k)], sibling block's at [64 (2k + 1)]), kept in LDS rather than
    // registers from the sum to the descent; level-major, so the 64 lanes
    // touch 64 consecutive doubles (no bank conflicts)
    double *lvl = wbase + (size_t)(64 / S) * lp.var_stride + lane;
    const uint32_t chunk_runs = lp.inst[lo].chunk_runs;
    const uint64_t chunk = g - lp.inst[lo].chunk_first;
    const uint32_t r_first = (uint32_t)(chunk * chunk_runs);
    const uint32_t r_end = (uint32_t)((uint64_t)r_first + chunk_runs < p.n_runs ? r_first + chunk_runs : p.n_runs);
    if (lane == 0) chunk_next = r_first;
    wave_sync();
    const int n = m.n_comp;

    int loc[C];
#pragma unroll
    for (int j = 0; j < C; ++j) loc[j] = -1;

    bool active = false, done = false;
    uint64_t run = 0, ridx = 0;
    double t = 0.0;
    uint32_t step = 0, trans = 0;
    int chan = 0, sender = -1;
    bool goal = false, goal_local = false;
    bool changed = true;  // state changed since the last absorption check
    uint32_t blk = 0;     // blocked slots at the last change of state
    uint64_t digest = 0;
    LaneQuery q{p.qvar, p.qinit, p.qinit};

    for (;;) {
        if (!active && !done) {
            unsigned int r = 0;
            if (sl == 0) r = atomicAdd(&chunk_next, 1u);
            r = __shfl(r, seg_base);
            if (r >= r_end) {
                done = true;
            } else {
                ridx = r;
                run = p.run_begin + round * p.n_runs + r;
                active = true;
                t = 0.0; step = 0; trans = 0; chan = 0; sender = -1;
                goal = false; goal_local = false; changed = true;
                q.qmax = p.qinit; q.qmin = p.qinit;
                digest = TRACE ? smc_mix(0, run) : 0;
                for (int v = sl; v < m.n_vars; v += S) vars[v] = m.var_init[v];
                if (sl == 0) vars[-1] = 0.0;
#pragma unroll
                for (int j = 0; j < C; ++j) {
                    const int c = sl * C + j;
                    loc[j] = c < n ? m.comp_init_loc[c] : -1;
                }
                wave_sync();
            }
        }
        if (__ballot(!done) == 0) break;
        if (done) continue;

        // ---- loop test (simulation.cu:835-849)
        bool hit = false, end = false;
        if (!(t < p.bound)) end = true;
        else if (p.reach && goal) { end = true; hit = true; }
        else if (step >= p.max_steps) end = true;

        if (!end) {
            double d = 0.0;
            int w = -1;
            // phase 0: pending broadcast receivers (SyncInSerial simulation.cu:357-407)
            // phase 1: race (compute_possible_delay :409-541, find_minimum_delay
            //          :544-671), clock advance, the winner's transition
#pragma clang loop unroll(disable)
            for (int phase = 0; phase < 2; ++phase) {
                if (phase == 0 && chan <= 0) continue;
                if (phase == 1) {
                    // a goal entered by this iteration's receivers counts
                    // only if this step, raced by every component, keeps
                    // t < bound (simulation.cu:409-415, 835-849): no
                    // absorption exit and no left-out null racers then
                    const bool full = p.reach && goal;
                    // blocked bits: after a change of state, in the slot loop
                    // below on models with frozen racers (it loads each slot's
                    // location anyway), here for the absorption scan otherwise
                    const bool reblk = changed && m.guarded && m.freeze;
                    if (changed) {
                        if (m.guarded && !m.freeze) blk = blocked_slots<C>(loc, vars, m, sl, n);
                        if (!full) {
                            changed = false;
                            // with frozen racers the race itself ends a run that nothing
                            // can leave (no event: blocked racers are out, the rest race at
                            // rate 0), so the scan is needed only without them
                            if (!m.freeze && absorbed<C>(loc, vars, m, sl, n, seg_mask, blk)) { end = true; break; }
                        }
                    }

                    double leaf[C], Mb[C];
                    uint32_t rref[C];     // rate expression of slots in rpend
                    int zero_j = C;
                    uint32_t bmask = 0, deadmask = 0, rpend = 0;
                    RateMemo memo;
#pragma unroll
                    for (int j = 0; j < C; ++j) {
                        leaf[j] = 0.0;
                        Mb[j] = 0.0;
                        rref[j] = 0xFFFFFFFFu;
                        const int c = sl * C + j;
                        if (c < n) {
                            const DevLoc L = load_loc(m.locs + loc[j]);
                            if (reblk && (L.flags & SMC_LF_GUARDED) && L.klass != SMC_K_ZERO) {
                                const uint32_t bit = 1u << j;
                                blk = location_blocked(L, vars, m) ? (blk | bit) : (blk & ~bit);
                            }
                            if (L.klass == SMC_K_ZERO) {
                                if (zero_j == C) zero_j = j;
                            } else {
                                bool bounded = false;
                                double M = DBL_MAX;
                                if (L.flags & SMC_LF_INL_INV) {
                                    double ttb = L.inl_value - (vars[-1] - vars[L.inl_var]);
                                    if (L.inl_op) ttb -= 1e-6;
                                    if (ttb >= 0.0) { M = ttb; bounded = true; }
                                } else for (int i = 0; i < L.n_inv; ++i) {
                                    const DevInv iv = load_rec(m.invs + L.inv_first + i);
                                    const double bound = iv.bound_const ? iv.bound_value : eval_ref(iv.bound, vars, m);
                                    double ttb = bound - (vars[-1] - vars[iv.var]);
                                    if (iv.strict) ttb -= 1e-6;
                                    if (ttb >= 0.0) { M = ttb < M ? ttb : M; bounded = true; }
                                }
                                if (bounded) {
                                    if (M == 0.0) { if (zero_j == C) zero_j = j; }
                                    else { Mb[j] = M; bmask |= 1u << j; }
                                } else if (m.freeze && ((L.flags & SMC_LF_DEAD) || ((blk >> j) & 1u))) {
                                    deadmask |= 1u << j;
                                } else if (L.rate_const) {
                                    leaf[j] = L.rate_value > 0.0 ? L.rate_value : 0.0;
                                } else {
                                    rref[j] = L.rate;
                                    rpend |= 1u << j;
                                }
                            }
                        }
                    }
                    pending_rates<C>(rpend, rref, leaf, vars, m);
                    // A frozen component (blocked, exponential) only makes
                    // null steps until the next transition.  Leaving it out of
                    // the race keeps the law of everything else: exponential
                    // racers are memoryless (rates constant between
                    // transitions), and a uniform racer redrawn at every step
                    // has hazard 1/(W - tau), a function of the absolute time
                    // alone (W: the end of its window).  It races as in the
                    // reference when the step decides the outcome (full).
                    if (full) {
                        for (uint32_t b = deadmask; b; b &= b - 1) {
                            const int j = __ffs(b) - 1;
                            const DevLoc L = load_loc(m.locs + slot_get<C>(loc, j));
                            const double r = location_rate(L, vars, m, memo);
                            slot_set<C>(leaf, j, r > 0.0 ? r : 0.0);
                        }
                    }
                    const uint64_t zb = __ballot(zero_j < C) & seg_mask;
                    if (zb) {
                        const int Lz = lowest_lane(zb);
                        w = (Lz - seg_base) * C + __shfl(zero_j, Lz);
                        d = 0.0;
                    } else {
                        // uniform delays under clock bounds, lowest slot first on ties;
                        // slots 2i and 2i+1 of a lane share one draw
                        double umin = INF;
                        int unif_j = -1;
                        for (uint32_t b = bmask; b;) {
                            const int j = __ffs(b) - 1;
                            const uint32_t c = (uint32_t)(sl * C + j);
                            const smc_u32x4 rw = smc_draw(p.seed, run, step, SMC_P_UNIFORM, smc_uniform_draw(c));
                            const double du = slot_get<C>(Mb, j) * smc_uniform_u(rw, c);
                            if (du < umin) { umin = du; unif_j = j; }
                            b &= b - 1;
                            if (C > 1 && !(c & 1u) && ((b >> (j + 1)) & 1u)) {
                                const double du1 = slot_get<C>(Mb, j + 1) * smc_uniform_u(rw, c + 1);
                                if (du1 < umin) { umin = du1; unif_j = j + 1; }
                                b &= b - 1;
                            }
                        }
                        // pairwise tree over the lane's C leaves (heap order,
                        // root tr[1]); only the root lives through the butterfly
                        double s;
                        {
                            double tr[2 * C];
#pragma unroll
                            for (int j = 0; j < C; ++j) tr[C + j] = leaf[j];
#pragma unroll
                            for (int i = C - 1; i >= 1; --i) tr[i] = tr[2 * i] + tr[2 * i + 1];
                            s = (C == 1) ? leaf[0] : tr[1];
                        }
#pragma unroll
                        for (int k = 0; k < 6; ++k) {
                            if ((1 << k) < S) {
                                const double sib = seg_xchg(s, k);
                                lvl[64 * (2 * k)] = s;
                                lvl[64 * (2 * k + 1)] = sib;
                                s = s + sib;
                            }
                        }
                        const double R = s;
                        double dexp = INF;
                        int wexp = 0x7FFFFFFF;
                        if (R > 0.0) {
                            const smc_u32x4 rw = smc_draw(p.seed, run, step, SMC_P_RACE, 0);
                            dexp = -smc_log2(smc_u01(rw.v[0], rw.v[1])) / R;
                            double v = smc_u01(rw.v[2], rw.v[3]) * R;
                            int b = 0;
#pragma unroll
                            for (int k = 5; k >= 0; --k) {
                                if ((1 << k) < S) {
                                    const bool in_left = ((sl >> k) & 1) == 0;
                                    const double own = lvl[64 * (2 * k)], sib = lvl[64 * (2 * k + 1)];
                                    const double left = in_left ? own : sib;
                                    const double right = in_left ? sib : own;
                                    if (!(v < left || right == 0.0)) { v = v - left; b += 1 << k; }
                                }
                            }
                            // the lane's tree again, from the same leaves (same sums)
                            double tr[2 * C];
#pragma unroll
                            for (int j = 0; j < C; ++j) tr[C + j] = leaf[j];
#pragma unroll
                            for (int i = C - 1; i >= 1; --i) tr[i] = tr[2 * i] + tr[2 * i + 1];
                            int node = 1;
#pragma unroll
                            for (int lv = 0; (1 << lv) < C; ++lv) {
                                double left = 0.0, right = 0.0;
#pragma unroll
                                for (int cand = (1 << lv); cand < (2 << lv); ++cand)
                                    if (node == cand) { left = tr[2 * cand]; right = tr[2 * cand + 1]; }
                                if (v < left || right == 0.0) node = 2 * node;
                                else { v = v - left; node = 2 * node + 1; }
                            }
                            const int Lw = lowest_lane(__ballot(sl == b) & seg_mask);
                            wexp = (Lw - seg_base) * C + __shfl(node - C, Lw);
                        }
                        double gu = umin;
                        if (__ballot(unif_j >= 0) & seg_mask) {  // some slot of the run drew a delay
#pragma unroll
                            for (int k = 0; k < 6; ++k)
                                if ((1 << k) < S) gu = fmin(gu, seg_xchg(gu, k));
                        }
                        int wu = 0x7FFFFFFF;
                        if (gu < INF) {
                            const int Lu = lowest_lane(__ballot(unif_j >= 0 && umin == gu) & seg_mask);
                            wu = (Lu - seg_base) * C + __shfl(unif_j, Lu);
                        }
                        if (gu < dexp || (gu == dexp && wu < wexp)) { d = gu; w = wu; }
                        else if (dexp < INF) { d = dexp; w = wexp; }
                        else { d = INF; w = -1; }  // no racer: no event
                        // the step that crosses the bound: the left-out null
                        // racers' first event after the bound may come first
                        if (!full && d < INF && !(t + d < p.bound) && (__ballot(deadmask != 0) & seg_mask)) {
                            const double Rd = dead_total<C>(deadmask, loc, vars, m, S);
                            if (Rd > 0.0) {
                                const smc_u32x4 rw = smc_draw(p.seed, run, step, SMC_P_DEAD, 0);
                                const double td = p.bound + (-smc_log2(smc_u01(rw.v[0], rw.v[1])) / Rd);
                                if (td < t + d) { d = td - t; w = -1; }
                            }
                        }
                    }
                    if (!(d < INF)) { end = true; break; }  // no event: the run ends at its clock
                    // clocks advance (find_minimum_delay :632-651): one store of
                    // the run's clock time
                    if (sl == 0) vars[-1] = t + d;
                    wave_sync();
                    if (w < 0) break;  // a null event of the left-out racers
                    // the winner's transition (check_enabled_edges :268-305,
                    // take_transition :27-181), once, on the lane holding it
                    const int wl = seg_base + w / C;
                    int e_fired = -1, new_chan = 0;
                    if (lane == wl) {
                        const int wj = w % C;
                        const int lj = slot_get<C>(loc, wj);
                        const uint64_t mkw = enabled_mask(lj, 0, vars, m);
                        if (mkw) {
                            const int first = (int)m.locs[lj].edge_first;
                            const int i = pick_edge(mkw, first, p.seed, run, step, SMC_P_WINNER, (uint32_t)w, vars, m);
                            // the network's edge index (digests name it; records are shared)
                            e_fired = TRACE ? m.loc_edge0[lj] + i : first + i;
                            slot_set<C>(loc, wj, fire_edge(first + i, lj, vars, m, q, goal_local, new_chan));
                        }
                    }
                    wave_sync();
                    e_fired = __shfl(e_fired, wl);
                    if (e_fired >= 0) {
                        ++trans;
                        changed = true;
                        if (TRACE)
                            digest = smc_mix(digest, ((uint64_t)step << 40) | ((uint64_t)w << 20) | (uint64_t)e_fired);
                        chan = __shfl(new_chan, wl);
                        sender = chan > 0 ? w : -1;
                    }
                    goal = (__ballot(goal_local) & seg_mask) != 0;
                    break;
                }

                // phase 0: enabled receive edges of every component but the sender
                changed = true;
                uint64_t mk[C];
                bool any = false;
#pragma unroll
                for (int j = 0; j < C; ++j) {
                    const int c = sl * C + j;
                    mk[j] = (c < n && c != sender) ? enabled_mask(loc[j], chan, vars, m) : 0ull;
                    any |= mk[j] != 0;
                }
                // fire them: all at once when they commute, else one at a
                // time in component order
                if (!fire_receivers_fast<C>(mk, loc, vars, m, q, goal_local, p.seed, run, step, sl, S, seg_base,
                                            seg_mask, trans, TRACE, digest))
                    fire_receivers_serial<C>(mk, loc, any, vars, m, q, goal_local, p.seed, run, step, lane, seg_base,
                                             seg_mask, trans, TRACE, digest);
                chan = 0;
                sender = -1;
                goal = (__ballot(goal_local) & seg_mask) != 0;
            }
            if (!end) { t = t + d; ++step; }
        }

        if (end) {
            // per-run result (simulation.cu:887-893, statistics.cuh:126-163)
            double qmax = q.qmax, qmin = q.qmin;
#pragma unroll
            for (int k = 0; k < 6; ++k)
                if ((1 << k) < S) {
                    qmax = fmax(qmax, seg_xchg(qmax, k));
                    qmin = fmin(qmin, seg_xchg(qmin, k));
                }
            const double value = p.is_max ? qmax : qmin;
            if (TRACE) {
                digest = smc_mix(digest, smc_digest_bits(t));
                for (int v = 0; v < m.n_vars; ++v)
                    digest = smc_mix(digest, smc_digest_bits(m.var_clock[v] ? vars[-1] - vars[v] : vars[v]));
                for (int c = 0; c < n; ++c) {
                    const int lv = __shfl(slot_get<C>(loc, c % C), seg_base + c / C);
                    digest = smc_mix(digest, (uint64_t)(uint32_t)lv);
                }
            }
            if (sl == 0) {
                uint8_t flag;
                if (p.reach) flag = hit ? 1 : 0;
                else flag = p.is_max ? (value > p.threshold) : (value < p.threshold);
                p.flag[ridx] = flag;
                p.value[ridx] = value;
                p.steps[ridx] = step;
                p.trans[ridx] = trans;
                if (p.time) p.time[ridx] = t;
                if (TRACE && p.digest) p.digest[ridx] = digest;
            }
            wave_sync();
            active = false;
        }
    }
  }
}

// Two register policies of the run kernel: unrestricted (two waves per SIMD
// for every slot count: 150-256 VGPRs), and at most 168 (three waves per
// SIMD; up to four slots per lane fit, more spill).  Which is faster depends
// on the packing; the autotuner measures both.
template <int C, bool TRACE, int TAB>
__global__ __launch_bounds__(64 * SMC_MAX_GROUP_WAVES) void sim_kernel(LaunchParams lp) {
    sim_body<C, TRACE, TAB>(lp);
}

template <int C, bool TRACE, int TAB>
__global__ __launch_bounds__(64 * SMC_MAX_GROUP_WAVES) __attribute__((amdgpu_waves_per_eu(3))) void sim_kernel_o3(
    LaunchParams lp) {
    sim_body<C, TRACE, TAB>(lp);
}

}  // namespace
