// is_advance.h — the tilted advance kernel of importance sampling on gfx950
// (compiled once per slot count and register policy by is_variant.hip) and
// the bf16 helpers of the MLP kernels.
//
// No counterpart in the reference (cuSMC estimates by plain Monte Carlo,
// statistics.cuh:126-163).  A proposal multiplies the rate of each
// exponential component by 2^theta, with theta per location given by a small
// MLP of the run's state; each run carries log2 of its likelihood ratio in
// FP64, so that hit * 2^log2lr is an unbiased estimate of the probability.
//
// Runs advance in lock step by macro-steps: one launch of is_advance_kernel
// moves every active run to its next transition (the race, clocks and the
// winner's edge exactly as the run kernel does them, with tilted rates), then
// writes the run's features for the MLP, whose forward pass (MFMA GEMMs
// below) produces the tilts of the next macro-step.
//
// Cross-entropy training replays the elite runs once with the tilts that
// generated them and records, per (run, macro-step), the features at the
// macro-step's start and two statistics per location L of the run's
// components: A_L, the number of races won by the component in L, and B_L,
// its base cumulative hazard over the macro-step's races (r ln2 d per race
// for an exponential racer, -ln(1 - d/M) for a uniform racer on the window
// [0, M]).  The log-density of the macro-step under tilts theta is then
//   ln q = sum_L (ln2 theta_L A_L - 2^theta_L B_L) + terms free of theta,
// so every training epoch is a pass of GEMMs over the recorded pairs.
//
// The proposal: an exponential component's rate r becomes r' = r 2^theta;
// a uniform (clock-bounded) component's hazard 1/(M - d) is multiplied by
// a = 2^theta (a delay-window tilt: survival (1 - d/M)^a on the same window,
// drawn as M (1 - (1 - u)^(1/a))).  Likelihood ratio of one race step
// (exponential base rates r_c summing to R, tilted to R'; delays drawn as
// -log2(u)/rate, so the effective rates carry a factor ln 2):
//   log2 LR = -theta_w + (R' - R) d + sum_{uniform c} (1 - a_c) log2(1 - d/M_c)
// where theta_w is present when an exponential or uniform racer won with its
// own draw.  Zero delays are not tilted and contribute nothing.
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.h"
#include "is_kernels.hpp"

namespace {

__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)((u >> 16) | ((u & 0xFFFF) ? 0x40 : 0));
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Features of a run: one-hot of each component's location, scaled
// variables, time / bound.  Each entry is written by exactly one lane.
template <int C>
__device__ __forceinline__ void write_features(const IsParams &P, uint16_t *f, const int (&loc)[C], const double *vars,
                                               double t, int sl, int S, int n) {
#pragma unroll
    for (int j = 0; j < C; ++j) {
        const int c = sl * C + j;
        if (c < n) {
            for (int l = P.m.comp_first_loc[c]; l < P.m.comp_first_loc[c + 1]; ++l)
                f[l] = loc[j] == l ? (uint16_t)0x3F80 : (uint16_t)0;
        }
    }
    const int base = P.m.n_locs;
    for (int v = sl; v < P.m.n_vars; v += S) {
        const double x = P.m.var_clock[v] ? vars[-1] - vars[v] : vars[v];
        f[base + v] = f32_to_bf16((float)(x * P.var_scale[v]));
    }
    if (sl == 0) {
        f[base + P.m.n_vars] = f32_to_bf16((float)(t / P.bound));
        for (int k = base + P.m.n_vars + 1; k < P.fpad; ++k) f[k] = 0;
    }
}

// Whether a var-probability query is already decided: the segment-wide
// running max (min) has crossed the threshold.  Every lane tracks the
// updates it executed, so the test reduces over the segment.
__device__ __forceinline__ bool settled(const LaneQuery &q, const IsParams &P, int S) {
    double v = P.is_max ? q.qmax : q.qmin;
#pragma unroll
    for (int k = 0; k < 6; ++k)
        if ((1 << k) < S) v = P.is_max ? fmax(v, seg_xchg(v, k)) : fmin(v, seg_xchg(v, k));
    return P.is_max ? v > P.threshold : v < P.threshold;
}

// ------------------------------------------------------------ advance

// Tilt lookups of the workgroup and its waves (all in LDS): the tilt index
// of every location, and 2^theta precomputed once per theta entry — for the
// whole launch when every run shares one theta vector, else per row.
struct TiltLds {
    const int32_t *tix;   // [lpad]
    const double *ex;     // [theta_len]: 2^theta, shared theta only
    double *row_ex;       // the wave's rows: [64/S][theta_cap] 2^theta
    float *row_th;        // [64/S][theta_cap] theta
    double *lvl;          // the wave's race-tree levels: [2 log2(S)][64] (own, sibling per level)
    double *rec;          // recording build: [C][64] rates, [C][64] int32 locations, [C][64] uint32 wins
};

// REC: the build that records macro-step statistics (elite replays); the
// sampling build leaves the recording state out of its registers.
template <int C, bool REC, class MT>
__device__ __forceinline__ void is_advance_rows(const IsParams &P, const MT &m, double *wlds, const TiltLds &tl,
                                                int group) {
    const int lane = threadIdx.x & 63;
    const int S = P.lanes;
    const int sl = lane & (S - 1);
    const int seg_base = lane - sl;
    const uint64_t seg_mask = (S == 64) ? ~0ull : (((1ull << S) - 1ull) << seg_base);
    const int row = group * (64 / S) + lane / S;
    if (row >= P.n_rows) return;  // whole segments only
    double *vars = wlds + (size_t)(lane / S) * P.var_stride + 1;  // [clock time][variables]
    const IsRows &st = P.st;
    const int n = m.n_comp;
    const double INF = __builtin_huge_val();

    uint8_t stat = st.status[row];
    if (!(stat & IS_ACTIVE)) return;  // its features and tilts are not read

    // load the run
    const uint64_t run = st.run[row];
    double t = st.t[row];
    double log2lr = st.log2lr[row];
    uint32_t step = st.step[row], trans = st.trans[row];
    int chan = st.chan[row], sender = st.sender[row];
    bool goal = (stat & IS_GOAL) != 0, goal_local = false;
    bool mid = (stat & IS_MID) != 0;
    LaneQuery q{P.qvar, st.qmax[row], st.qmin[row]};
    uint64_t digest = st.digest[row];
    for (int v = sl; v < m.n_vars; v += S) vars[v] = st.vars[(size_t)row * m.n_vars + v];
    if (sl == 0) vars[-1] = t;
    int loc[C];
    // recording build: per slot the starting location, the base cumulative
    // hazard over the macro-step's races and the wins, in LDS ([slot][lane])
    double *hz = tl.rec + lane;
    int32_t *loc0 = (int32_t *)(tl.rec + 64 * C) + lane;
    uint32_t *wins = (uint32_t *)(tl.rec + 64 * C) + 64 * C + lane;
#pragma unroll
    for (int j = 0; j < C; ++j) {
        const int c = sl * C + j;
        loc[j] = c < n ? (int)st.loc[(size_t)row * n + c] : -1;
        if (REC) {
            loc0[64 * j] = loc[j];
            hz[64 * j] = 0.0;
            wins[64 * j] = 0;
        }
    }
    wave_sync();
    // this row's tilts: theta and 2^theta by theta column
    const float *theta = nullptr;
    const double *tex = nullptr;
    if (P.theta) {
        if (P.theta_stride == 0) {
            theta = P.theta;
            tex = tl.ex;
        } else {
            float *th = tl.row_th + (size_t)(lane / S) * P.theta_cap;
            double *ex = tl.row_ex + (size_t)(lane / S) * P.theta_cap;
            const float *src = P.theta + (size_t)row * P.theta_stride;
            for (int i = sl; i < P.theta_len; i += S) {
                const float v = src[i];
                th[i] = v;
                ex[i] = smc_exp2((double)v);
            }
            wave_sync();
            theta = th;
            tex = ex;
        }
    }
    const int rec = (REC && P.rec_pos) ? P.rec_pos[row] : -1;
    if (rec >= 0) {
        write_features<C>(P, P.rec_X + (size_t)rec * P.fpad, loc, vars, t, sl, S, n);
        if (sl == 0) P.rec_w[rec] = P.weight[st.job[row]];
    }
    bool hit = false, end = false, fired_any = false;
    bool changed = true;  // state changed since the last absorption check
    uint32_t blk = 0;     // blocked slots at the last change of state

    for (;;) {
        if (!mid) {
            // loop test (simulation.cu:835-849), with the early stop of a
            // var-probability query once its outcome is settled
            if (!(t < P.bound)) end = true;
            else if (P.reach && goal) { end = true; hit = true; }
            else if (P.early_stop && !P.reach && settled(q, P, S)) { end = true; hit = true; }
            else if (step >= P.max_steps) end = true;
            if (end) break;
            if (fired_any && P.stop_at_transition) {
                // process the pending broadcast now so the features show it;
                // the next launch resumes at the race (mid-iteration)
                if (chan > 0) {
                    uint64_t mk[C];
                    bool any = false;
#pragma unroll
                    for (int j = 0; j < C; ++j) {
                        const int c = sl * C + j;
                        mk[j] = (c < n && c != sender) ? enabled_mask(loc[j], chan, vars, m) : 0ull;
                        any |= mk[j] != 0;
                    }
                    if (!fire_receivers_fast<C>(mk, loc, vars, m, q, goal_local, P.seed, run, step, sl, S, seg_base,
                                                seg_mask, trans, P.trace != 0, digest))
                        fire_receivers_serial<C>(mk, loc, any, vars, m, q, goal_local, P.seed, run, step, lane,
                                                 seg_base, seg_mask, trans, P.trace != 0, digest);
                    chan = 0;
                    sender = -1;
                    goal = (__ballot(goal_local) & seg_mask) != 0;
                }
                mid = true;
                break;
            }
        }
        double d = 0.0;
        int w = -1;
#pragma clang loop unroll(disable)
        for (int phase = mid ? 1 : 0; phase < 2; ++phase) {
            if (phase == 0 && chan <= 0) continue;
            bool exp_won = false;
            if (phase == 1) {
                // a goal entered by receivers: this step is raced in full
                // and decides the hit (as in the run kernel)
                const bool full = P.reach && goal;
                // blocked bits: after a change of state, in the slot loop
                // below on models with frozen racers, here otherwise
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

                double leaf[C], leaft[C], Mb[C];
                uint32_t rref[C];  // rate expression of slots in rpend
                RateMemo memo;
                int zero_j = C;
                uint32_t bmask = 0, deadmask = 0, rpend = 0;
#pragma unroll
                for (int j = 0; j < C; ++j) {
                    leaf[j] = 0.0; leaft[j] = 0.0; Mb[j] = 0.0;
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
                {
                    // racing exponential slots: rates (expressions evaluated once per
                    // distinct expression), then their tilts
                    uint32_t racing = rpend;
                    pending_rates<C>(rpend, rref, leaf, vars, m);
#pragma unroll
                    for (int j = 0; j < C; ++j) {
                        if (leaf[j] > 0.0 || ((racing >> j) & 1u)) {
                            leaft[j] = leaf[j];
                            if (theta && leaf[j] > 0.0) {
                                const int ti = tl.tix[loc[j]];
                                if (ti >= 0) leaft[j] = leaf[j] * tex[ti];
                            }
                        }
                    }
                }
                // frozen components race only in a full race (as in the run
                // kernel)
                if (full) {
                    for (uint32_t b = deadmask; b; b &= b - 1) {
                        const int j = __ffs(b) - 1;
                        const int lj = slot_get<C>(loc, j);
                        const DevLoc L = load_loc(m.locs + lj);
                        const double r = location_rate(L, vars, m, memo);
                        const double lf = r > 0.0 ? r : 0.0;
                        double lt = lf;
                        if (theta && lf > 0.0) {
                            const int ti = tl.tix[lj];
                            if (ti >= 0) lt = lf * tex[ti];
                        }
                        slot_set<C>(leaf, j, lf);
                        slot_set<C>(leaft, j, lt);
                    }
                }
                const uint64_t zb = __ballot(zero_j < C) & seg_mask;
                if (zb) {
                    const int Lz = lowest_lane(zb);
                    w = (Lz - seg_base) * C + __shfl(zero_j, Lz);
                    d = 0.0;
                } else {
                    // uniform delays; a tilted window (hazard times a = 2^theta)
                    // draws M (1 - (1 - u)^(1/a))
                    double umin = INF;
                    int unif_j = -1;
                    for (uint32_t b = bmask; b;) {
                        const int j = __ffs(b) - 1;
                        const uint32_t c = (uint32_t)(sl * C + j);
                        const smc_u32x4 rw = smc_draw(P.seed, run, step, SMC_P_UNIFORM, smc_uniform_draw(c));
                        // slot j, then slot j + 1 when it shares the draw
                        const int nj = (C > 1 && !(c & 1u) && ((b >> (j + 1)) & 1u)) ? 2 : 1;
                        for (int k = 0; k < nj; ++k) {
                            const int jk = j + k;
                            const double u = smc_uniform_u(rw, c + (uint32_t)k);
                            const double M = slot_get<C>(Mb, jk);
                            const int ti = theta ? tl.tix[slot_get<C>(loc, jk)] : -1;
                            const bool tilted = ti >= 0 && theta[ti] != 0.0f;
                            const double du = tilted ? M * (1.0 - smc_exp2(smc_log2(1.0 - u) / tex[ti])) : M * u;
                            if (du < umin) { umin = du; unif_j = jk; }
                            b &= b - 1;
                        }
                    }
                    // tilted race: the run kernel's tree over the tilted rates;
                    // the butterfly's partial sums wait in LDS for the descent
                    double *lvl = tl.lvl + lane;
                    double s;
                    {
                        double tr[2 * C];
#pragma unroll
                        for (int j = 0; j < C; ++j) tr[C + j] = leaft[j];
#pragma unroll
                        for (int i = C - 1; i >= 1; --i) tr[i] = tr[2 * i] + tr[2 * i + 1];
                        s = (C == 1) ? leaft[0] : tr[1];
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
                    const double Rt = s;
                    const double R = theta ? tree_total<C>(leaf, S) : Rt;
                    double dexp = INF;
                    int wexp = 0x7FFFFFFF;
                    if (Rt > 0.0) {
                        const smc_u32x4 rw = smc_draw(P.seed, run, step, SMC_P_RACE, 0);
                        dexp = -smc_log2(smc_u01(rw.v[0], rw.v[1])) / Rt;
                        double v = smc_u01(rw.v[2], rw.v[3]) * Rt;
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
                        for (int j = 0; j < C; ++j) tr[C + j] = leaft[j];
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
                    bool u_won = false;
                    if (gu < dexp || (gu == dexp && wu < wexp)) { d = gu; w = wu; u_won = true; }
                    else if (dexp < INF) { d = dexp; w = wexp; exp_won = true; }
                    else { d = INF; w = -1; }  // no racer: no event
                    // the step that crosses the bound: the left-out null
                    // racers' first event after it (base rates, no tilt)
                    if (!full && d < INF && !(t + d < P.bound) && (__ballot(deadmask != 0) & seg_mask)) {
                        const double Rd = dead_total<C>(deadmask, loc, vars, m, S);
                        if (Rd > 0.0) {
                            const smc_u32x4 rw = smc_draw(P.seed, run, step, SMC_P_DEAD, 0);
                            const double td = P.bound + (-smc_log2(smc_u01(rw.v[0], rw.v[1])) / Rd);
                            if (td < t + d) { d = td - t; w = -1; exp_won = false; u_won = false; }
                        }
                    }
                    if (theta && d < INF) {
                        // log2 LR of this step: -theta_w (the winner's own
                        // draw) + (R' - R) d + the tilted uniform racers'
                        // (1 - a) log2(1 - d/M), summed on the race tree
                        double th_w = 0.0;
                        if (exp_won || u_won) {
                            const int wl_lane = seg_base + w / C;
                            const int wj = w % C;
                            double tw = 0.0;
                            if (lane == wl_lane) {
                                const int lw = slot_get<C>(loc, wj);
                                const int ti = tl.tix[lw];
                                if (ti >= 0) tw = (double)theta[ti];
                            }
                            th_w = __shfl(tw, wl_lane);
                        }
                        log2lr = log2lr - th_w + (Rt - R) * d;
                        if (__ballot(bmask != 0) & seg_mask) {
                            double ut[C];
#pragma unroll
                            for (int j = 0; j < C; ++j) {
                                ut[j] = 0.0;
                                if ((bmask >> j) & 1u) {
                                    const int ti = tl.tix[loc[j]];
                                    if (ti >= 0 && theta[ti] != 0.0f)
                                        ut[j] = (1.0 - tex[ti]) * smc_log2(1.0 - d / Mb[j]);
                                }
                            }
                            log2lr = log2lr + tree_total<C>(ut, S);
                        }
                    }
                    if ((exp_won || u_won) && rec >= 0) {
                        const int wj = w % C;
                        if (lane == seg_base + w / C) wins[64 * wj] += 1u;
                    }
                    if (REC && rec >= 0 && d < INF) {
                        // each racer's base cumulative hazard over this step
                        const double LN2 = 0.69314718055994530942;
#pragma unroll
                        for (int j = 0; j < C; ++j) {
                            if (leaf[j] > 0.0) hz[64 * j] = hz[64 * j] + LN2 * leaf[j] * d;
                            else if ((bmask >> j) & 1u) hz[64 * j] = hz[64 * j] + -LN2 * smc_log2(1.0 - d / Mb[j]);
                        }
                    }
                }
                if (!(d < INF)) { end = true; break; }  // no event: the run ends at its clock
                if (sl == 0) vars[-1] = t + d;  // clocks advance
                wave_sync();
                if (w < 0) break;  // a null event of the left-out racers
                // the winner's transition, once, on the lane holding it
                const int wl = seg_base + w / C;
                int e_fired = -1, new_chan = 0;
                if (lane == wl) {
                    const int wj = w % C;
                    const int lj = slot_get<C>(loc, wj);
                    const uint64_t mkw = enabled_mask(lj, 0, vars, m);
                    if (mkw) {
                        const int first = (int)m.locs[lj].edge_first;
                        const int i = pick_edge(mkw, first, P.seed, run, step, SMC_P_WINNER, (uint32_t)w, vars, m);
                        // the network's edge index (digests name it; records are shared)
                        e_fired = P.trace ? m.loc_edge0[lj] + i : first + i;
                        slot_set<C>(loc, wj, fire_edge(first + i, lj, vars, m, q, goal_local, new_chan));
                    }
                }
                wave_sync();
                e_fired = __shfl(e_fired, wl);
                if (e_fired >= 0) {
                    ++trans;
                    changed = true;
                    if (P.trace)
                        digest = smc_mix(digest, ((uint64_t)step << 40) | ((uint64_t)w << 20) | (uint64_t)e_fired);
                    chan = __shfl(new_chan, wl);
                    sender = chan > 0 ? w : -1;
                    fired_any = true;
                }
                goal = (__ballot(goal_local) & seg_mask) != 0;
                break;
            }
            // phase 0: pending broadcast receivers
            changed = true;
            uint64_t mk[C];
            bool any = false;
#pragma unroll
            for (int j = 0; j < C; ++j) {
                const int c = sl * C + j;
                mk[j] = (c < n && c != sender) ? enabled_mask(loc[j], chan, vars, m) : 0ull;
                any |= mk[j] != 0;
            }
            if (!fire_receivers_fast<C>(mk, loc, vars, m, q, goal_local, P.seed, run, step, sl, S, seg_base, seg_mask,
                                        trans, P.trace != 0, digest))
                fire_receivers_serial<C>(mk, loc, any, vars, m, q, goal_local, P.seed, run, step, lane, seg_base,
                                         seg_mask, trans, P.trace != 0, digest);
            chan = 0;
            sender = -1;
            goal = (__ballot(goal_local) & seg_mask) != 0;
        }
        mid = false;
        if (end) break;
        t = t + d;
        ++step;
    }

    // store the run
    const bool still = !end;
    if (end) {
        double qmax = q.qmax, qmin = q.qmin;
#pragma unroll
        for (int k = 0; k < 6; ++k)
            if ((1 << k) < S) { qmax = fmax(qmax, seg_xchg(qmax, k)); qmin = fmin(qmin, seg_xchg(qmin, k)); }
        q.qmax = qmax; q.qmin = qmin;
        if (!hit && !P.reach) hit = P.is_max ? (qmax > P.threshold) : (qmin < P.threshold);
        if (P.trace) {
            digest = smc_mix(digest, smc_digest_bits(t));
            for (int v = 0; v < m.n_vars; ++v)
                digest = smc_mix(digest, smc_digest_bits(m.var_clock[v] ? vars[-1] - vars[v] : vars[v]));
            for (int c = 0; c < n; ++c) {
                const int lv = __shfl(slot_get<C>(loc, c % C), seg_base + c / C);
                digest = smc_mix(digest, (uint64_t)(uint32_t)lv);
            }
        }
    } else {
        double qmax = q.qmax, qmin = q.qmin;
#pragma unroll
        for (int k = 0; k < 6; ++k)
            if ((1 << k) < S) { qmax = fmax(qmax, seg_xchg(qmax, k)); qmin = fmin(qmin, seg_xchg(qmin, k)); }
        q.qmax = qmax; q.qmin = qmin;
    }
#pragma unroll
    for (int j = 0; j < C; ++j) {
        const int c = sl * C + j;
        if (c < n) st.loc[(size_t)row * n + c] = (uint16_t)loc[j];
    }
    for (int v = sl; v < m.n_vars; v += S) st.vars[(size_t)row * m.n_vars + v] = vars[v];
    goal = (__ballot(goal_local) & seg_mask) != 0 || goal;
    if (sl == 0) {
        st.t[row] = t;
        st.log2lr[row] = log2lr;
        st.step[row] = step;
        st.trans[row] = trans;
        st.chan[row] = chan;
        st.sender[row] = sender;
        st.qmax[row] = q.qmax;
        st.qmin[row] = q.qmin;
        st.digest[row] = digest;
        st.status[row] = (uint8_t)((still ? IS_ACTIVE : 0) | (hit ? IS_HIT : 0) | (goal ? IS_GOAL : 0) |
                                   (still && mid ? IS_MID : 0));
        if (still && P.active_count) atomicAdd(P.active_count, 1u);
    }
    if (P.feat) write_features<C>(P, P.feat + (size_t)row * P.fpad, loc, vars, t, sl, S, n);
    if (rec >= 0) {
        // statistics of the macro-step for the locations its components
        // started in (a component changes location only at the macro-step's
        // end): wins and base cumulative hazard
        float *A = P.rec_A + (size_t)rec * P.lpad;
        float *B = P.rec_B + (size_t)rec * P.lpad;
#pragma unroll
        for (int j = 0; j < C; ++j) {
            const int c = sl * C + j;
            if (c < n) {
                for (int l = m.comp_first_loc[c]; l < m.comp_first_loc[c + 1]; ++l) {
                    float a = 0.f, b = 0.f;
                    if (l == loc0[64 * j] && P.tiltable[l]) {
                        a = (float)wins[64 * j];
                        b = (float)hz[64 * j];
                    }
                    A[l] = a;
                    B[l] = b;
                }
            }
        }
        if (sl == 0)
            for (int k = m.n_locs; k < P.lpad; ++k) { A[k] = 0.f; B[k] = 0.f; }
    }
}

// A workgroup holds blockDim/64 waves; each takes groups of 64/S rows
// (blockIdx.x * waves + wave, then strided by the grid).  With TAB the
// workgroup first stages the model's hot tables in LDS (as the run kernel).
// LDS: [tables: TAB only][per wave: variables of its rows].
template <int C, bool TAB, bool REC>
__device__ __forceinline__ void is_advance_body(const IsParams &P) {
    extern __shared__ __attribute__((aligned(16))) double lds[];
    const int wv = threadIdx.x >> 6, W = blockDim.x >> 6;
    // LDS: [tables: TAB only][tilt index: lpad int32][2^theta: lpad doubles]
    //      then per wave [variables][rows' 2^theta][rows' theta][race-tree
    //      levels: 2 log2(S) doubles per lane][recording build: 16 C bytes
    //      per lane]
    double *shared = lds + (TAB ? P.table_bytes / 8 : 0);
    int32_t *tix = (int32_t *)shared;
    double *ex = shared + P.lpad / 2;
    double *wlds = shared + (size_t)P.lpad * 3 / 2 + (size_t)wv * (P.wave_lds / 8);
    const int R = 64 / P.lanes;
    TiltLds tl;
    tl.tix = tix;
    tl.ex = ex;
    tl.row_ex = wlds + (size_t)R * P.var_stride;
    tl.row_th = (float *)(tl.row_ex + (size_t)R * P.theta_cap);
    tl.lvl = (double *)(tl.row_th + (size_t)R * P.theta_cap);
    int levels = 0;
    while ((1 << levels) < P.lanes) ++levels;
    tl.rec = tl.lvl + (size_t)2 * levels * 64;
    for (int i = threadIdx.x; i < P.lpad; i += blockDim.x) tix[i] = P.tilt_index[i];
    if (P.theta && P.theta_stride == 0)
        for (int i = threadIdx.x; i < P.theta_len; i += blockDim.x) ex[i] = smc_exp2((double)P.theta[i]);
    const int groups = (P.n_rows + R - 1) / R;
    if constexpr (TAB) {
        stage_tables(P.tables, P.table_bytes, lds);  // ends with the workgroup barrier
        const DevModelT<SmcLdsTables> m = lds_model(P.m, P.tables, lds);
        for (int g = blockIdx.x * W + wv; g < groups; g += gridDim.x * W) {
            is_advance_rows<C, REC>(P, m, wlds, tl, g);
            wave_sync();
        }
    } else {
        __syncthreads();
        for (int g = blockIdx.x * W + wv; g < groups; g += gridDim.x * W) {
            is_advance_rows<C, REC>(P, P.m, wlds, tl, g);
            wave_sync();
        }
    }
}

// Registers unrestricted: 190-250 VGPRs, two waves per SIMD.
template <int C, bool TAB, bool REC>
__global__ __launch_bounds__(256) void is_advance_kernel(IsParams P) {
    is_advance_body<C, TAB, REC>(P);
}

}  // namespace
