// is_kernels.hip — importance sampling for rare-event queries on gfx950.
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
// components: A_L, the number of races won by the component in L, and
// B_L = ln2 * dt * r_L, its base rate times the macro-step's duration.  The
// log-density of the macro-step under tilts theta is then
//   ln q = sum_L ln2 (A_L theta_L - B_L 2^theta_L) + terms free of theta,
// so every training epoch is a pass of GEMMs over the recorded pairs.
//
// Likelihood ratio of one race step (exponential race with base rates r_c,
// tilted rates r'_c = r_c 2^theta_c, sums R and R', delay d, delays drawn
// as -log2(u)/rate so the effective rates carry a factor ln 2):
//   LR = (r_w / r'_w) exp(-(R - R') ln2 d)   =>   log2 LR = -theta_w + (R' - R) d
// where the first term is present only when an exponential component won.
// Uniform and zero delays are not tilted and contribute nothing.
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
    for (int v = sl; v < P.m.n_vars; v += S) f[base + v] = f32_to_bf16((float)(vars[v] * P.var_scale[v]));
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

template <int C>
__global__ __launch_bounds__(64) void is_advance_kernel(IsParams P) {
    extern __shared__ double lds[];
    const int lane = threadIdx.x;
    const int S = P.lanes;
    const int sl = lane & (S - 1);
    const int seg_base = lane - sl;
    const uint64_t seg_mask = (S == 64) ? ~0ull : (((1ull << S) - 1ull) << seg_base);
    const int row = blockIdx.x * (64 / S) + lane / S;
    if (row >= P.n_rows) return;  // whole segments only
    double *vars = lds + (size_t)(lane / S) * P.var_stride;
    const DevModel &m = P.m;
    const IsRows &st = P.st;
    const int n = m.n_comp;
    const double INF = __builtin_huge_val();
    const double LN2 = 0.69314718055994530942;

    uint8_t stat = st.status[row];
    if (!(stat & IS_ACTIVE)) {
        if (P.feat) {
            uint16_t *f = P.feat + (size_t)row * P.fpad;
            for (int k = sl; k < P.fpad; k += S) f[k] = 0;
        }
        return;
    }

    // load the run
    const uint64_t run = st.run[row];
    double t = st.t[row];
    const double t_start = t;
    double log2lr = st.log2lr[row];
    uint32_t step = st.step[row], trans = st.trans[row];
    int chan = st.chan[row], sender = st.sender[row];
    bool goal = (stat & IS_GOAL) != 0, goal_local = false;
    bool mid = (stat & IS_MID) != 0;
    LaneQuery q{P.qvar, st.qmax[row], st.qmin[row]};
    uint64_t digest = st.digest[row];
    for (int v = sl; v < m.n_vars; v += S) vars[v] = st.vars[(size_t)row * m.n_vars + v];
    int loc[C], loc0[C];
    double r0[C];
    uint32_t wins[C];
#pragma unroll
    for (int j = 0; j < C; ++j) {
        const int c = sl * C + j;
        loc[j] = c < n ? (int)st.loc[(size_t)row * n + c] : -1;
        loc0[j] = loc[j];
        r0[j] = 0.0;
        wins[j] = 0;
    }
    wave_sync();
    const float *theta = P.theta ? P.theta + (size_t)row * P.theta_stride : nullptr;
    const int rec = P.rec_pos ? P.rec_pos[row] : -1;
    if (rec >= 0) {
        write_features<C>(P, P.rec_X + (size_t)rec * P.fpad, loc, vars, t, sl, S, n);
        if (sl == 0) P.rec_w[rec] = P.weight[row];
    }
    bool first_race = true;
    bool hit = false, end = false, fired_any = false;

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
                    int dummy = 0;
                    for (uint64_t lw = __ballot(any) & seg_mask; lw; lw &= lw - 1) {
                        const int L = lowest_lane(lw);
#pragma clang loop unroll(disable)
                        for (int j = 0; j < C; ++j) {
                            if (!((__ballot(slot_get<C>(mk, j) != 0) >> L) & 1ull)) continue;
                            const int c = (L - seg_base) * C + j;
                            int e_fired = -1;
                            if (lane == L) {
                                const int lj = slot_get<C>(loc, j);
                                const int first = (int)m.locs[lj].edge_first;
                                const int i = pick_edge(slot_get<C>(mk, j), first, P.seed, run, step, SMC_P_RECEIVER,
                                                        (uint32_t)c, vars, m);
                                e_fired = first + i;
                                slot_set<C>(loc, j, fire_edge(e_fired, vars, m, q, goal_local, dummy));
                            }
                            wave_sync();
                            ++trans;
                            const int e = __shfl(e_fired, L);
                            digest = smc_mix(digest, ((uint64_t)step << 40) | ((uint64_t)c << 20) | (uint64_t)e);
                        }
                    }
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
                double leaf[C], leaft[C], Mb[C];
                int zero_j = C;
                uint32_t bmask = 0;
#pragma unroll
                for (int j = 0; j < C; ++j) {
                    leaf[j] = 0.0; leaft[j] = 0.0; Mb[j] = 0.0;
                    const int c = sl * C + j;
                    if (c < n) {
                        const DevLoc L = m.locs[loc[j]];
                        if (L.klass == SMC_K_ZERO) {
                            if (zero_j == C) zero_j = j;
                        } else {
                            bool bounded = false;
                            double M = DBL_MAX;
                            for (int i = 0; i < L.n_inv; ++i) {
                                const DevInv iv = m.invs[L.inv_first + i];
                                const double bound = iv.bound_const ? iv.bound_value : eval_ref(iv.bound, vars, m);
                                double ttb = bound - vars[iv.var];
                                if (iv.strict) ttb -= 1e-6;
                                if (ttb >= 0.0) { M = ttb < M ? ttb : M; bounded = true; }
                            }
                            if (bounded) {
                                if (M == 0.0) { if (zero_j == C) zero_j = j; }
                                else { Mb[j] = M; bmask |= 1u << j; }
                            } else {
                                const double r = L.rate_const ? L.rate_value : eval_ref(L.rate, vars, m);
                                leaf[j] = r > 0.0 ? r : 0.0;
                                leaft[j] = leaf[j];
                                if (theta && leaf[j] > 0.0 && P.tiltable[loc[j]])
                                    leaft[j] = leaf[j] * smc_exp2((double)theta[loc[j]]);
                            }
                        }
                    }
                }
                if (first_race) {
#pragma unroll
                    for (int j = 0; j < C; ++j) r0[j] = leaf[j];
                    first_race = false;
                }
                const uint64_t zb = __ballot(zero_j < C) & seg_mask;
                if (zb) {
                    const int Lz = lowest_lane(zb);
                    w = (Lz - seg_base) * C + __shfl(zero_j, Lz);
                    d = 0.0;
                } else {
                    double umin = INF;
                    int unif_j = -1;
                    for (uint32_t b = bmask; b; b &= b - 1) {
                        const int j = __ffs(b) - 1;
                        const smc_u32x4 rw = smc_draw(P.seed, run, step, SMC_P_UNIFORM, (uint32_t)(sl * C + j));
                        const double du = slot_get<C>(Mb, j) * smc_u01(rw.v[0], rw.v[1]);
                        if (du < umin) { umin = du; unif_j = j; }
                    }
                    // tilted race: the run kernel's tree over the tilted rates
                    double tr[2 * C];
#pragma unroll
                    for (int j = 0; j < C; ++j) tr[C + j] = leaft[j];
#pragma unroll
                    for (int i = C - 1; i >= 1; --i) tr[i] = tr[2 * i] + tr[2 * i + 1];
                    double s = (C == 1) ? leaft[0] : tr[1];
                    double own[6], sib[6];
#pragma unroll
                    for (int k = 0; k < 6; ++k) {
                        own[k] = s;
                        sib[k] = 0.0;
                        if ((1 << k) < S) { sib[k] = seg_xchg(s, k); s = s + sib[k]; }
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
                                const double left = in_left ? own[k] : sib[k];
                                const double right = in_left ? sib[k] : own[k];
                                if (!(v < left || right == 0.0)) { v = v - left; b += 1 << k; }
                            }
                        }
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
#pragma unroll
                    for (int k = 0; k < 6; ++k)
                        if ((1 << k) < S) gu = fmin(gu, seg_xchg(gu, k));
                    int wu = 0x7FFFFFFF;
                    if (gu < INF) {
                        const int Lu = lowest_lane(__ballot(unif_j >= 0 && umin == gu) & seg_mask);
                        wu = (Lu - seg_base) * C + __shfl(unif_j, Lu);
                    }
                    if (gu < dexp || (gu == dexp && wu < wexp)) { d = gu; w = wu; }
                    else { d = dexp; w = wexp; exp_won = true; }
                    if (theta && d < INF) {
                        // log2 LR of this step: -theta_w + (R' - R) d
                        double th_w = 0.0;
                        if (exp_won) {
                            const int wl_lane = seg_base + w / C;
                            const int wj = w % C;
                            double tw = 0.0;
                            if (lane == wl_lane) {
                                const int lw = slot_get<C>(loc, wj);
                                if (P.tiltable[lw]) tw = (double)theta[lw];
                            }
                            th_w = __shfl(tw, wl_lane);
                        }
                        log2lr = log2lr - th_w + (Rt - R) * d;
                    }
                    if (exp_won && rec >= 0) {
                        const int wj = w % C;
                        if (lane == seg_base + w / C) {
#pragma unroll
                            for (int j = 0; j < C; ++j) wins[j] += (j == wj) ? 1u : 0u;
                        }
                    }
                }
                if (!(d < INF)) { t = INF; end = true; break; }
                for (int i = sl; i < m.n_clocks; i += S) vars[m.clocks[i]] += d;
                wave_sync();
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
                        e_fired = first + i;
                        slot_set<C>(loc, wj, fire_edge(e_fired, vars, m, q, goal_local, new_chan));
                    }
                }
                wave_sync();
                e_fired = __shfl(e_fired, wl);
                if (e_fired >= 0) {
                    ++trans;
                    digest = smc_mix(digest, ((uint64_t)step << 40) | ((uint64_t)w << 20) | (uint64_t)e_fired);
                    chan = __shfl(new_chan, wl);
                    sender = chan > 0 ? w : -1;
                    fired_any = true;
                }
                goal = (__ballot(goal_local) & seg_mask) != 0;
                break;
            }
            // phase 0: pending broadcast receivers
            uint64_t mk[C];
            bool any = false;
#pragma unroll
            for (int j = 0; j < C; ++j) {
                const int c = sl * C + j;
                mk[j] = (c < n && c != sender) ? enabled_mask(loc[j], chan, vars, m) : 0ull;
                any |= mk[j] != 0;
            }
            int dummy_chan = 0;
            for (uint64_t lanes_with = __ballot(any) & seg_mask; lanes_with; lanes_with &= lanes_with - 1) {
                const int L = lowest_lane(lanes_with);
#pragma clang loop unroll(disable)
                for (int j = 0; j < C; ++j) {
                    if (!((__ballot(slot_get<C>(mk, j) != 0) >> L) & 1ull)) continue;
                    const int c = (L - seg_base) * C + j;
                    int e_fired = -1;
                    if (lane == L) {
                        const int lj = slot_get<C>(loc, j);
                        const int first = (int)m.locs[lj].edge_first;
                        const int i = pick_edge(slot_get<C>(mk, j), first, P.seed, run, step, SMC_P_RECEIVER,
                                                (uint32_t)c, vars, m);
                        e_fired = first + i;
                        slot_set<C>(loc, j, fire_edge(e_fired, vars, m, q, goal_local, dummy_chan));
                    }
                    wave_sync();
                    ++trans;
                    const int e = __shfl(e_fired, L);
                    digest = smc_mix(digest, ((uint64_t)step << 40) | ((uint64_t)c << 20) | (uint64_t)e);
                }
            }
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
        digest = smc_mix(digest, smc_digest_bits(t));
        for (int v = 0; v < m.n_vars; ++v) digest = smc_mix(digest, smc_digest_bits(vars[v]));
        for (int c = 0; c < n; ++c) {
            const int lv = __shfl(slot_get<C>(loc, c % C), seg_base + c / C);
            digest = smc_mix(digest, (uint64_t)(uint32_t)lv);
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
        // started in (rates are constant within a macro-step)
        float *A = P.rec_A + (size_t)rec * P.lpad;
        float *B = P.rec_B + (size_t)rec * P.lpad;
        const double dt = (t < INF ? t : P.bound) - t_start;
#pragma unroll
        for (int j = 0; j < C; ++j) {
            const int c = sl * C + j;
            if (c < n) {
                for (int l = m.comp_first_loc[c]; l < m.comp_first_loc[c + 1]; ++l) {
                    float a = 0.f, b = 0.f;
                    if (l == loc0[j] && P.tiltable[l]) {
                        a = (float)wins[j];
                        b = (float)(LN2 * dt * r0[j]);
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

// Initial state of a row (simulation.cu:780-800: initial locations and
// values, time 0), one workgroup of 64 threads per row.
__global__ __launch_bounds__(64) void is_init_kernel(IsParams P, const uint64_t *runs, uint64_t run_begin,
                                                     int n_valid, const uint16_t *feat0) {
    const int row = blockIdx.x;
    const int tid = threadIdx.x;
    const IsRows &st = P.st;
    const DevModel &m = P.m;
    const bool valid = row < n_valid;
    for (int c = tid; c < m.n_comp; c += 64) st.loc[(size_t)row * m.n_comp + c] = (uint16_t)m.comp_init_loc[c];
    for (int v = tid; v < m.n_vars; v += 64) st.vars[(size_t)row * m.n_vars + v] = m.var_init[v];
    if (P.feat)
        for (int k = tid; k < P.fpad; k += 64) P.feat[(size_t)row * P.fpad + k] = valid ? feat0[k] : (uint16_t)0;
    if (tid == 0) {
        const uint64_t run = valid ? (runs ? runs[row] : run_begin + (uint64_t)row) : 0;
        st.run[row] = run;
        st.t[row] = 0.0;
        st.log2lr[row] = 0.0;
        st.qmax[row] = P.qinit;
        st.qmin[row] = P.qinit;
        st.step[row] = 0;
        st.trans[row] = 0;
        st.chan[row] = 0;
        st.sender[row] = -1;
        st.digest[row] = smc_mix(0, run);
        st.status[row] = valid ? (uint8_t)IS_ACTIVE : (uint8_t)0;
    }
}

// Record slots: an exclusive scan of the active flags in row order, offset
// by *base; one workgroup of 1024 threads, each owning a contiguous block of
// rows.  Slots at or beyond cap are not recorded.
__global__ __launch_bounds__(1024) void record_slots_kernel(const uint8_t *status, int n_rows, int32_t *pos,
                                                            unsigned int *base, unsigned int cap) {
    __shared__ unsigned int part[1024];
    const int tid = threadIdx.x;
    const int per = (n_rows + 1023) / 1024;
    const int r0 = tid * per, r1 = min(n_rows, r0 + per);
    unsigned int cnt = 0;
    for (int r = r0; r < r1; ++r) cnt += (status[r] & IS_ACTIVE) ? 1u : 0u;
    part[tid] = cnt;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // Hillis-Steele inclusive scan
        const unsigned int v = tid >= off ? part[tid - off] : 0u;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    const unsigned int b0 = *base;
    unsigned int at = b0 + part[tid] - cnt;
    for (int r = r0; r < r1; ++r) {
        int p = -1;
        if (status[r] & IS_ACTIVE) {
            if (at < cap) p = (int)at;
            ++at;
        }
        pos[r] = p;
    }
    __syncthreads();
    if (tid == 1023) *base = b0 + part[1023];
}

// d(-sum w ln q)/d theta at the MLP outputs of the recorded pairs:
//   G_L = -w ln2 (A_L - B_L 2^theta_L)
// and 0 where theta is clamped or the location is not tiltable.
__global__ void ce_grad_kernel(const float *theta, const float *A, const float *B, const float *w,
                               const uint8_t *tiltable, float tilt_max, int lpad, size_t n, float *G) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int l = (int)(i % (size_t)lpad);
    const size_t row = i / (size_t)lpad;
    const float th = theta[i];
    float g = 0.f;
    if (tiltable[l] && fabsf(th) < tilt_max) {
        const double LN2 = 0.69314718055994530942;
        g = (float)(-(double)w[row] * LN2 * ((double)A[i] - (double)B[i] * smc_exp2((double)th)));
    }
    G[i] = g;
}

// theta per location from the outputs of the location's tilt group.
__global__ void expand_groups_kernel(const float *tg, int gpad, const int32_t *group, int n_locs, int lpad,
                                     size_t n, float *tl) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int l = (int)(i % (size_t)lpad);
    const size_t row = i / (size_t)lpad;
    tl[i] = l < n_locs ? tg[row * gpad + group[l]] : 0.f;
}

// Gradient of a group's output: the sum over its member locations, in the
// fixed member order (CSR member_first / members).
__global__ void group_sum_kernel(const float *gl, int lpad, const int32_t *member_first, const int32_t *members,
                                 int n_groups, int gpad, size_t n, float *gg, uint16_t *ggb) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int gidx = (int)(i % (size_t)gpad);
    const size_t row = i / (size_t)gpad;
    float s = 0.f;
    if (gidx < n_groups)
        for (int k = member_first[gidx]; k < member_first[gidx + 1]; ++k) s += gl[row * lpad + members[k]];
    gg[i] = s;
    ggb[i] = f32_to_bf16(s);
}

// Column sums over a slice of rows (blockIdx.y), 64 columns per workgroup.
__global__ __launch_bounds__(64) void colsum_kernel(const float *X, int rows, int ld, int N, int rows_per,
                                                    float *part) {
    const int col = blockIdx.x * 64 + threadIdx.x;
    if (col >= N) return;
    const int r0 = blockIdx.y * rows_per, r1 = min(rows, r0 + rows_per);
    float s = 0.f;
    for (int r = r0; r < r1; ++r) s += X[(size_t)r * ld + col];
    part[(size_t)blockIdx.y * N + col] = s;
}

// dst[i] (+)= sum_z part[z][i], z in order.
__global__ void fold_splits_kernel(const float *part, int splits, size_t n, float *dst, int accumulate) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int z = 0; z < splits; ++z) s += part[(size_t)z * n + i];
    dst[i] = accumulate ? dst[i] + s : s;
}

// ------------------------------------------------------------ GEMM
//
// C[M x N] (+)= op(A)[M x K] . op(B)[K x N], bf16 inputs, fp32 accumulation
// on MFMA 16x16x16.  A workgroup of four waves computes a 64 x 64 tile; each
// wave a 32 x 32 quarter as 2 x 2 MFMA blocks; K advances 16 at a time
// through LDS.  Operand layout of v_mfma_f32_16x16x16_bf16: lane l holds
// A[l % 16][4 (l / 16) .. +3] and B[4 (l / 16) .. +3][l % 16]; the result
// lane l holds rows 4 (l / 16) .. +3 of column l % 16.  M, N are multiples
// of 64 and K of 16 (the host pads every buffer).  Each output element sums
// K in a fixed order, so a row's result does not depend on the other rows.

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g) {
    __shared__ uint16_t As[64][16 + 4];
    __shared__ uint16_t Bs[16][64 + 4];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, l = tid & 63;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int kz = (int)blockIdx.z * g.kchunk;
    const int kend = min(g.K, kz + g.kchunk);
    float *const Cout = g.C ? g.C + (size_t)blockIdx.z * g.zstride : nullptr;
    for (int k0 = kz; k0 < kend; k0 += 16) {
        // stage A (64 x 16) and B (16 x 64): 1024 elements each, 4 per thread
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int idx = tid * 4 + e;
            if (TA) {
                const int i = idx & 63, k = idx >> 6;       // coalesced along M
                As[i][k] = g.A[(size_t)(k0 + k) * g.lda + (m0 + i)];
            } else {
                const int k = idx & 15, i = idx >> 4;       // coalesced along K
                As[i][k] = g.A[(size_t)(m0 + i) * g.lda + (k0 + k)];
            }
            if (TB) {
                const int k = idx & 15, j = idx >> 4;       // coalesced along K
                Bs[k][j] = g.B[(size_t)(n0 + j) * g.ldb + (k0 + k)];
            } else {
                const int j = idx & 63, k = idx >> 6;       // coalesced along N
                Bs[k][j] = g.B[(size_t)(k0 + k) * g.ldb + (n0 + j)];
            }
        }
        __syncthreads();
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            bf16x4 fa;
            const int row = wr + a * 16 + (l & 15), kb = 4 * (l >> 4);
            fa[0] = (short)As[row][kb]; fa[1] = (short)As[row][kb + 1];
            fa[2] = (short)As[row][kb + 2]; fa[3] = (short)As[row][kb + 3];
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                bf16x4 fb;
                const int col = wc + b * 16 + (l & 15);
                fb[0] = (short)Bs[kb][col]; fb[1] = (short)Bs[kb + 1][col];
                fb[2] = (short)Bs[kb + 2][col]; fb[3] = (short)Bs[kb + 3][col];
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fa, fb, acc[a][b], 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = m0 + wr + a * 16 + 4 * (l >> 4) + i;
                const int col = n0 + wc + b * 16 + (l & 15);
                float v = acc[a][b][i];
                const size_t o = (size_t)row * g.ldc + col;
                if (g.accumulate) v += Cout[o];
                if (g.bias) v += g.bias[col];
                if (g.relu) v = v > 0.f ? v : 0.f;
                if (g.mask) v = g.mask[(size_t)row * g.ldm + col] != 0 ? v : 0.f;  // relu backward (bf16 bits)
                if (g.clamp_max > 0.f) {
                    v = v > g.clamp_max ? g.clamp_max : (v < -g.clamp_max ? -g.clamp_max : v);
                    if (g.col_mask && !g.col_mask[col]) v = 0.f;
                }
                if (Cout) Cout[o] = v;
                if (g.Cb) g.Cb[o] = f32_to_bf16(v);
            }
}

__global__ void to_bf16_kernel(const float *x, uint16_t *y, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f32_to_bf16(x[i]);
}

// Adam on a flat parameter vector (gradient descent on the loss).
__global__ void adam_kernel(float *w, const float *gsum, float *m1, float *m2, size_t n, float lr, float b1, float b2,
                            float eps, float bc1, float bc2, uint16_t *wb) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gr = gsum[i];
    const float a = b1 * m1[i] + (1.f - b1) * gr;
    const float b = b2 * m2[i] + (1.f - b2) * gr * gr;
    m1[i] = a;
    m2[i] = b;
    const float upd = lr * (a / bc1) / (sqrtf(b / bc2) + eps);
    const float nw = w[i] - upd;
    w[i] = nw;
    if (wb) wb[i] = f32_to_bf16(nw);
}

}  // namespace

namespace smc {

hipError_t is_launch_advance(const IsParams &p, int slots, size_t lds, hipStream_t s) {
    const int R = 64 / p.lanes;
    const int grid = (p.n_rows + R - 1) / R;
    if (grid <= 0) return hipSuccess;
    switch (slots) {
    case 1: hipLaunchKernelGGL(is_advance_kernel<1>, dim3(grid), dim3(64), lds, s, p); break;
    case 2: hipLaunchKernelGGL(is_advance_kernel<2>, dim3(grid), dim3(64), lds, s, p); break;
    case 4: hipLaunchKernelGGL(is_advance_kernel<4>, dim3(grid), dim3(64), lds, s, p); break;
    case 8: hipLaunchKernelGGL(is_advance_kernel<8>, dim3(grid), dim3(64), lds, s, p); break;
    case 16: hipLaunchKernelGGL(is_advance_kernel<16>, dim3(grid), dim3(64), lds, s, p); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

static hipError_t gemm_launch(const GemmArgs &g, bool ta, bool tb, int zsplit, hipStream_t s) {
    dim3 grid(g.N / 64, g.M / 64, zsplit);
    if (!ta && !tb) hipLaunchKernelGGL((gemm_kernel<false, false>), grid, dim3(256), 0, s, g);
    else if (ta && !tb) hipLaunchKernelGGL((gemm_kernel<true, false>), grid, dim3(256), 0, s, g);
    else if (!ta && tb) hipLaunchKernelGGL((gemm_kernel<false, true>), grid, dim3(256), 0, s, g);
    else hipLaunchKernelGGL((gemm_kernel<true, true>), grid, dim3(256), 0, s, g);
    return hipGetLastError();
}

hipError_t gemm(const GemmArgs &g0, bool ta, bool tb, hipStream_t s) {
    if (g0.M % 64 || g0.N % 64 || g0.K % 16 || g0.M <= 0 || g0.N <= 0 || g0.K <= 0) return hipErrorInvalidValue;
    GemmArgs g = g0;
    g.kchunk = g.K;
    g.zstride = 0;
    return gemm_launch(g, ta, tb, 1, s);
}

hipError_t gemm_splitk(const GemmArgs &g0, bool ta, bool tb, int splits, float *work, hipStream_t s) {
    if (g0.M % 64 || g0.N % 64 || g0.K % 16 || g0.M <= 0 || g0.N <= 0 || g0.K <= 0 || !g0.C) return hipErrorInvalidValue;
    if (g0.bias || g0.relu || g0.mask || g0.clamp_max > 0.f || g0.Cb || g0.ldc != g0.N) return hipErrorInvalidValue;
    if (splits <= 1) return gemm(g0, ta, tb, s);
    int kchunk = (g0.K + splits - 1) / splits;
    kchunk = (kchunk + 15) / 16 * 16;
    const int z = (g0.K + kchunk - 1) / kchunk;
    GemmArgs g = g0;
    g.C = work;
    g.accumulate = 0;
    g.kchunk = kchunk;
    g.zstride = (size_t)g0.M * g0.N;
    hipError_t e = gemm_launch(g, ta, tb, z, s);
    if (e != hipSuccess) return e;
    const size_t n = (size_t)g0.M * g0.N;
    hipLaunchKernelGGL(fold_splits_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, work, z, n, g0.C,
                       g0.accumulate);
    return hipGetLastError();
}

hipError_t colsum(const float *X, int rows, int ld, int N, float *dst, int splits, float *work, hipStream_t s) {
    if (rows <= 0 || N <= 0) return hipErrorInvalidValue;
    if (splits < 1) splits = 1;
    const int per = (rows + splits - 1) / splits;
    const int z = (rows + per - 1) / per;
    hipLaunchKernelGGL(colsum_kernel, dim3((unsigned)((N + 63) / 64), (unsigned)z), dim3(64), 0, s, X, rows, ld, N,
                       per, work);
    hipLaunchKernelGGL(fold_splits_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, work, z, (size_t)N,
                       dst, 1);
    return hipGetLastError();
}

hipError_t expand_groups(const float *tg, int gpad, const int32_t *group, int n_locs, int lpad, int rows, float *tl,
                         hipStream_t s) {
    const size_t n = (size_t)rows * lpad;
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(expand_groups_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, tg, gpad, group,
                       n_locs, lpad, n, tl);
    return hipGetLastError();
}

hipError_t group_sum(const float *gl, int lpad, const int32_t *member_first, const int32_t *members, int n_groups,
                     int gpad, int rows, float *gg, uint16_t *ggb, hipStream_t s) {
    const size_t n = (size_t)rows * gpad;
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(group_sum_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, gl, lpad, member_first,
                       members, n_groups, gpad, n, gg, ggb);
    return hipGetLastError();
}

hipError_t is_init_rows(const IsParams &p, const uint64_t *runs, uint64_t run_begin, int n_valid,
                        const uint16_t *feat0, hipStream_t s) {
    if (p.n_rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(is_init_kernel, dim3(p.n_rows), dim3(64), 0, s, p, runs, run_begin, n_valid, feat0);
    return hipGetLastError();
}

hipError_t is_record_slots(const uint8_t *status, int n_rows, int32_t *pos, unsigned int *base, unsigned int cap,
                           hipStream_t s) {
    hipLaunchKernelGGL(record_slots_kernel, dim3(1), dim3(1024), 0, s, status, n_rows, pos, base, cap);
    return hipGetLastError();
}

hipError_t is_ce_grad(const float *theta, const float *A, const float *B, const float *w, const uint8_t *tiltable,
                      float tilt_max, int rows, int lpad, float *G, hipStream_t s) {
    const size_t n = (size_t)rows * lpad;
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(ce_grad_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, theta, A, B, w, tiltable,
                       tilt_max, lpad, n, G);
    return hipGetLastError();
}

hipError_t to_bf16(const float *x, uint16_t *y, size_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, y, n);
    return hipGetLastError();
}

hipError_t adam_step(float *w, const float *g, float *m1, float *m2, size_t n, float lr, int t, uint16_t *wb,
                     hipStream_t s) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float bc1 = 1.f - powf(b1, (float)t), bc2 = 1.f - powf(b2, (float)t);
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, w, g, m1, m2, n, lr, b1, b2,
                       eps, bc1, bc2, wb);
    return hipGetLastError();
}

}  // namespace smc
