// is.cpp — learned importance sampling (smc_is_* in include/smc_capi.h).
//
// No counterpart in the reference: cuSMC estimates every probability by plain
// Monte Carlo (statistics.cuh:126-163, smc.cu:234-257).  Everything here runs
// on the model's stream:
//   sampling    lock-step batches of rows; per macro-step the MLP forward pass
//               (three MFMA GEMMs) turns each row's features into log2 tilts
//               and is_advance_kernel moves every row to its next transition
//   training    multilevel cross-entropy: sample a round, keep the runs whose
//               score reaches the round's level, replay them once under the
//               tilts that generated them while recording per macro-step the
//               features and the statistics (A = wins, B = ln2 dt r) per
//               location, then run epochs of forward, output gradient,
//               backward GEMMs and Adam over the recorded pairs
//   estimation  batches of runs under the trained proposal until the 95%
//               half-width of mean(hit 2^log2lr) is within the target
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "capi_internal.hpp"
#include "is_kernels.hpp"

using namespace smc::capi;

namespace {

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) hipFree(p);
        p = nullptr;
        n = 0;
    }
    void ensure(size_t k, const char *what) {
        if (k <= n) return;
        release();
        hip_check(hipMalloc(&p, std::max<size_t>(k, 1) * sizeof(T)), what);
        n = k;
    }
};

size_t rup(size_t x, size_t a) { return (x + a - 1) / a * a; }

// splitmix64 and Box-Muller: weight initialisation only.
struct InitRng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return ((double)(next() >> 11) + 0.5) * 0x1.0p-53; }
    double normal() { return std::sqrt(-2.0 * std::log(uniform())) * std::cos(6.283185307179586 * uniform()); }
};

uint16_t host_bf16(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)((u >> 16) | ((u & 0xFFFF) ? 0x40 : 0));
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

enum class Tilt { NONE, FIXED, MLP };

// Row state of a lock-step batch.  Two sets: compaction copies the rows
// still running from one into the other.
struct RowSet {
    DevBuf<uint16_t> loc;
    DevBuf<double> vars, t, log2lr, qmax, qmin;
    DevBuf<uint32_t> step, trans, job;
    DevBuf<int32_t> chan, sender;
    DevBuf<uint8_t> status;
    DevBuf<uint64_t> digest, run;
    DevBuf<uint16_t> X;  // bf16 features of the rows' current states

    void ensure(size_t r, int n_comp, int n_vars, int fpad, bool feats) {
        loc.ensure(r * n_comp, "hipMalloc rows");
        vars.ensure(r * (size_t)std::max(n_vars, 1), "hipMalloc rows");
        for (DevBuf<double> *b : {&t, &log2lr, &qmax, &qmin}) b->ensure(r, "hipMalloc rows");
        step.ensure(r, "hipMalloc rows");
        trans.ensure(r, "hipMalloc rows");
        job.ensure(r, "hipMalloc rows");
        chan.ensure(r, "hipMalloc rows");
        sender.ensure(r, "hipMalloc rows");
        status.ensure(r, "hipMalloc rows");
        digest.ensure(r, "hipMalloc rows");
        run.ensure(r, "hipMalloc rows");
        if (feats) X.ensure(r * fpad, "hipMalloc features");
    }
    IsRows view() const {
        return IsRows{loc.p, vars.p, t.p, log2lr.p, qmax.p, qmin.p, step.p, trans.p, chan.p, sender.p,
                      status.p, digest.p, run.p, job.p};
    }
};

// Per-run results of a batch, by job.
struct OutSet {
    DevBuf<uint8_t> status;
    DevBuf<double> log2lr, score, t;
    DevBuf<uint32_t> step, trans;
    DevBuf<uint64_t> digest;
    void ensure(size_t n) {
        status.ensure(n, "hipMalloc results");
        for (DevBuf<double> *b : {&log2lr, &score, &t}) b->ensure(n, "hipMalloc results");
        step.ensure(n, "hipMalloc results");
        trans.ensure(n, "hipMalloc results");
        digest.ensure(n, "hipMalloc results");
    }
    IsOut view() const { return IsOut{status.p, log2lr.p, score.p, t.p, step.p, trans.p, digest.p}; }
};

constexpr uint64_t kTrainRunBase = 1ull << 48;  // run indices of training rounds
constexpr int kMaxBatchRows = 1 << 18;
constexpr int kTrainChunk = 1 << 21;  // pairs per training pass (all of a round's, usually)
constexpr double kZ95 = 1.959963984540054;

}  // namespace

struct smc_is {
    smc_model *m = nullptr;
    smc_query q{};
    smc_is_config cfg{};
    int S = 1, C = 1, R = 64, stride = 1;
    size_t lds = 0;
    uint32_t wave_lds = 0;
    bool tab = false;            // advance kernel stages the model tables in LDS
    int max_groups = 0;          // its persistent workgroups when it does
    int n_locs = 0, n_vars = 0, n_comp = 0, F = 0, fpad = 0, H = 128, Lpad = 0;
    int n_groups = 0, Gpad = 0;  // tilt groups: the MLP's outputs
    int G16 = 0;                 // groups rounded up to 16: the outputs the fused forward pass computes
    bool fused_fwd = false;      // forward pass in one kernel (weights fit in LDS)
    DevBuf<int32_t> tix_loc, tix_group;  // per location: its theta column (per location / per group), -1 if untiltable
    DevBuf<uint8_t> mlp_img;     // weights packed as the fused pass stages them in LDS
    bool packed = false;         // mlp_img matches the current weights
    size_t o_w1 = 0, o_b1 = 0, o_w2 = 0, o_b2 = 0, o_w3 = 0, o_b3 = 0, n_params = 0;
    double qinit = 0.0;
    std::vector<uint8_t> h_tiltable;
    // parameters: fp32 master, bf16 copy for the GEMMs, gradient, Adam moments
    DevBuf<float> w, g, m1, m2;
    DevBuf<uint16_t> wb;
    DevBuf<uint8_t> tiltable, gtiltable;
    DevBuf<int32_t> group, member_first, members;
    std::vector<int32_t> h_group;
    DevBuf<float> var_scale, theta_fixed;
    DevBuf<uint16_t> feat0;
    // rows of a lock-step batch, their results, per-row scratch
    RowSet rows_a, rows_b;
    OutSet out;
    DevBuf<int32_t> pos, cpos;
    DevBuf<uint64_t> runs_in;
    DevBuf<float> weight;
    DevBuf<uint16_t> h1, h2;
    DevBuf<float> thg;
    // recorded (run, macro-step) pairs
    DevBuf<uint16_t> rX;
    DevBuf<float> rA, rB, rw, rAg, rBg;  // per location, and summed per tilt group
    size_t rec_cap = 0;
    // training activations of one chunk of pairs
    DevBuf<uint16_t> th1, th2, Gb, dh2b, dh1b, xin;
    DevBuf<float> tth, tthg, dh2, dh1, work, xf;
    DevBuf<unsigned int> counters;  // [0] active rows, [1] record slots used, [2] compaction count
    unsigned int *h_counters = nullptr;  // [0..1] copies of counters[0..1]; [2..3] active counts of alternate macro-steps
    hipEvent_t step_done[2] = {nullptr, nullptr};
    int adam_t = 0;
    bool zero_output = true;  // output layer all zero: theta == 0 exactly
    bool trained = false;

    ~smc_is() {
        if (m) hipSetDevice(m->device);
        if (h_counters) hipHostFree(h_counters);
        for (hipEvent_t e : step_done) if (e) hipEventDestroy(e);
    }

    hipStream_t stream() const { return m->stream; }

    void ensure_rows(int rows, bool feats) {
        const size_t r = (size_t)rows;
        rows_a.ensure(r, n_comp, n_vars, fpad, feats);
        out.ensure(r);
        pos.ensure(r, "hipMalloc rows");
        cpos.ensure(r, "hipMalloc rows");
        runs_in.ensure(r, "hipMalloc rows");
        weight.ensure(r, "hipMalloc rows");
        if (feats) {
            if (!fused_fwd) {
                h1.ensure(r * H, "hipMalloc activations");
                h2.ensure(r * H, "hipMalloc activations");
            }
            thg.ensure(r * Gpad, "hipMalloc tilts");
        }
    }

    // per_location: also the per-location gradient buffer (the MLP test entry)
    void ensure_train(int rows, bool per_location = false) {
        const size_t r = (size_t)rows;
        th1.ensure(r * H, "hipMalloc training");
        th2.ensure(r * H, "hipMalloc training");
        dh2b.ensure(r * H, "hipMalloc training");
        dh1b.ensure(r * H, "hipMalloc training");
        Gb.ensure(r * Gpad, "hipMalloc training");
        if (per_location) tth.ensure(r * Lpad, "hipMalloc training");
        tthg.ensure(r * Gpad, "hipMalloc training");
        dh2.ensure(r * H, "hipMalloc training");
        dh1.ensure(r * H, "hipMalloc training");
        work.ensure(std::max<size_t>((size_t)2048 * 4096, (size_t)1024 * std::max(Gpad, H)), "hipMalloc workspace");
    }

    IsParams params(int rows, uint64_t seed) const {
        IsParams P{};
        P.m = m->dm;
        P.st = rows_a.view();
        P.n_rows = rows;
        P.lanes = S;
        P.var_stride = stride;
        P.lpad = Lpad;
        P.fpad = fpad;
        P.seed = seed;
        P.bound = q.time_bound;
        P.threshold = q.threshold;
        P.qinit = qinit;
        P.reach = q.kind == SMC_QUERY_REACH;
        P.qvar = q.kind == SMC_QUERY_VAR ? q.var : -1;
        P.is_max = q.is_max ? 1 : 0;
        P.max_steps = q.max_steps ? q.max_steps : SMC_DEFAULT_MAX_STEPS;
        P.tiltable = tiltable.p;
        P.tilt_index = tix_loc.p;
        P.tables = tab ? (decltype(P.tables))m->arena : nullptr;
        P.table_bytes = tab ? m->table_bytes : 0;
        P.wave_lds = wave_lds;
        P.theta_cap = Gpad;
        P.theta_len = 0;
        P.var_scale = var_scale.p;
        P.weight = weight.p;
        P.active_count = counters.p;
        return P;
    }

    // theta = MLP(X) for `rows` rows (a multiple of 64): the group outputs in
    // og (rows x ldo; returns ldo: G16 when `compact` and the fused pass
    // runs, else Gpad), expanded to every location in out (rows x Lpad) when
    // out is given; a1, a2 keep the bf16 hidden activations for the backward
    // pass (optional with the fused pass).
    int forward(const uint16_t *Xin, int rows, uint16_t *a1, uint16_t *a2, float *og, bool compact, float *out) {
        int ldo = Gpad;
        if (fused_fwd) {
            if (!packed) {
                mlp_img.ensure(smc::mlp_image_bytes(fpad, H, G16), "hipMalloc packed weights");
                hip_check(smc::mlp_pack(w.p, o_w1, o_b1, o_w2, o_b2, o_w3, o_b3, fpad, H, Gpad, G16, mlp_img.p,
                                        stream()), "pack weights");
                packed = true;
            }
            MlpFwdArgs f{};
            f.X = Xin;
            f.image = mlp_img.p;
            f.col_mask = gtiltable.p;
            f.a1 = a1; f.a2 = a2; f.og = og;
            f.rows = rows; f.fpad = fpad; f.H = H; f.gpad = Gpad; f.g16 = G16;
            f.ldo = ldo = compact ? G16 : Gpad;
            f.clamp_max = cfg.tilt_max;
            hip_check(smc::mlp_forward(f, stream()), "MLP forward");
        } else {
            if (!a1) a1 = h1.p;
            if (!a2) a2 = h2.p;
            GemmArgs g1{};
            g1.A = Xin; g1.lda = fpad; g1.B = wb.p + o_w1; g1.ldb = H; g1.Cb = a1; g1.ldc = H;
            g1.M = rows; g1.N = H; g1.K = fpad; g1.bias = w.p + o_b1; g1.relu = 1;
            hip_check(smc::gemm(g1, false, false, stream()), "MLP layer 1");
            GemmArgs g2{};
            g2.A = a1; g2.lda = H; g2.B = wb.p + o_w2; g2.ldb = H; g2.Cb = a2; g2.ldc = H;
            g2.M = rows; g2.N = H; g2.K = H; g2.bias = w.p + o_b2; g2.relu = 1;
            hip_check(smc::gemm(g2, false, false, stream()), "MLP layer 2");
            GemmArgs g3{};
            g3.A = a2; g3.lda = H; g3.B = wb.p + o_w3; g3.ldb = Gpad; g3.C = og; g3.ldc = Gpad;
            g3.M = rows; g3.N = Gpad; g3.K = H; g3.bias = w.p + o_b3; g3.clamp_max = cfg.tilt_max;
            g3.col_mask = gtiltable.p;
            hip_check(smc::gemm(g3, false, false, stream()), "MLP layer 3");
        }
        if (out) hip_check(smc::expand_groups(og, ldo, group.p, n_locs, Lpad, rows, out, stream()), "expand tilts");
        return ldo;
    }

    // K splits for the weight gradients (K = rows): about 512 workgroups,
    // each layer at least 1024 rows deep, so the fold stays small
    static int splits_for(int M, int N, int K) {
        const int tiles = (M / 64) * (N / 64);
        int s = std::min(K / 1024, 512 / std::max(tiles, 1));
        return std::max(1, std::min(s, 128));
    }

    // Accumulate into g the parameter gradient of sum(G * theta) for rows of
    // a chunk, G given per location in tth: the gradient of a group output is
    // the sum over its locations.
    void backward(const uint16_t *Xc, int rows) {
        hip_check(smc::group_sum(tth.p, Lpad, member_first.p, members.p, n_groups, Gpad, rows, tthg.p, Gb.p,
                                 stream()), "group gradient");
        backward_groups(Xc, rows);
    }

    // The same from the group outputs' gradient: tthg (fp32) and Gb (bf16);
    // th1/th2 hold the activations.
    void backward_groups(const uint16_t *Xc, int rows) {
        hipStream_t s = stream();
        const int csplit = std::max(1, std::min(1024, rows / 128));
        GemmArgs a{};
        a.A = th2.p; a.lda = H; a.B = Gb.p; a.ldb = Gpad; a.C = g.p + o_w3; a.ldc = Gpad;
        a.M = H; a.N = Gpad; a.K = rows; a.accumulate = 1;
        hip_check(smc::gemm_splitk(a, true, false, splits_for(H, Gpad, rows), work.p, s), "grad W3");
        hip_check(smc::colsum(tthg.p, rows, Gpad, Gpad, g.p + o_b3, csplit, work.p, s), "grad b3");
        GemmArgs b{};
        b.A = Gb.p; b.lda = Gpad; b.B = wb.p + o_w3; b.ldb = Gpad; b.C = dh2.p; b.Cb = dh2b.p; b.ldc = H;
        b.M = rows; b.N = H; b.K = Gpad; b.mask = th2.p; b.ldm = H;
        hip_check(smc::gemm(b, false, true, s), "grad h2");
        GemmArgs c{};
        c.A = th1.p; c.lda = H; c.B = dh2b.p; c.ldb = H; c.C = g.p + o_w2; c.ldc = H;
        c.M = H; c.N = H; c.K = rows; c.accumulate = 1;
        hip_check(smc::gemm_splitk(c, true, false, splits_for(H, H, rows), work.p, s), "grad W2");
        hip_check(smc::colsum(dh2.p, rows, H, H, g.p + o_b2, csplit, work.p, s), "grad b2");
        GemmArgs d{};
        d.A = dh2b.p; d.lda = H; d.B = wb.p + o_w2; d.ldb = H; d.C = dh1.p; d.Cb = dh1b.p; d.ldc = H;
        d.M = rows; d.N = H; d.K = H; d.mask = th1.p; d.ldm = H;
        hip_check(smc::gemm(d, false, true, s), "grad h1");
        GemmArgs e{};
        e.A = Xc; e.lda = fpad; e.B = dh1b.p; e.ldb = H; e.C = g.p + o_w1; e.ldc = H;
        e.M = fpad; e.N = H; e.K = rows; e.accumulate = 1;
        hip_check(smc::gemm_splitk(e, true, false, splits_for(fpad, H, rows), work.p, s), "grad W1");
        hip_check(smc::colsum(dh1.p, rows, H, H, g.p + o_b1, csplit, work.p, s), "grad b1");
    }

    struct Batch {
        Tilt tilt = Tilt::NONE;
        bool lockstep = true, early_stop = true, record = false, trace = false;
        const uint64_t *d_runs = nullptr;  // run of each row, or run_begin + row
        uint64_t run_begin = 0, seed = 0;
        int n_valid = 0;
    };

    // Run the rows of one batch to their end; the results land in `out`
    // by job (the row's index at the start).  Returns the number of launches.
    int run_rows(const Batch &b) {
        hipStream_t s = stream();
        int rows = (int)rup((size_t)b.n_valid, 64);
        const bool feats = b.tilt == Tilt::MLP;
        if (feats && !b.lockstep) throw std::runtime_error("the MLP proposal needs lock-step advance");
        ensure_rows(rows, feats);
        RowSet *A = &rows_a, *B = &rows_b;
        IsParams P = params(rows, b.seed);
        P.feat = feats ? A->X.p : nullptr;
        hip_check(smc::is_init_rows(P, b.d_runs, b.run_begin, b.n_valid, feat0.p, s), "init rows");
        P.early_stop = b.early_stop ? 1 : 0;
        P.trace = b.trace ? 1 : 0;
        P.stop_at_transition = b.lockstep ? 1 : 0;
        if (b.tilt == Tilt::FIXED) {
            P.theta = theta_fixed.p;
            P.theta_stride = 0;
            P.theta_len = Lpad;
            P.tilt_index = tix_loc.p;
        }
        size_t lds_use = lds;
        if (b.record) {
            P.rec_X = rX.p; P.rec_A = rA.p; P.rec_B = rB.p; P.rec_w = rw.p;
            hip_check(hipMemsetAsync(counters.p + 1, 0, sizeof(unsigned int), s), "clear record count");
            // the recording build reads the tables from global memory and
            // keeps its per-slot statistics in LDS (16 C bytes per lane)
            P.tables = nullptr;
            P.table_bytes = 0;
            P.wave_lds = wave_lds + (uint32_t)(16 * C * 64);
            lds_use = (size_t)Lpad * 12 + P.wave_lds;
        }
        // Lock-step macro-steps.  The host never waits for the step it just
        // launched: after launching step k it reads the active count of step
        // k-1 (an upper bound on the rows still live after step k), so the
        // GPU always has the next step queued.  A count of zero ends the loop
        // (step k then found no live row); compaction sizes the new row set
        // from the same bound.
        int launches = 0;
        for (int k = 0;; ++k) {
            if (feats) {
                // the tilts from the features of the rows' states (the
                // advance kernel then overwrites the features, in order)
                P.theta_stride = forward(A->X.p, rows, nullptr, nullptr, thg.p, true, nullptr);
                P.theta_len = P.theta_stride;
                P.theta = thg.p;
                P.tilt_index = tix_group.p;
            }
            if (b.record) {
                hip_check(smc::is_record_slots(A->status.p, rows, pos.p, counters.p + 1, (unsigned int)rec_cap, s),
                          "record slots");
                P.rec_pos = pos.p;
            }
            unsigned int *cnt = counters.p + 3 + (k & 1);
            hip_check(hipMemsetAsync(cnt, 0, sizeof(unsigned int), s), "clear active count");
            P.active_count = cnt;
            hip_check(smc::is_launch_advance(P, C, lds_use, max_groups, s), "advance kernel launch");
            ++launches;
            if (!b.lockstep) break;
            hip_check(hipMemcpyAsync(h_counters + 2 + (k & 1), cnt, sizeof(unsigned int), hipMemcpyDeviceToHost, s),
                      "copy active count");
            hip_check(hipEventRecord(step_done[k & 1], s), "event");
            if (k == 0) continue;
            hip_check(hipEventSynchronize(step_done[(k - 1) & 1]), "advance");
            const int active = (int)h_counters[2 + ((k - 1) & 1)];
            if (active == 0) break;
            if (active * 2 <= rows && rows > 1024) {
                // compaction: results of the finished rows out, live rows
                // to the front of the other set, in row order
                hip_check(smc::is_collect(P, out.view(), s), "collect results");
                hip_check(hipMemsetAsync(counters.p + 2, 0, sizeof(unsigned int), s), "clear compaction count");
                hip_check(smc::is_record_slots(A->status.p, rows, cpos.p, counters.p + 2, 0xFFFFFFFFu, s),
                          "compaction slots");
                const int nrows = (int)rup((size_t)active, 64);
                B->ensure((size_t)nrows, n_comp, n_vars, fpad, feats);
                hip_check(hipMemsetAsync(B->status.p, 0, (size_t)nrows, s), "clear rows");
                hip_check(hipMemsetAsync(B->job.p, 0xFF, (size_t)nrows * 4, s), "clear rows");
                hip_check(smc::is_compact(P, cpos.p, B->view(), feats ? A->X.p : nullptr, feats ? B->X.p : nullptr,
                                          s), "compact rows");
                std::swap(A, B);
                rows = nrows;
                P.st = A->view();
                P.n_rows = rows;
                P.feat = feats ? A->X.p : nullptr;
            }
        }
        hip_check(smc::is_collect(P, out.view(), s), "collect results");
        hip_check(hipStreamSynchronize(s), "advance");
        return launches;
    }

    size_t rec_pairs = 0;  // recorded pairs (padded to 64 with zero pairs)

    void ensure_records() {
        if (rec_cap) return;
        // pairs of (features, wins, ln2 dt r) within 8 GiB
        const size_t per_pair = (size_t)fpad * 2 + (size_t)Lpad * 8 + (size_t)Gpad * 8 + 4;
        rec_cap = std::min<size_t>(((size_t)8 << 30) / per_pair, (size_t)1 << 23) / 64 * 64;
        rX.ensure(rec_cap * fpad, "hipMalloc records");
        rA.ensure(rec_cap * Lpad, "hipMalloc records");
        rB.ensure(rec_cap * Lpad, "hipMalloc records");
        rw.ensure(rec_cap, "hipMalloc records");
        rAg.ensure(rec_cap * Gpad, "hipMalloc records");
        rBg.ensure(rec_cap * Gpad, "hipMalloc records");
    }

    // Replay runs (device array runs_in, weights in `weight`) under the
    // batch's proposal and record their macro-steps; returns the pair count.
    size_t record(Batch b) {
        ensure_records();
        b.record = true;
        b.d_runs = runs_in.p;
        run_rows(b);
        hipStream_t st = stream();
        hip_check(hipMemcpyAsync(h_counters, counters.p, 2 * sizeof(unsigned int), hipMemcpyDeviceToHost, st), "copy count");
        hip_check(hipStreamSynchronize(st), "record");
        const size_t pairs = std::min<size_t>(h_counters[1], rec_cap);
        const size_t padded = rup(pairs, 64);
        if (padded > pairs) {
            const size_t z = padded - pairs;
            hip_check(hipMemsetAsync(rX.p + pairs * fpad, 0, z * fpad * 2, st), "clear records");
            hip_check(hipMemsetAsync(rA.p + pairs * Lpad, 0, z * Lpad * 4, st), "clear records");
            hip_check(hipMemsetAsync(rB.p + pairs * Lpad, 0, z * Lpad * 4, st), "clear records");
            hip_check(hipMemsetAsync(rw.p + pairs, 0, z * 4, st), "clear records");
        }
        rec_pairs = padded;
        // statistics per tilt group (theta is shared inside a group), once
        for (size_t off = 0; off < padded; off += kTrainChunk) {
            const int rows_c = (int)std::min<size_t>(kTrainChunk, padded - off);
            hip_check(smc::group_sum(rA.p + off * Lpad, Lpad, member_first.p, members.p, n_groups, Gpad, rows_c,
                                     rAg.p + off * Gpad, nullptr, st), "group wins");
            hip_check(smc::group_sum(rB.p + off * Lpad, Lpad, member_first.p, members.p, n_groups, Gpad, rows_c,
                                     rBg.p + off * Gpad, nullptr, st), "group rates");
        }
        return pairs;
    }

    // g = d(-sum_pairs w ln q)/d parameters over the recorded pairs.
    void ce_gradient() {
        hipStream_t st = stream();
        hip_check(hipMemsetAsync(g.p, 0, n_params * 4, st), "clear gradient");
        if (rec_pairs) ensure_train((int)std::min<size_t>(kTrainChunk, rec_pairs));
        for (size_t off = 0; off < rec_pairs; off += kTrainChunk) {
            const int rows_c = (int)std::min<size_t>(kTrainChunk, rec_pairs - off);
            const uint16_t *Xc = rX.p + off * fpad;
            forward(Xc, rows_c, th1.p, th2.p, tthg.p, false, nullptr);
            hip_check(smc::is_ce_grad(tthg.p, rAg.p + off * Gpad, rBg.p + off * Gpad, rw.p + off, gtiltable.p,
                                      cfg.tilt_max, rows_c, Gpad, tthg.p, st), "output gradient");
            hip_check(smc::to_bf16(tthg.p, Gb.p, (size_t)rows_c * Gpad, st), "output gradient to bf16");
            backward_groups(Xc, rows_c);
        }
    }

    template <class T>
    void fetch(std::vector<T> &dst, const DevBuf<T> &src, size_t n) {
        dst.resize(n);
        if (n) hip_check(hipMemcpyAsync(dst.data(), src.p, n * sizeof(T), hipMemcpyDeviceToHost, stream()), "copy rows");
    }

    void upload_weights(const std::vector<float> &hw) {
        hip_check(hipMemcpyAsync(w.p, hw.data(), n_params * 4, hipMemcpyHostToDevice, stream()), "upload weights");
        hip_check(smc::to_bf16(w.p, wb.p, n_params, stream()), "weights to bf16");
        packed = false;
        hip_check(hipStreamSynchronize(stream()), "upload weights");
    }
};

extern "C" {

int smc_is_create(smc_model *m, const smc_query *q, const smc_is_config *cfg_in, smc_is **out) {
    if (!m || !q || !out) return fail(SMC_E_INVALID, "null argument");
    *out = nullptr;
    return guarded([&]() -> int {
        const smc::ModelImage &img = m->img;
        if (q->kind != SMC_QUERY_REACH && q->kind != SMC_QUERY_VAR) throw std::runtime_error("unknown query kind");
        if (q->kind == SMC_QUERY_VAR && (q->var < 0 || q->var >= img.n_vars))
            throw std::runtime_error("query variable id out of range");
        if (q->kind == SMC_QUERY_VAR && q->is_estimate)
            throw std::runtime_error("importance sampling estimates probabilities: use a var query with is_estimate = 0");
        if (!(q->time_bound > 0.0)) throw std::runtime_error("time bound must be positive");
        auto s = std::make_unique<smc_is>();
        s->m = m;
        s->q = *q;
        smc_is_config c = cfg_in ? *cfg_in : smc_is_config{};
        if (c.hidden <= 0) c.hidden = 128;
        if (c.rounds <= 0) c.rounds = 5;
        if (c.runs_per_round == 0) c.runs_per_round = 100000;
        if (!(c.elite_frac > 0.0 && c.elite_frac < 1.0)) c.elite_frac = 0.1;
        if (c.epochs <= 0) c.epochs = 20;
        if (!(c.learning_rate > 0.f)) c.learning_rate = 1e-3f;
        if (!(c.tilt_max > 0.f)) c.tilt_max = 4.f;
        if (c.hidden % 64 || c.hidden > 1024) throw std::runtime_error("hidden width must be a multiple of 64, at most 1024");
        if (c.runs_per_round > (uint64_t)kMaxBatchRows * 8) throw std::runtime_error("limit: runs_per_round too large");
        s->cfg = c;
        hip_check(hipSetDevice(m->device), "hipSetDevice");

        // launch shape of the advance kernel
        // lanes per run: the model's tuned packing (smc_autotune) when there
        // is one, else the widest
        const int P = img.n_leaves;
        if (P > 512) throw std::runtime_error("limit: importance sampling takes networks of at most 512 components");
        s->S = P < 64 ? P : 64;
        if (m->has_tuned && m->tuned.lanes_per_run > 0 && m->tuned.lanes_per_run <= s->S &&
            P / m->tuned.lanes_per_run <= 8) {
            s->S = m->tuned.lanes_per_run;
        }
        // SMC_IS_LANES=<power of two> overrides the lanes per run (diagnosis:
        // the sampler's best packing may differ from the plain simulator's)
        if (const char *lv = std::getenv("SMC_IS_LANES")) {
            const int L = std::atoi(lv);
            if (L >= 1 && L <= 64 && (L & (L - 1)) == 0 && L <= P && P / L <= 8) s->S = L;
        }
        s->C = P / s->S;
        if (s->C > 8) throw std::runtime_error("internal: more than 8 slots per lane");
        s->R = 64 / s->S;
        s->stride = (img.n_vars + 1) | 1;  // clock time, then the variables

        s->n_locs = (int)img.locs.size();
        s->n_vars = img.n_vars;
        s->n_comp = img.n_comp;
        s->F = s->n_locs + s->n_vars + 1;
        s->fpad = (int)rup((size_t)s->F, 64);
        s->H = c.hidden;
        s->Lpad = (int)rup((size_t)s->n_locs, 64);

        // tilt groups: structurally identical components (same locations,
        // delay classes, rates and edge shapes — instances of one template)
        // share the tilt of corresponding locations; per_location gives
        // every location its own
        s->h_group.assign(s->n_locs, 0);
        if (c.per_location) {
            for (int l = 0; l < s->n_locs; ++l) s->h_group[l] = l;
            s->n_groups = s->n_locs;
        } else {
            std::map<std::vector<int64_t>, int> base_of;
            int next = 0;
            for (int cpt = 0; cpt < img.n_comp; ++cpt) {
                const int l0 = img.comp_first_loc[cpt], l1 = img.comp_first_loc[cpt + 1];
                std::vector<int64_t> sig{l1 - l0};
                for (int l = l0; l < l1; ++l) {
                    const DevLoc &Lc = img.locs[l];
                    int64_t rate_key;
                    if (Lc.rate_const) memcpy(&rate_key, &Lc.rate_value, 8);
                    else rate_key = (int64_t)Lc.rate ^ ((int64_t)1 << 62);
                    sig.insert(sig.end(), {(int64_t)Lc.klass, (int64_t)Lc.type, (int64_t)Lc.n_inv, rate_key,
                                           (int64_t)Lc.edge_n});
                    for (int e = (int)Lc.edge_first; e < (int)Lc.edge_first + Lc.edge_n; ++e) {
                        sig.push_back((int64_t)l + img.edges[e].dst - l0);  // destinations are relative
                        sig.push_back((int64_t)img.edges[e].channel);
                    }
                }
                auto it = base_of.find(sig);
                int base;
                if (it == base_of.end()) {
                    base = next;
                    next += l1 - l0;
                    base_of.emplace(std::move(sig), base);
                } else {
                    base = it->second;
                }
                for (int l = l0; l < l1; ++l) s->h_group[l] = base + (l - l0);
            }
            s->n_groups = next;
        }
        s->Gpad = (int)rup((size_t)std::max(s->n_groups, 1), 64);
        s->G16 = (int)rup((size_t)std::max(s->n_groups, 1), 16);
        // advance kernel LDS: per wave the variables of its 64/S rows, their
        // tilt tables (theta and 2^theta, up to Gpad entries) and the race
        // tree's levels (2 log2(S) doubles per lane), after a shared part
        // (tilt index and a shared 2^theta vector, Lpad each)
        int levels = 0;
        while ((1 << levels) < s->S) ++levels;
        s->wave_lds = (uint32_t)rup((size_t)s->R * s->stride * sizeof(double) + (size_t)s->R * s->Gpad * 12 +
                                        (size_t)2 * levels * 64 * sizeof(double), 16);
        if (s->wave_lds > 60000) throw std::runtime_error("limit: variables exceed the LDS budget");
        {
            const size_t shared = (size_t)s->Lpad * 12;
            // hot tables staged in LDS when the model's tuned launch does so
            // (or SMC_IS_TABLES=1/0 says) and four waves fit beside them
            const char *tv = std::getenv("SMC_IS_TABLES");
            bool want = m->has_tuned && m->tuned.tables == 1;
            if (tv) want = std::string(tv) == "1";
            const size_t need = (size_t)m->table_bytes + shared + 4 * (size_t)s->wave_lds;
            s->tab = want && m->table_bytes > 0 && need <= 160 * 1024;
            s->lds = s->tab ? need : shared + s->wave_lds;
            if (s->lds > 160 * 1024) throw std::runtime_error("limit: the advance kernel's LDS exceeds the budget");
            // two waves per SIMD: two workgroups of four waves per CU
            const int per_cu = s->tab ? std::max(1, std::min(2, (int)((160 * 1024) / need))) : 0;
            s->max_groups = per_cu * m->cu_count;
        }
        // the fused forward pass when its weights fit in LDS;
        // SMC_IS_FORWARD=gemm forces the three-GEMM path (the two agree bit
        // for bit: same products in the same order)
        {
            const char *fw = std::getenv("SMC_IS_FORWARD");
            s->fused_fwd = !(fw && std::string(fw) == "gemm") &&
                           smc::mlp_forward_supported(s->fpad, s->H, s->G16, s->cfg.tilt_max);
        }
        const size_t fp = s->fpad, H = s->H, L = s->Lpad, Gp = s->Gpad;
        s->o_w1 = 0;
        s->o_b1 = fp * H;
        s->o_w2 = s->o_b1 + H;
        s->o_b2 = s->o_w2 + H * H;
        s->o_w3 = s->o_b2 + H;
        s->o_b3 = s->o_w3 + H * Gp;
        s->n_params = s->o_b3 + Gp;
        s->qinit = (q->kind == SMC_QUERY_VAR && img.var_kind[q->var] == SMC_VAR_INT) ? img.var_init[q->var] : 0.0;

        // tiltable locations: every location that is not zero-delay
        s->h_tiltable.assign(L, 0);
        for (int l = 0; l < s->n_locs; ++l) s->h_tiltable[l] = img.locs[l].klass != SMC_K_ZERO;
        // feature scales: clocks by the time bound, other variables by the
        // largest initial magnitude or the component count
        std::vector<float> scale(std::max(s->n_vars, 1), 1.f);
        double imax = (double)img.n_comp;
        for (int v = 0; v < img.n_vars; ++v)
            if (img.var_kind[v] != SMC_VAR_CLOCK) imax = std::max(imax, std::fabs(img.var_init[v]));
        for (int v = 0; v < img.n_vars; ++v)
            scale[v] = img.var_kind[v] == SMC_VAR_CLOCK ? (float)(1.0 / q->time_bound) : (float)(1.0 / std::max(1.0, imax));
        std::vector<uint16_t> f0(fp, 0);
        for (int cpt = 0; cpt < img.n_comp; ++cpt) f0[img.comp_init_loc[cpt]] = 0x3F80;
        for (int v = 0; v < img.n_vars; ++v)
            f0[s->n_locs + v] = host_bf16((float)(img.var_init[v] * (double)scale[v]));

        hipStream_t st = m->stream;
        s->tiltable.ensure(L, "hipMalloc tiltable");
        std::vector<uint8_t> gt(Gp, 0);
        std::vector<int32_t> mfirst(s->n_groups + 1, 0), mem(s->n_locs);
        for (int l = 0; l < s->n_locs; ++l) {
            gt[s->h_group[l]] |= s->h_tiltable[l];
            ++mfirst[s->h_group[l] + 1];
        }
        for (int k = 0; k < s->n_groups; ++k) mfirst[k + 1] += mfirst[k];
        {
            std::vector<int32_t> fill(mfirst.begin(), mfirst.end() - 1);
            for (int l = 0; l < s->n_locs; ++l) mem[fill[s->h_group[l]]++] = l;
        }
        s->gtiltable.ensure(Gp, "hipMalloc tiltable");
        {
            std::vector<int32_t> tl(L, -1), tg(L, -1);
            for (int l = 0; l < s->n_locs; ++l)
                if (s->h_tiltable[l]) { tl[l] = l; tg[l] = s->h_group[l]; }
            s->tix_loc.ensure(L, "hipMalloc tilt index");
            s->tix_group.ensure(L, "hipMalloc tilt index");
            hip_check(hipMemcpyAsync(s->tix_loc.p, tl.data(), L * 4, hipMemcpyHostToDevice, st), "upload");
            hip_check(hipMemcpyAsync(s->tix_group.p, tg.data(), L * 4, hipMemcpyHostToDevice, st), "upload");
            hip_check(hipStreamSynchronize(st), "upload");
        }
        s->group.ensure(std::max(s->n_locs, 1), "hipMalloc groups");
        s->member_first.ensure(mfirst.size(), "hipMalloc groups");
        s->members.ensure(std::max(s->n_locs, 1), "hipMalloc groups");
        hip_check(hipMemcpyAsync(s->gtiltable.p, gt.data(), Gp, hipMemcpyHostToDevice, st), "upload");
        hip_check(hipMemcpyAsync(s->group.p, s->h_group.data(), s->n_locs * 4, hipMemcpyHostToDevice, st), "upload");
        hip_check(hipMemcpyAsync(s->member_first.p, mfirst.data(), mfirst.size() * 4, hipMemcpyHostToDevice, st),
                  "upload");
        hip_check(hipMemcpyAsync(s->members.p, mem.data(), mem.size() * 4, hipMemcpyHostToDevice, st), "upload");
        s->var_scale.ensure(scale.size(), "hipMalloc scales");
        s->theta_fixed.ensure(L, "hipMalloc tilts");
        s->feat0.ensure(fp, "hipMalloc features");
        // counters: [0] active rows, [1] recorded pairs, [2] compaction slots,
        // [3..4] active rows after alternate macro-steps
        s->counters.ensure(5, "hipMalloc counters");
        hip_check(hipHostMalloc((void **)&s->h_counters, 4 * sizeof(unsigned int), hipHostMallocDefault), "hipHostMalloc");
        for (hipEvent_t &e : s->step_done)
            hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
        hip_check(hipMemcpyAsync(s->tiltable.p, s->h_tiltable.data(), L, hipMemcpyHostToDevice, st), "upload");
        hip_check(hipMemcpyAsync(s->var_scale.p, scale.data(), scale.size() * 4, hipMemcpyHostToDevice, st), "upload");
        hip_check(hipMemcpyAsync(s->feat0.p, f0.data(), fp * 2, hipMemcpyHostToDevice, st), "upload");

        // parameters: He initialisation of the hidden layers; the output
        // layer starts at zero, so the untrained proposal is the base law
        s->w.ensure(s->n_params, "hipMalloc weights");
        s->wb.ensure(s->n_params, "hipMalloc weights");
        s->g.ensure(s->n_params, "hipMalloc gradient");
        s->m1.ensure(s->n_params, "hipMalloc Adam");
        s->m2.ensure(s->n_params, "hipMalloc Adam");
        std::vector<float> hw(s->n_params, 0.f);
        InitRng rng{c.seed ^ 0x5DEECE66Dull};
        const double s1 = std::sqrt(2.0 / s->F), s2 = std::sqrt(2.0 / (double)H);
        for (int f = 0; f < s->F; ++f)
            for (size_t h = 0; h < H; ++h) hw[s->o_w1 + f * H + h] = (float)(rng.normal() * s1);
        for (size_t i = 0; i < H * H; ++i) hw[s->o_w2 + i] = (float)(rng.normal() * s2);
        hip_check(hipMemsetAsync(s->m1.p, 0, s->n_params * 4, st), "clear Adam");
        hip_check(hipMemsetAsync(s->m2.p, 0, s->n_params * 4, st), "clear Adam");
        s->upload_weights(hw);
        *out = s.release();
        return SMC_OK;
    });
}

void smc_is_destroy(smc_is *s) { delete s; }

int smc_is_shape(const smc_is *s, int32_t *features_padded, int32_t *hidden, int32_t *outputs_padded,
                 int32_t *locations_padded, uint64_t *n_params) {
    if (!s) return fail(SMC_E_INVALID, "null argument");
    if (features_padded) *features_padded = s->fpad;
    if (hidden) *hidden = s->H;
    if (outputs_padded) *outputs_padded = s->Gpad;
    if (locations_padded) *locations_padded = s->Lpad;
    if (n_params) *n_params = s->n_params;
    return SMC_OK;
}

int smc_is_groups(const smc_is *s, int32_t *group_of_location) {
    if (!s) {
        fail(SMC_E_INVALID, "null argument");
        return -1;
    }
    if (group_of_location) memcpy(group_of_location, s->h_group.data(), s->h_group.size() * 4);
    return s->n_groups;
}

int smc_is_get_params(const smc_is *s, float *out) {
    if (!s || !out) return fail(SMC_E_INVALID, "null argument");
    return guarded([&]() -> int {
        hip_check(hipSetDevice(s->m->device), "hipSetDevice");
        hip_check(hipMemcpyAsync(out, s->w.p, s->n_params * 4, hipMemcpyDeviceToHost, s->stream()), "copy weights");
        hip_check(hipStreamSynchronize(s->stream()), "copy weights");
        return SMC_OK;
    });
}

int smc_is_set_params(smc_is *s, const float *in) {
    if (!s || !in) return fail(SMC_E_INVALID, "null argument");
    return guarded([&]() -> int {
        hip_check(hipSetDevice(s->m->device), "hipSetDevice");
        s->upload_weights(std::vector<float>(in, in + s->n_params));
        s->zero_output = false;
        return SMC_OK;
    });
}

int smc_is_mlp_eval(smc_is *s, const float *x, int rows, const float *gout, float *theta_out, float *grad_out) {
    if (!s || !x || rows <= 0) return fail(SMC_E_INVALID, "null argument or no rows");
    return guarded([&]() -> int {
        hip_check(hipSetDevice(s->m->device), "hipSetDevice");
        hipStream_t st = s->stream();
        const int rp = (int)rup((size_t)rows, 64);
        const size_t nx = (size_t)rp * s->fpad, nt = (size_t)rp * s->Lpad;
        s->ensure_train(rp, true);
        s->xf.ensure(std::max(nx, nt), "hipMalloc inputs");
        s->xin.ensure(nx, "hipMalloc inputs");
        hip_check(hipMemsetAsync(s->xf.p, 0, nx * 4, st), "clear inputs");
        hip_check(hipMemcpyAsync(s->xf.p, x, (size_t)rows * s->fpad * 4, hipMemcpyHostToDevice, st), "upload inputs");
        hip_check(smc::to_bf16(s->xf.p, s->xin.p, nx, st), "inputs to bf16");
        s->forward(s->xin.p, rp, s->th1.p, s->th2.p, s->tthg.p, false, s->tth.p);
        std::vector<float> th((size_t)rows * s->Lpad);
        hip_check(hipMemcpyAsync(th.data(), s->tth.p, th.size() * 4, hipMemcpyDeviceToHost, st), "copy tilts");
        hip_check(hipStreamSynchronize(st), "MLP forward");
        if (theta_out) memcpy(theta_out, th.data(), th.size() * 4);
        if (gout) {
            // the output gradient, zero where the output is clamped or untiltable
            std::vector<float> gm((size_t)rp * s->Lpad, 0.f);
            for (int r = 0; r < rows; ++r)
                for (int l = 0; l < s->Lpad; ++l) {
                    const size_t i = (size_t)r * s->Lpad + l;
                    if (s->h_tiltable[l] && std::fabs(th[i]) < s->cfg.tilt_max) gm[i] = gout[i];
                }
            hip_check(hipMemcpyAsync(s->tth.p, gm.data(), nt * 4, hipMemcpyHostToDevice, st), "upload output gradient");
            hip_check(hipMemsetAsync(s->g.p, 0, s->n_params * 4, st), "clear gradient");
            s->backward(s->xin.p, rp);
            if (grad_out)
                hip_check(hipMemcpyAsync(grad_out, s->g.p, s->n_params * 4, hipMemcpyDeviceToHost, st), "copy gradient");
            hip_check(hipStreamSynchronize(st), "MLP backward");
        }
        return SMC_OK;
    });
}

static void accumulate(const std::vector<uint8_t> &stat, const std::vector<double> &lr, const std::vector<uint32_t> &stp,
                       const std::vector<uint32_t> &trn, size_t n, double *sw, double *sw2, uint64_t *hits,
                       uint64_t *steps, uint64_t *transitions) {
    for (size_t i = 0; i < n; ++i) {
        if (stat[i] & IS_HIT) {
            const double w = std::exp2(lr[i]);
            *sw += w;
            *sw2 += w * w;
            ++*hits;
        }
        *steps += stp[i];
        *transitions += trn[i];
    }
}

static void fill_is_result(uint64_t n, uint64_t hits, double sw, double sw2, uint64_t steps, uint64_t trans,
                           double ms, smc_is_result *res) {
    memset(res, 0, sizeof *res);
    res->n_runs = n;
    res->hits = hits;
    const double N = (double)n;
    res->estimate = n ? sw / N : 0.0;
    const double var = n > 1 ? std::max(0.0, (sw2 - sw * sw / N) / (N - 1.0)) : 0.0;
    res->rel_half_width = res->estimate > 0.0 ? kZ95 * std::sqrt(var / N) / res->estimate : INFINITY;
    res->ess = sw2 > 0.0 ? sw * sw / sw2 : 0.0;
    res->steps = steps;
    res->transitions = trans;
    res->ms = ms;
}

int smc_is_sample(smc_is *s, const float *theta, int use_mlp, uint64_t seed, uint64_t run_begin, uint64_t n_runs,
                  int lockstep, int early_stop, smc_is_outputs *out, smc_is_result *res) {
    if (!s || !res) return fail(SMC_E_INVALID, "null argument");
    return guarded([&]() -> int {
        auto t0 = std::chrono::steady_clock::now();
        if (n_runs == 0) throw std::runtime_error("n_runs must be positive");
        hip_check(hipSetDevice(s->m->device), "hipSetDevice");
        hipStream_t st = s->stream();
        smc_is::Batch b;
        b.tilt = theta ? Tilt::FIXED : (use_mlp ? Tilt::MLP : Tilt::NONE);
        b.lockstep = lockstep != 0;
        b.early_stop = early_stop != 0;
        b.trace = out && out->digest;
        b.seed = seed;
        if (theta) {
            std::vector<float> th(s->Lpad, 0.f);
            memcpy(th.data(), theta, (size_t)s->n_locs * 4);
            hip_check(hipMemcpyAsync(s->theta_fixed.p, th.data(), th.size() * 4, hipMemcpyHostToDevice, st), "upload tilts");
        }
        double sw = 0, sw2 = 0;
        uint64_t hits = 0, steps = 0, trans = 0;
        std::vector<uint8_t> stat;
        std::vector<double> lr, tm;
        std::vector<uint32_t> stp, trn;
        std::vector<uint64_t> dig;
        for (uint64_t done = 0; done < n_runs;) {
            const int nb = (int)std::min<uint64_t>(n_runs - done, kMaxBatchRows);
            b.run_begin = run_begin + done;
            b.n_valid = nb;
            s->run_rows(b);
            s->fetch(stat, s->out.status, nb);
            s->fetch(lr, s->out.log2lr, nb);
            s->fetch(stp, s->out.step, nb);
            s->fetch(trn, s->out.trans, nb);
            if (out && out->digest) s->fetch(dig, s->out.digest, nb);
            if (out && out->time) s->fetch(tm, s->out.t, nb);
            hip_check(hipStreamSynchronize(st), "copy rows");
            accumulate(stat, lr, stp, trn, nb, &sw, &sw2, &hits, &steps, &trans);
            if (out) {
                for (int i = 0; i < nb; ++i) {
                    if (out->flag) out->flag[done + i] = (stat[i] & IS_HIT) ? 1 : 0;
                    if (out->log2lr) out->log2lr[done + i] = lr[i];
                    if (out->steps) out->steps[done + i] = stp[i];
                    if (out->transitions) out->transitions[done + i] = trn[i];
                    if (out->digest) out->digest[done + i] = dig[i];
                    if (out->time) out->time[done + i] = tm[i];
                }
            }
            done += (uint64_t)nb;
        }
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        fill_is_result(n_runs, hits, sw, sw2, steps, trans, ms, res);
        return SMC_OK;
    });
}

int smc_is_train(smc_is *s, smc_is_round *rounds) {
    if (!s) return fail(SMC_E_INVALID, "null argument");
    return guarded([&]() -> int {
        hip_check(hipSetDevice(s->m->device), "hipSetDevice");
        hipStream_t st = s->stream();
        const smc_is_config &c = s->cfg;
        const int N = (int)c.runs_per_round;
        const bool reach = s->q.kind == SMC_QUERY_REACH;
        s->ensure_records();

        std::vector<uint8_t> stat;
        std::vector<double> lr, qm;
        std::vector<uint32_t> stp, trn;
        for (int r = 0; r < c.rounds; ++r) {
            auto t0 = std::chrono::steady_clock::now();
            // 1. sample a round under the current proposal (each round its own
            // runs); before any training the output layer is zero, theta is
            // exactly 0, and the runs are the base law's in one launch
            smc_is::Batch b;
            b.tilt = s->zero_output ? Tilt::NONE : Tilt::MLP;
            b.lockstep = !s->zero_output;
            b.seed = c.seed;
            b.run_begin = kTrainRunBase + (uint64_t)r * c.runs_per_round;
            std::vector<uint64_t> run_of;
            std::vector<double> score;
            std::vector<uint8_t> hit;
            std::vector<double> l2;
            double sw = 0, sw2 = 0;
            uint64_t hits = 0, steps = 0, trans = 0;
            for (int done = 0; done < N;) {
                const int nb = std::min(N - done, kMaxBatchRows);
                b.n_valid = nb;
                b.run_begin = kTrainRunBase + (uint64_t)r * c.runs_per_round + (uint64_t)done;
                s->run_rows(b);
                s->fetch(stat, s->out.status, nb);
                s->fetch(lr, s->out.log2lr, nb);
                s->fetch(qm, s->out.score, nb);
                s->fetch(stp, s->out.step, nb);
                s->fetch(trn, s->out.trans, nb);
                hip_check(hipStreamSynchronize(st), "copy rows");
                accumulate(stat, lr, stp, trn, nb, &sw, &sw2, &hits, &steps, &trans);
                for (int i = 0; i < nb; ++i) {
                    const bool h = (stat[i] & IS_HIT) != 0;
                    run_of.push_back(b.run_begin + (uint64_t)i);
                    hit.push_back(h);
                    l2.push_back(lr[i]);
                    // larger is closer to the event
                    score.push_back(reach ? (h ? 1.0 : 0.0) : (s->q.is_max ? qm[i] : -qm[i]));
                }
                done += nb;
            }
            // 2. level and elites
            const double thr = reach ? 0.5 : (s->q.is_max ? s->q.threshold : -s->q.threshold);
            std::vector<double> sorted = score;
            const size_t kq = std::min(sorted.size() - 1, (size_t)((1.0 - c.elite_frac) * (double)sorted.size()));
            std::nth_element(sorted.begin(), sorted.begin() + kq, sorted.end());
            const double gamma = sorted[kq];
            std::vector<int> elite;
            double level;
            if ((double)hits >= c.elite_frac * N || gamma > thr || reach) {
                for (int i = 0; i < N; ++i) if (hit[i]) elite.push_back(i);
                level = s->q.threshold;
            } else {
                std::vector<int> ge, gt;
                for (int i = 0; i < N; ++i) {
                    if (score[i] >= gamma) ge.push_back(i);
                    if (score[i] > gamma) gt.push_back(i);
                }
                const bool strict = ge.size() > 2 * c.elite_frac * N && gt.size() >= 0.25 * c.elite_frac * N;
                elite = strict ? gt : ge;
                level = s->q.is_max ? gamma : -gamma;
            }
            smc_is_round info{};
            info.level = level;
            info.elites = elite.size();
            info.hits = hits;
            info.estimate = sw / (double)N;
            if (elite.size() > (size_t)kMaxBatchRows) elite.resize(kMaxBatchRows);
            if (!elite.empty()) {
                // 3. CE weights: likelihood ratios of the elites, normalised
                double lmax = -INFINITY;
                for (int i : elite) lmax = std::max(lmax, l2[i]);
                std::vector<double> wv(elite.size());
                double wsum = 0.0;
                double wsq = 0.0;
                for (size_t k = 0; k < elite.size(); ++k) {
                    wv[k] = std::exp2(l2[elite[k]] - lmax);
                    wsum += wv[k];
                    wsq += wv[k] * wv[k];
                }
                info.ess = wsum * wsum / wsq;
                const int ne = (int)elite.size();
                const int rows = (int)rup((size_t)ne, 64);
                std::vector<uint64_t> eruns(rows, 0);
                std::vector<float> ew(rows, 0.f);
                for (int k = 0; k < ne; ++k) { eruns[k] = run_of[elite[k]]; ew[k] = (float)(wv[k] / wsum); }
                s->ensure_rows(rows, true);
                hip_check(hipMemcpyAsync(s->runs_in.p, eruns.data(), rows * 8, hipMemcpyHostToDevice, st), "upload elites");
                hip_check(hipMemcpyAsync(s->weight.p, ew.data(), rows * 4, hipMemcpyHostToDevice, st), "upload weights");
                // 4. replay the elites under the same proposal, recording pairs
                smc_is::Batch rb;
                rb.tilt = s->zero_output ? Tilt::NONE : Tilt::MLP;
                rb.seed = c.seed;
                rb.n_valid = ne;
                info.pairs = s->record(rb);
                // 5. epochs over the recorded pairs
                for (int e = 0; e < c.epochs && s->rec_pairs > 0; ++e) {
                    s->ce_gradient();
                    hip_check(smc::adam_step(s->w.p, s->g.p, s->m1.p, s->m2.p, s->n_params, c.learning_rate,
                                             ++s->adam_t, s->wb.p, st), "Adam step");
                    s->packed = false;
                    s->zero_output = false;
                }
                hip_check(hipStreamSynchronize(st), "training");
            }
            info.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            if (rounds) rounds[r] = info;
        }
        s->trained = true;
        return SMC_OK;
    });
}

int smc_is_record(smc_is *s, const float *theta, uint64_t seed, const uint64_t *runs, int n_runs, const float *w,
                  int early_stop, uint16_t *X, float *A, float *B, uint64_t cap, uint64_t *pairs) {
    if (!s || !runs || n_runs <= 0) return fail(SMC_E_INVALID, "null argument or no runs");
    return guarded([&]() -> int {
        hip_check(hipSetDevice(s->m->device), "hipSetDevice");
        hipStream_t st = s->stream();
        if (n_runs > kMaxBatchRows) throw std::runtime_error("limit: too many runs to record at once");
        const int rows = (int)rup((size_t)n_runs, 64);
        s->ensure_rows(rows, theta == nullptr);
        std::vector<uint64_t> rr(rows, 0);
        std::vector<float> ww(rows, 0.f);
        for (int i = 0; i < n_runs; ++i) { rr[i] = runs[i]; ww[i] = w ? w[i] : 1.f; }
        hip_check(hipMemcpyAsync(s->runs_in.p, rr.data(), rows * 8, hipMemcpyHostToDevice, st), "upload runs");
        hip_check(hipMemcpyAsync(s->weight.p, ww.data(), rows * 4, hipMemcpyHostToDevice, st), "upload weights");
        if (theta) {
            std::vector<float> th(s->Lpad, 0.f);
            memcpy(th.data(), theta, (size_t)s->n_locs * 4);
            hip_check(hipMemcpyAsync(s->theta_fixed.p, th.data(), th.size() * 4, hipMemcpyHostToDevice, st), "upload tilts");
        }
        smc_is::Batch b;
        b.tilt = theta ? Tilt::FIXED : Tilt::MLP;
        b.seed = seed;
        b.early_stop = early_stop != 0;
        b.n_valid = n_runs;
        const size_t n = s->record(b);
        if (pairs) *pairs = n;
        const size_t k = std::min<size_t>(n, cap);
        if (X) hip_check(hipMemcpyAsync(X, s->rX.p, k * s->fpad * 2, hipMemcpyDeviceToHost, st), "copy records");
        if (A) hip_check(hipMemcpyAsync(A, s->rA.p, k * s->Lpad * 4, hipMemcpyDeviceToHost, st), "copy records");
        if (B) hip_check(hipMemcpyAsync(B, s->rB.p, k * s->Lpad * 4, hipMemcpyDeviceToHost, st), "copy records");
        hip_check(hipStreamSynchronize(st), "copy records");
        return SMC_OK;
    });
}

int smc_is_ce_gradient(smc_is *s, float *grad_out) {
    if (!s || !grad_out) return fail(SMC_E_INVALID, "null argument");
    return guarded([&]() -> int {
        hip_check(hipSetDevice(s->m->device), "hipSetDevice");
        s->ensure_records();
        s->ce_gradient();
        hip_check(hipMemcpyAsync(grad_out, s->g.p, s->n_params * 4, hipMemcpyDeviceToHost, s->stream()), "copy gradient");
        hip_check(hipStreamSynchronize(s->stream()), "gradient");
        return SMC_OK;
    });
}

int smc_is_estimate(smc_is *s, uint64_t seed, uint64_t run_begin, uint64_t batch_runs, uint64_t max_runs,
                    double rel_target, smc_is_result *out) {
    if (!s || !out) return fail(SMC_E_INVALID, "null argument");
    return guarded([&]() -> int {
        auto t0 = std::chrono::steady_clock::now();
        hip_check(hipSetDevice(s->m->device), "hipSetDevice");
        if (batch_runs == 0) batch_runs = 65536;
        if (batch_runs > (uint64_t)kMaxBatchRows) batch_runs = kMaxBatchRows;
        if (max_runs == 0) max_runs = 1ull << 32;
        if (!(rel_target > 0.0)) rel_target = 0.1;
        smc_is::Batch b;
        b.tilt = Tilt::MLP;
        b.seed = seed;
        double sw = 0, sw2 = 0;
        uint64_t hits = 0, steps = 0, trans = 0, done = 0;
        std::vector<uint8_t> stat;
        std::vector<double> lr;
        std::vector<uint32_t> stp, trn;
        while (done < max_runs) {
            const int nb = (int)std::min<uint64_t>(batch_runs, max_runs - done);
            b.run_begin = run_begin + done;
            b.n_valid = nb;
            s->run_rows(b);
            s->fetch(stat, s->out.status, nb);
            s->fetch(lr, s->out.log2lr, nb);
            s->fetch(stp, s->out.step, nb);
            s->fetch(trn, s->out.trans, nb);
            hip_check(hipStreamSynchronize(s->stream()), "copy rows");
            accumulate(stat, lr, stp, trn, nb, &sw, &sw2, &hits, &steps, &trans);
            done += (uint64_t)nb;
            smc_is_result tmp;
            fill_is_result(done, hits, sw, sw2, steps, trans, 0.0, &tmp);
            if (hits >= 10 && tmp.rel_half_width <= rel_target) break;
        }
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        fill_is_result(done, hits, sw, sw2, steps, trans, ms, out);
        return SMC_OK;
    });
}

}  // extern "C"
