// capi.cpp — the C ABI (include/smc_capi.h).
#include "smc_capi.h"

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "capi_internal.hpp"
#include "noalloc.hpp"
#include "uppaal.hpp"

namespace smc::capi {
thread_local std::string g_last_error;
}  // namespace smc::capi
using namespace smc::capi;

struct smc_parsed {
    std::unique_ptr<smc::Parsed> p;
    smc_network view{};
};

namespace {

// Launch shape for one group of instances.
struct Plan {
    int S = 1, C = 1, R = 64, stride = 1;
    int occ = 0;  // run-kernel build: 0 registers unrestricted, 2 two waves per SIMD
    bool tab = false;     // tables staged in LDS
    int group_waves = 1;  // waves per workgroup
    uint32_t table_bytes = 0, wave_lds = 0;
    size_t lds = 0;
    int waves = 1;
    uint32_t chunk_runs = 64;
    uint64_t total_chunks = 0;
};

constexpr size_t kLdsPerCu = 160 * 1024;

Plan plan_launch(int P, int max_vars, int lanes, int occ, int tables, uint32_t table_bytes, int waves_req,
                 int waves_per_cu, uint64_t total_runs, uint64_t max_runs_per_inst, bool trace, int cu_count) {
    Plan pl;
    if (occ != 0 && occ != 2) throw std::runtime_error("occupancy must be 0 or 2");
    if (tables != 0 && tables != 1) throw std::runtime_error("tables must be 0 or 1");
    pl.occ = occ;
    int S = lanes ? lanes : (P < 64 ? P : 64);
    if (S <= 0 || S > 64 || (S & (S - 1))) throw std::runtime_error("lanes_per_run must be a power of two <= 64");
    if (S > P) S = P;
    pl.S = S;
    pl.C = P / S;
    if (pl.C > 16) throw std::runtime_error("limit: more than 16 components per lane; use more lanes per run");
    pl.R = 64 / S;
    pl.stride = (max_vars > 0 ? max_vars : 1) | 1;
    // per wave: variables of the R packed runs, then two doubles per lane
    // for each butterfly level of the race tree
    int levels = 0;
    while ((1 << levels) < S) ++levels;
    const size_t wave = ((size_t)pl.R * pl.stride * sizeof(double) + (size_t)2 * levels * 64 * sizeof(double) + 15) & ~(size_t)15;
    if (wave > 60000) throw std::runtime_error("limit: variables of the packed runs exceed the LDS budget");
    pl.wave_lds = (uint32_t)wave;
    pl.tab = tables == 1 && !trace;  // digest runs (parity checks) read the tables from global memory
    if (pl.tab) {
        if (table_bytes == 0 || table_bytes % 16) throw std::runtime_error("internal: table size");
        pl.table_bytes = table_bytes;
        pl.group_waves = SMC_MAX_GROUP_WAVES;
        if (table_bytes + (size_t)pl.group_waves * wave > kLdsPerCu)
            throw std::runtime_error("limit: the model tables do not fit in LDS");
    }
    pl.lds = pl.table_bytes + (size_t)pl.group_waves * wave;
    int waves = waves_req;
    if (waves <= 0) {
        int blocks = 0;
        hip_check(smc::sim_occupancy(pl.C, trace, pl.occ, pl.tab, pl.group_waves, pl.lds, &blocks), "occupancy query");
        int per_cu = blocks * pl.group_waves;
        if (waves_per_cu > 0 && waves_per_cu < per_cu) per_cu = waves_per_cu;
        if (per_cu <= 0) {
            if (pl.tab) throw std::runtime_error("limit: the model tables do not fit in LDS");
            per_cu = 1;
        }
        waves = cu_count * per_cu;
    }
    waves = (waves + pl.group_waves - 1) / pl.group_waves * pl.group_waves;
    // chunks: enough per wave to balance uneven run lengths, at most 64 rounds of the wave's segments
    uint64_t per_wave = (total_runs + (uint64_t)waves - 1) / (uint64_t)waves;
    uint64_t k = (per_wave + 4ull * pl.R - 1) / (4ull * pl.R);
    if (k < 1) k = 1;
    if (k > 64) k = 64;
    uint64_t chunk = k * pl.R;
    if (chunk > max_runs_per_inst && max_runs_per_inst > 0) chunk = max_runs_per_inst;
    if (chunk < 1) chunk = 1;
    pl.chunk_runs = (uint32_t)chunk;
    pl.waves = waves;
    return pl;
}

void validate_query(const smc_query *q, int n_vars) {
    if (q->kind != SMC_QUERY_REACH && q->kind != SMC_QUERY_VAR) throw std::runtime_error("unknown query kind");
    if (q->kind == SMC_QUERY_VAR && (q->var < 0 || q->var >= n_vars))
        throw std::runtime_error("query variable id out of range");
    if (!(q->time_bound >= 0.0)) throw std::runtime_error("time bound must be a non-negative number");
}

SimParams make_params(const smc_query *q, const smc::ModelImage &img, uint64_t seed, uint64_t run_begin,
                      uint64_t n_runs, const RunBuffers &b, bool want_time, bool trace) {
    SimParams p{};
    p.seed = seed;
    p.run_begin = run_begin;
    p.n_runs = n_runs;
    p.bound = q->time_bound;
    p.threshold = q->threshold;
    p.reach = q->kind == SMC_QUERY_REACH;
    p.qvar = q->kind == SMC_QUERY_VAR ? q->var : -1;
    p.qinit = (p.qvar >= 0 && img.var_kind[p.qvar] == SMC_VAR_INT) ? img.var_init[p.qvar] : 0.0;
    p.is_max = q->is_max ? 1 : 0;
    p.max_steps = q->max_steps ? q->max_steps : SMC_DEFAULT_MAX_STEPS;
    p.flag = b.flag;
    p.value = b.value;
    p.steps = b.steps;
    p.trans = b.trans;
    p.digest = trace ? b.digest : nullptr;
    p.time = want_time ? b.time : nullptr;
    return p;
}

void fill_result(const smc_query *q, const ReducePartial &t, uint64_t n_runs, smc_result *out) {
    const double N = (double)n_runs;
    out->n_runs = n_runs;
    out->hits = t.hits;
    out->sum = t.sum;
    out->sumsq = t.sumsq;
    out->vmin = t.vmin;
    out->vmax = t.vmax;
    out->steps = t.steps;
    out->transitions = t.trans;
    out->truncated = t.truncated;
    out->ch_eps = smc_ch_epsilon(n_runs, 0.05);
    if (q->kind == SMC_QUERY_VAR && q->is_estimate) {
        out->estimate = t.sum / N;
        double var = n_runs > 1 ? std::max(0.0, (t.sumsq - t.sum * t.sum / N) / (N - 1.0)) : 0.0;
        out->clt_half = 1.959963984540054 * std::sqrt(var / N);
    } else {
        double pr = (double)t.hits / N;
        out->estimate = pr;
        out->clt_half = 1.959963984540054 * std::sqrt(pr * (1.0 - pr) / N);
    }
}

uint32_t reduce_blocks(uint64_t n_runs) {
    uint64_t b = (n_runs + 4095) / 4096;
    if (b < 1) b = 1;
    if (b > 1024) b = 1024;
    return (uint32_t)b;
}

}  // namespace

extern "C" {

const char *smc_last_error(void) { return g_last_error.c_str(); }
const char *smc_version(void) { return "cusmc_amd 0.1 (gfx950)"; }

double smc_ch_epsilon(uint64_t n, double delta) {
    if (n == 0 || !(delta > 0.0) || !(delta < 1.0)) return NAN;
    return std::sqrt(std::log(2.0 / delta) / (2.0 * (double)n));
}

// ------------------------------------------------------------ front end

static int parse_common(bool from_file, const char *src, const char *loc_query, smc_parsed **out) {
    if (!src || !out) return fail(SMC_E_INVALID, "null argument");
    *out = nullptr;
    try {
        auto res = std::make_unique<smc_parsed>();
        std::string q = loc_query ? loc_query : "";
        res->p = from_file ? smc::parse_uppaal_file(src, q) : smc::parse_uppaal_string(src, q);
        res->view = res->p->net.view();
        *out = res.release();
        return SMC_OK;
    } catch (const std::bad_alloc &) {
        return fail(SMC_E_ALLOC, "out of host memory");
    } catch (const std::exception &e) {
        return fail(SMC_E_PARSE, e.what());
    }
}

int smc_parse_uppaal(const char *path, const char *loc_query, smc_parsed **out) {
    return parse_common(true, path, loc_query, out);
}
int smc_parse_uppaal_string(const char *xml, const char *loc_query, smc_parsed **out) {
    return parse_common(false, xml, loc_query, out);
}
const smc_network *smc_parsed_network(const smc_parsed *p) { return p ? &p->view : nullptr; }
static const char *name_at(const std::vector<std::string> &v, int i) {
    return (i >= 0 && i < (int)v.size()) ? v[i].c_str() : "";
}
const char *smc_parsed_component_name(const smc_parsed *p, int c) { return p ? name_at(p->p->net.comp_names, c) : ""; }
const char *smc_parsed_location_name(const smc_parsed *p, int l) { return p ? name_at(p->p->net.loc_names, l) : ""; }
const char *smc_parsed_variable_name(const smc_parsed *p, int v) { return p ? name_at(p->p->net.var_names, v) : ""; }
int smc_parsed_goal_count(const smc_parsed *p) { return p ? p->p->goal_count : 0; }
void smc_parsed_destroy(smc_parsed *p) { delete p; }

// ------------------------------------------------------------ models

int smc_model_create(const smc_network *net, int device, smc_model **out) {
    if (!net || !out) return fail(SMC_E_INVALID, "null argument");
    *out = nullptr;
    return guarded([&]() -> int {
        auto m = std::make_unique<smc_model>();
        m->img = smc::compile_model(*net);
        if (m->img.n_leaves > 1024) throw std::runtime_error("limit: more than 1024 components per network");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
            throw HipError("no HIP device available");
        if (device < 0 || device >= ndev) throw std::runtime_error("device index out of range");
        m->device = device;
        hip_check(hipSetDevice(device), "hipSetDevice");
        hipDeviceProp_t prop;
        hip_check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
        m->cu_count = prop.multiProcessorCount;
        hip_check(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking), "hipStreamCreate");
        for (auto &e : m->ev) hip_check(hipEventCreate(&e), "hipEventCreate");

        // one arena for every table
        const smc::ModelImage &im = m->img;
        struct Part { const void *src; size_t bytes; size_t off; };
        std::vector<Part> parts = {
            {im.locs.data(), im.locs.size() * sizeof(DevLoc), 0},
            {im.invs.data(), im.invs.size() * sizeof(DevInv), 0},
            {im.edges.data(), im.edges.size() * sizeof(DevEdge), 0},
            {im.guards.data(), im.guards.size() * sizeof(DevGuard), 0},
            {im.upds.data(), im.upds.size() * sizeof(DevUpd), 0},
            {im.consts.data(), im.consts.size() * sizeof(double), 0},
            {im.code.data(), im.code.size() * sizeof(DevInsn), 0},
            {im.var_init.data(), im.var_init.size() * sizeof(double), 0},
            {im.comp_init_loc.data(), im.comp_init_loc.size() * sizeof(int32_t), 0},
            {im.clocks.data(), im.clocks.size() * sizeof(uint16_t), 0},
            {im.comp_first_loc.data(), im.comp_first_loc.size() * sizeof(int32_t), 0},
        };
        size_t total = 0;
        for (auto &p : parts) { p.off = total; total = align_up(total + p.bytes, 256); }
        std::vector<uint8_t> host(total, 0);
        for (auto &p : parts) if (p.bytes) memcpy(host.data() + p.off, p.src, p.bytes);
        hip_check(hipMalloc(&m->arena, total), "hipMalloc model arena");
        hip_check(hipMemcpy(m->arena, host.data(), total, hipMemcpyHostToDevice), "upload model");
        auto at = [&](int i) { return (const void *)((const uint8_t *)m->arena + parts[i].off); };
        m->table_bytes = (uint32_t)parts[7].off;  // locs .. code: what the run kernel reads every step
        DevModel &d = m->dm;
        d.locs = (decltype(d.locs))at(0);
        d.invs = (decltype(d.invs))at(1);
        d.edges = (decltype(d.edges))at(2);
        d.guards = (decltype(d.guards))at(3);
        d.upds = (decltype(d.upds))at(4);
        d.consts = (decltype(d.consts))at(5);
        d.code = (decltype(d.code))at(6);
        d.var_init = (decltype(d.var_init))at(7);
        d.comp_init_loc = (decltype(d.comp_init_loc))at(8);
        d.clocks = (decltype(d.clocks))at(9);
        d.comp_first_loc = (decltype(d.comp_first_loc))at(10);
        d.n_comp = im.n_comp;
        d.n_vars = im.n_vars;
        int nclk = 0;
        for (int k : im.var_kind) nclk += k == SMC_VAR_CLOCK;
        d.n_clocks = nclk;
        d.n_leaves = im.n_leaves;
        d.n_locs = (int)im.locs.size();
        d.n_edges = net->n_edges;
        d.freeze = im.freeze ? 1 : 0;
        d.guarded = im.guarded ? 1 : 0;
        hip_check(hipMalloc(&m->d_inst, sizeof(InstDesc)), "hipMalloc instance");
        hip_check(hipMalloc(&m->d_counter, sizeof(unsigned long long)), "hipMalloc counter");
        hip_check(hipMalloc(&m->d_partials, 1024 * sizeof(ReducePartial)), "hipMalloc partials");
        hip_check(hipMalloc(&m->d_acc, sizeof(ReducePartial)), "hipMalloc totals");
        *out = m.release();
        return SMC_OK;
    });
}

void smc_model_destroy(smc_model *m) { delete m; }

int smc_run(smc_model *m, const smc_query *q, uint64_t seed, uint64_t run_begin, uint64_t n_runs,
            const smc_launch *launch, smc_run_outputs *outputs, smc_result *out) {
    if (!m || !q || !out) return fail(SMC_E_INVALID, "null argument");
    return guarded([&]() -> int {
        auto wall0 = std::chrono::steady_clock::now();
        memset(out, 0, sizeof *out);
        validate_query(q, m->img.n_vars);
        if (n_runs == 0) throw std::runtime_error("n_runs must be positive");
        if (n_runs > (1ull << 32) - 1) throw std::runtime_error("limit: at most 2^32-1 runs per call");
        if (run_begin > (~0ull) - n_runs) throw std::runtime_error("run range overflows");
        const bool trace = outputs && outputs->digest;
        const smc_launch *use =
            (launch && (launch->lanes_per_run || launch->waves || launch->waves_per_cu || launch->occupancy ||
                        launch->tables))
                ? launch : (m->has_tuned ? &m->tuned : nullptr);
        hip_check(hipSetDevice(m->device), "hipSetDevice");
        Plan pl = plan_launch(m->img.n_leaves, m->img.n_vars, use ? use->lanes_per_run : 0, use ? use->occupancy : 0,
                              use ? use->tables : 0, m->table_bytes, use ? use->waves : 0,
                              use ? use->waves_per_cu : 0, n_runs, n_runs, trace, m->cu_count);
        m->runs.ensure(n_runs);
        pl.total_chunks = (n_runs + pl.chunk_runs - 1) / pl.chunk_runs;
        int waves = pl.waves;
        if ((uint64_t)waves > pl.total_chunks)
            waves = (int)((pl.total_chunks + pl.group_waves - 1) / pl.group_waves * pl.group_waves);

        m->h_inst.m = m->dm;
        m->h_inst.p = make_params(q, m->img, seed, run_begin, n_runs, m->runs, outputs && outputs->time, trace);
        m->h_inst.chunk_first = 0;
        LaunchParams lp{};
        lp.inst = m->d_inst;
        lp.n_inst = 1;
        lp.lanes = pl.S;
        lp.var_stride = pl.stride;
        lp.chunk_runs = pl.chunk_runs;
        lp.total_chunks = pl.total_chunks;
        lp.counter = m->d_counter;
        lp.round = nullptr;
        lp.tables = pl.tab ? (decltype(lp.tables))m->arena : nullptr;
        lp.table_bytes = pl.table_bytes;
        lp.wave_lds = pl.wave_lds;
        const uint32_t blocks = reduce_blocks(n_runs);

        hip_check(hipMemcpyAsync(m->d_inst, &m->h_inst, sizeof(InstDesc), hipMemcpyHostToDevice, m->stream), "upload instance");
        hip_check(smc::launch_prep(m->d_counter, nullptr, 0, m->stream), "prep kernel launch");
        hip_check(hipEventRecord(m->ev[0], m->stream), "event");
        hip_check(smc::launch_sim(lp, pl.C, trace, pl.occ, waves, pl.group_waves, pl.lds, m->stream), "run kernel launch");
        hip_check(hipEventRecord(m->ev[1], m->stream), "event");
        hip_check(smc::launch_reduce(m->d_inst, 1, blocks, m->d_partials, m->d_acc, 0, m->stream), "reduce kernel launch");
        hip_check(hipEventRecord(m->ev[2], m->stream), "event");
        hip_check(hipMemcpyAsync(&m->h_acc, m->d_acc, sizeof(ReducePartial), hipMemcpyDeviceToHost, m->stream), "copy totals");
        if (outputs) {
            const RunBuffers &b = m->runs;
            if (outputs->flag) hip_check(hipMemcpyAsync(outputs->flag, b.flag, n_runs, hipMemcpyDeviceToHost, m->stream), "copy flags");
            if (outputs->value) hip_check(hipMemcpyAsync(outputs->value, b.value, n_runs * 8, hipMemcpyDeviceToHost, m->stream), "copy values");
            if (outputs->steps) hip_check(hipMemcpyAsync(outputs->steps, b.steps, n_runs * 4, hipMemcpyDeviceToHost, m->stream), "copy steps");
            if (outputs->transitions) hip_check(hipMemcpyAsync(outputs->transitions, b.trans, n_runs * 4, hipMemcpyDeviceToHost, m->stream), "copy transitions");
            if (outputs->digest) hip_check(hipMemcpyAsync(outputs->digest, b.digest, n_runs * 8, hipMemcpyDeviceToHost, m->stream), "copy digests");
            if (outputs->time) hip_check(hipMemcpyAsync(outputs->time, b.time, n_runs * 8, hipMemcpyDeviceToHost, m->stream), "copy times");
        }
        hip_check(hipStreamSynchronize(m->stream), "run");
        float k_ms = 0.f, r_ms = 0.f;
        hip_check(hipEventElapsedTime(&k_ms, m->ev[0], m->ev[1]), "event time");
        hip_check(hipEventElapsedTime(&r_ms, m->ev[1], m->ev[2]), "event time");
        fill_result(q, m->h_acc, n_runs, out);
        out->kernel_ms = k_ms;
        out->reduce_ms = r_ms;
        out->lanes_per_run = pl.S;
        out->waves = waves;
        out->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
        return SMC_OK;
    });
}

int smc_run_noalloc(smc_model *m, const smc_query *q, uint64_t seed, uint64_t run_begin, uint64_t n_runs,
                    const smc_launch *launch, smc_result *out) {
    if (!m || !q || !out) return fail(SMC_E_INVALID, "null argument");
    return guarded([&]() -> int {
        auto wall0 = std::chrono::steady_clock::now();
        memset(out, 0, sizeof *out);
        validate_query(q, m->img.n_vars);
        if (n_runs == 0) throw std::runtime_error("n_runs must be positive");
        if (n_runs > SMC_NOALLOC_MAX_RUNS)
            throw std::runtime_error("limit: smc_run_noalloc takes at most 1048576 runs per call (its scratch is fixed)");
        if (run_begin > (~0ull) - n_runs) throw std::runtime_error("run range overflows");
        const smc_launch *use =
            (launch && (launch->lanes_per_run || launch->waves || launch->waves_per_cu || launch->occupancy ||
                        launch->tables))
                ? launch : (m->has_tuned ? &m->tuned : nullptr);
        hip_check(hipSetDevice(m->device), "hipSetDevice");
        void *scr = nullptr;
        hip_check(smc::noalloc_scratch(m->device, &scr), "scratch address");
        uint8_t *base = (uint8_t *)scr;
        Plan pl = plan_launch(m->img.n_leaves, m->img.n_vars, use ? use->lanes_per_run : 0, use ? use->occupancy : 0,
                              use ? use->tables : 0, m->table_bytes, use ? use->waves : 0,
                              use ? use->waves_per_cu : 0, n_runs, n_runs, false, m->cu_count);
        pl.total_chunks = (n_runs + pl.chunk_runs - 1) / pl.chunk_runs;
        int waves = pl.waves;
        if ((uint64_t)waves > pl.total_chunks)
            waves = (int)((pl.total_chunks + pl.group_waves - 1) / pl.group_waves * pl.group_waves);

        // the per-run buffers are the scratch's (the whole run range of
        // this call is written by the run kernel before the reduction reads)
        RunBuffers b;
        b.flag = base + smc::kNoallocFlag;
        b.value = (double *)(base + smc::kNoallocValue);
        b.steps = (uint32_t *)(base + smc::kNoallocSteps);
        b.trans = (uint32_t *)(base + smc::kNoallocTrans);
        InstDesc *d_inst = (InstDesc *)(base + smc::kNoallocInst);
        unsigned long long *d_counter = (unsigned long long *)(base + smc::kNoallocCounter);
        ReducePartial *d_acc = (ReducePartial *)(base + smc::kNoallocAcc);
        ReducePartial *d_partials = (ReducePartial *)(base + smc::kNoallocPartials);
        const uint32_t blocks = reduce_blocks(n_runs);
        if (blocks > smc::kNoallocPartialsMax) throw std::runtime_error("internal: reduction blocks");

        m->h_inst.m = m->dm;
        m->h_inst.p = make_params(q, m->img, seed, run_begin, n_runs, b, false, false);
        m->h_inst.chunk_first = 0;
        LaunchParams lp{};
        lp.inst = d_inst;
        lp.n_inst = 1;
        lp.lanes = pl.S;
        lp.var_stride = pl.stride;
        lp.chunk_runs = pl.chunk_runs;
        lp.total_chunks = pl.total_chunks;
        lp.counter = d_counter;
        lp.round = nullptr;
        lp.tables = pl.tab ? (decltype(lp.tables))m->arena : nullptr;
        lp.table_bytes = pl.table_bytes;
        lp.wave_lds = pl.wave_lds;

        hip_check(hipMemcpyAsync(d_inst, &m->h_inst, sizeof(InstDesc), hipMemcpyHostToDevice, m->stream), "upload instance");
        hip_check(smc::launch_prep(d_counter, nullptr, 0, m->stream), "prep kernel launch");
        hip_check(hipEventRecord(m->ev[0], m->stream), "event");
        hip_check(smc::launch_sim(lp, pl.C, false, pl.occ, waves, pl.group_waves, pl.lds, m->stream), "run kernel launch");
        hip_check(hipEventRecord(m->ev[1], m->stream), "event");
        hip_check(smc::launch_reduce(d_inst, 1, blocks, d_partials, d_acc, 0, m->stream), "reduce kernel launch");
        hip_check(hipEventRecord(m->ev[2], m->stream), "event");
        hip_check(hipMemcpyAsync(&m->h_acc, d_acc, sizeof(ReducePartial), hipMemcpyDeviceToHost, m->stream), "copy totals");
        hip_check(hipStreamSynchronize(m->stream), "run");
        float k_ms = 0.f, r_ms = 0.f;
        hip_check(hipEventElapsedTime(&k_ms, m->ev[0], m->ev[1]), "event time");
        hip_check(hipEventElapsedTime(&r_ms, m->ev[1], m->ev[2]), "event time");
        fill_result(q, m->h_acc, n_runs, out);
        out->kernel_ms = k_ms;
        out->reduce_ms = r_ms;
        out->lanes_per_run = pl.S;
        out->waves = waves;
        out->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
        return SMC_OK;
    });
}

int smc_autotune(smc_model *m, const smc_query *q, uint64_t seed, uint64_t sample_runs, smc_launch *best) {
    if (!m || !q) return fail(SMC_E_INVALID, "null argument");
    return guarded([&]() -> int {
        if (sample_runs == 0) sample_runs = 65536;
        const int P = m->img.n_leaves;
        std::vector<smc_launch> cands;
        for (int S = P < 64 ? P : 64; S >= 1; S >>= 1) {
            if (P / S > 16) break;
            for (int occ : {0, 2})
                for (int tab : {0, 1}) cands.push_back(smc_launch{S, 0, 0, occ, tab});
        }
        // each candidate runs at least twice the runs its grid holds at once,
        // so the sample measures the steady state rather than the tail
        double best_cost = 1e300;
        smc_launch pick = cands.front();
        for (const smc_launch &c : cands) {
            uint64_t n = sample_runs;
            try {
                Plan pl = plan_launch(P, m->img.n_vars, c.lanes_per_run, c.occupancy, c.tables, m->table_bytes, 0, 0,
                                      sample_runs, sample_runs, false, m->cu_count);
                n = std::max<uint64_t>(n, 2ull * (uint64_t)pl.waves * (uint64_t)pl.R);
            } catch (const std::exception &) {
                continue;  // e.g. LDS budget exceeded at this packing
            }
            smc_result r;
            int rc = smc_run(m, q, seed, 0, std::min<uint64_t>(n, 4096), &c, nullptr, &r);  // warm
            if (rc != SMC_OK) continue;
            rc = smc_run(m, q, seed + 1, 0, n, &c, nullptr, &r);
            const double cost = r.kernel_ms / (double)n;
            if (rc == SMC_OK && cost < best_cost) { best_cost = cost; pick = c; }
        }
        if (!(best_cost < 1e300)) throw std::runtime_error("no launch shape could run this model");
        m->tuned = pick;
        m->has_tuned = true;
        if (best) *best = pick;
        return SMC_OK;
    });
}

// ------------------------------------------------------------ batches

}  // extern "C"

struct smc_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<hipStream_t> gstreams;
    std::vector<hipEvent_t> joins;
    hipEvent_t fork = nullptr, ev0 = nullptr, ev1 = nullptr;
    std::vector<smc_batch_item> items;
    std::vector<RunBuffers> bufs;
    struct Group {
        Plan pl;
        std::vector<int> members;
        InstDesc *d_inst = nullptr;
        unsigned long long *d_counter = nullptr;
        LaunchParams lp{};
    };
    std::vector<Group> groups;
    InstDesc *d_all = nullptr;
    ReducePartial *d_partials = nullptr, *d_acc = nullptr;
    uint64_t *d_round = nullptr;
    uint32_t blocks = 1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    ~smc_batch() {
        hipSetDevice(device);
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        for (auto &b : bufs) b.release();
        for (auto &g : groups) { if (g.d_inst) hipFree(g.d_inst); if (g.d_counter) hipFree(g.d_counter); }
        for (void *p : {(void *)d_all, (void *)d_partials, (void *)d_acc, (void *)d_round}) if (p) hipFree(p);
        for (auto e : joins) hipEventDestroy(e);
        for (hipEvent_t e : {fork, ev0, ev1}) if (e) hipEventDestroy(e);
        for (auto s : gstreams) hipStreamDestroy(s);
        if (stream) hipStreamDestroy(stream);
    }
};

extern "C" {

int smc_batch_create(const smc_batch_item *items, int n_items, int tune, smc_batch **out) {
    if (!items || n_items <= 0 || !out) return fail(SMC_E_INVALID, "null argument or empty batch");
    *out = nullptr;
    return guarded([&]() -> int {
        auto b = std::make_unique<smc_batch>();
        b->items.assign(items, items + n_items);
        b->device = items[0].model ? items[0].model->device : 0;
        for (int i = 0; i < n_items; ++i) {
            const smc_batch_item &it = items[i];
            if (!it.model) throw std::runtime_error("batch item without a model");
            if (it.model->device != b->device) throw std::runtime_error("batch items on different devices");
            validate_query(&it.query, it.model->img.n_vars);
            if (it.n_runs == 0 || it.n_runs > (1ull << 32) - 1) throw std::runtime_error("batch item n_runs out of range");
        }
        hip_check(hipSetDevice(b->device), "hipSetDevice");
        hip_check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
        for (hipEvent_t *e : {&b->fork, &b->ev0, &b->ev1}) hip_check(hipEventCreate(e), "hipEventCreate");
        const int cu = items[0].model->cu_count;

        // group by launch shape: tuned (per size class) or default
        std::map<std::tuple<int, int, int>, std::vector<int>> by_shape;  // (P, S, occupancy) -> members
        std::map<int, smc_launch> tuned;                            // P -> lanes, occupancy
        if (tune) {
            std::map<int, int> rep;  // P -> item with the most components... first seen
            for (int i = 0; i < n_items; ++i) {
                const int P = items[i].model->img.n_leaves;
                auto it = rep.find(P);
                if (it == rep.end() || items[i].model->img.n_comp > items[it->second].model->img.n_comp) rep[P] = i;
            }
            for (auto &kv : rep) {
                const smc_batch_item &it = items[kv.second];
                smc_launch best{};
                uint64_t sample = std::min<uint64_t>(it.n_runs, 16384);
                if (smc_autotune(it.model, &it.query, it.seed, sample, &best) != SMC_OK)
                    throw std::runtime_error(std::string("autotune failed: ") + g_last_error);
                tuned[kv.first] = best;
            }
        }
        for (int i = 0; i < n_items; ++i) {
            const smc_model *m = items[i].model;
            const int P = m->img.n_leaves;
            const smc_launch *t = tuned.count(P) ? &tuned[P] : (m->has_tuned ? &m->tuned : nullptr);
            int S = t ? t->lanes_per_run : 0;
            if (!S) S = P < 64 ? P : 64;
            by_shape[{P, S, t ? t->occupancy : 0}].push_back(i);
        }
        b->bufs.resize(n_items);
        std::vector<InstDesc> all(n_items);
        for (auto &kv : by_shape) {
            smc_batch::Group g;
            g.members = kv.second;
            uint64_t total = 0, maxr = 0;
            int max_vars = 1;
            for (int i : g.members) {
                total += items[i].n_runs;
                maxr = std::max<uint64_t>(maxr, items[i].n_runs);
                max_vars = std::max(max_vars, items[i].model->img.n_vars);
            }
            g.pl = plan_launch(std::get<0>(kv.first), max_vars, std::get<1>(kv.first), std::get<2>(kv.first), 0, 0, 0,
                               0, total, maxr, false, cu);
            std::vector<InstDesc> inst;
            uint64_t chunk = 0;
            for (int i : g.members) {
                const smc_batch_item &it = items[i];
                b->bufs[i].ensure(it.n_runs);
                InstDesc d{};
                d.m = it.model->dm;
                d.p = make_params(&it.query, it.model->img, it.seed, it.run_begin, it.n_runs, b->bufs[i], false, false);
                d.chunk_first = chunk;
                chunk += (it.n_runs + g.pl.chunk_runs - 1) / g.pl.chunk_runs;
                inst.push_back(d);
                all[i] = d;
            }
            g.pl.total_chunks = chunk;
            if ((uint64_t)g.pl.waves > chunk) g.pl.waves = (int)chunk;
            hip_check(hipMalloc(&g.d_inst, inst.size() * sizeof(InstDesc)), "hipMalloc group");
            hip_check(hipMemcpy(g.d_inst, inst.data(), inst.size() * sizeof(InstDesc), hipMemcpyHostToDevice), "upload group");
            hip_check(hipMalloc(&g.d_counter, sizeof(unsigned long long)), "hipMalloc counter");
            hip_check(hipMemset(g.d_counter, 0, sizeof(unsigned long long)), "clear counter");
            b->groups.push_back(g);
        }
        uint64_t maxr = 0;
        for (auto &it : b->items) maxr = std::max<uint64_t>(maxr, it.n_runs);
        b->blocks = std::min<uint32_t>(reduce_blocks(maxr), 64);
        hip_check(hipMalloc(&b->d_all, n_items * sizeof(InstDesc)), "hipMalloc instances");
        hip_check(hipMemcpy(b->d_all, all.data(), n_items * sizeof(InstDesc), hipMemcpyHostToDevice), "upload instances");
        hip_check(hipMalloc(&b->d_partials, (size_t)n_items * b->blocks * sizeof(ReducePartial)), "hipMalloc partials");
        hip_check(hipMalloc(&b->d_acc, n_items * sizeof(ReducePartial)), "hipMalloc totals");
        hip_check(hipMalloc(&b->d_round, sizeof(uint64_t)), "hipMalloc round");
        hip_check(hipMemset(b->d_round, 0, sizeof(uint64_t)), "clear round");
        for (auto &g : b->groups) {
            g.lp.inst = g.d_inst;
            g.lp.n_inst = (int32_t)g.members.size();
            g.lp.lanes = g.pl.S;
            g.lp.var_stride = g.pl.stride;
            g.lp.chunk_runs = g.pl.chunk_runs;
            g.lp.total_chunks = g.pl.total_chunks;
            g.lp.counter = g.d_counter;
            g.lp.round = b->d_round;
            g.lp.tables = nullptr;
            g.lp.table_bytes = 0;
            g.lp.wave_lds = g.pl.wave_lds;
        }
        b->gstreams.resize(b->groups.size());
        b->joins.resize(b->groups.size());
        for (size_t k = 0; k < b->groups.size(); ++k) {
            hip_check(hipStreamCreateWithFlags(&b->gstreams[k], hipStreamNonBlocking), "hipStreamCreate");
            hip_check(hipEventCreateWithFlags(&b->joins[k], hipEventDisableTiming), "hipEventCreate");
        }
        hip_check(hipDeviceSynchronize(), "batch setup");

        // capture one round: groups run concurrently, then the reduction
        // accumulates, then the counters are cleared and the round advances
        hip_check(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "begin capture");
        hip_check(hipEventRecord(b->fork, b->stream), "fork");
        for (size_t k = 0; k < b->groups.size(); ++k) {
            auto &g = b->groups[k];
            hip_check(hipStreamWaitEvent(b->gstreams[k], b->fork, 0), "fork wait");
            hip_check(smc::launch_sim(g.lp, g.pl.C, false, g.pl.occ, g.pl.waves, g.pl.group_waves, g.pl.lds, b->gstreams[k]),
                      "capture run kernel");
            hip_check(hipEventRecord(b->joins[k], b->gstreams[k]), "join");
            hip_check(hipStreamWaitEvent(b->stream, b->joins[k], 0), "join wait");
        }
        hip_check(smc::launch_reduce(b->d_all, n_items, b->blocks, b->d_partials, b->d_acc, 1, b->stream), "capture reduce");
        for (size_t k = 0; k < b->groups.size(); ++k)
            hip_check(smc::launch_prep(b->groups[k].d_counter, k == 0 ? b->d_round : nullptr, k == 0 ? 1 : 0, b->stream),
                      "capture prep");
        hip_check(hipStreamEndCapture(b->stream, &b->graph), "end capture");
        hip_check(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "instantiate graph");
        *out = b.release();
        return SMC_OK;
    });
}

int smc_batch_run(smc_batch *b, int rounds, smc_result *results, smc_batch_stats *stats) {
    if (!b || rounds <= 0) return fail(SMC_E_INVALID, "null batch or non-positive rounds");
    return guarded([&]() -> int {
        const int n = (int)b->items.size();
        hip_check(hipSetDevice(b->device), "hipSetDevice");
        std::vector<ReducePartial> init(n);
        for (auto &t : init) { t = ReducePartial{}; t.vmin = HUGE_VAL; t.vmax = -HUGE_VAL; }
        hip_check(hipMemcpyAsync(b->d_acc, init.data(), n * sizeof(ReducePartial), hipMemcpyHostToDevice, b->stream), "init totals");
        hip_check(hipEventRecord(b->ev0, b->stream), "event");
        for (int r = 0; r < rounds; ++r) hip_check(hipGraphLaunch(b->exec, b->stream), "graph launch");
        hip_check(hipEventRecord(b->ev1, b->stream), "event");
        std::vector<ReducePartial> acc(n);
        hip_check(hipMemcpyAsync(acc.data(), b->d_acc, n * sizeof(ReducePartial), hipMemcpyDeviceToHost, b->stream), "copy totals");
        hip_check(hipStreamSynchronize(b->stream), "batch run");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, b->ev0, b->ev1), "event time");
        uint64_t runs = 0, steps = 0, trans = 0;
        for (int i = 0; i < n; ++i) {
            const uint64_t nr = b->items[i].n_runs * (uint64_t)rounds;
            runs += nr;
            steps += acc[i].steps;
            trans += acc[i].trans;
            if (results) {
                memset(&results[i], 0, sizeof(smc_result));
                fill_result(&b->items[i].query, acc[i], nr, &results[i]);
                results[i].kernel_ms = ms;
            }
        }
        if (stats) {
            stats->graph_ms = ms;
            stats->runs = runs;
            stats->steps = steps;
            stats->transitions = trans;
            stats->groups = (int32_t)b->groups.size();
            stats->rounds = rounds;
        }
        return SMC_OK;
    });
}

void smc_batch_destroy(smc_batch *b) { delete b; }

}  // extern "C"
