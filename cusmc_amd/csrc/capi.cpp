// capi.cpp — the C ABI (include/smc_capi.h).
#include "smc_capi.h"

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
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
    int occ = 0;  // run-kernel build: 0 registers unrestricted, 3 three waves per SIMD
    int tab = 0;          // tables staged in LDS: 0 none, 1 all hot tables, 2 all but the locations
    int group_waves = 1;  // waves per workgroup
    uint32_t table_bytes = 0, wave_lds = 0;
    size_t lds = 0;
    int waves = 1;
    uint32_t chunk_runs = 64;
    uint64_t total_chunks = 0;
};

constexpr size_t kLdsPerCu = 160 * 1024;

Plan plan_launch(int P, int max_vars, int lanes, int occ, int tables, uint32_t table_bytes, uint32_t small_bytes,
                 int waves_req,
                 int waves_per_cu, uint64_t total_runs, uint64_t max_runs_per_inst, bool trace, int cu_count) {
    Plan pl;
    if (occ != 0 && occ != 3) throw std::runtime_error("occupancy must be 0 or 3");
    if (tables < 0 || tables > 2) throw std::runtime_error("tables must be 0, 1 or 2");
    pl.occ = occ;
    int S = lanes ? lanes : (P < 64 ? P : 64);
    if (S <= 0 || S > 64 || (S & (S - 1))) throw std::runtime_error("lanes_per_run must be a power of two <= 64");
    if (S > P) S = P;
    pl.S = S;
    pl.C = P / S;
    if (pl.C > 16) throw std::runtime_error("limit: more than 16 components per lane; use more lanes per run");
    pl.R = 64 / S;
    pl.stride = (max_vars + 1) | 1;  // clock time, then the variables
    // per wave: variables of the R packed runs, then two doubles per lane
    // for each butterfly level of the race tree
    int levels = 0;
    while ((1 << levels) < S) ++levels;
    const size_t wave = ((size_t)pl.R * pl.stride * sizeof(double) + (size_t)2 * levels * 64 * sizeof(double) + 15) & ~(size_t)15;
    if (wave > 60000) throw std::runtime_error("limit: variables of the packed runs exceed the LDS budget");
    pl.wave_lds = (uint32_t)wave;
    pl.tab = trace ? 0 : tables;  // digest runs (parity checks) read the tables from global memory
    if (pl.tab) {
        const uint32_t bytes = pl.tab == 1 ? table_bytes : small_bytes;
        if (bytes == 0 || bytes % 16) throw std::runtime_error("internal: table size");
        pl.table_bytes = bytes;
        pl.group_waves = SMC_MAX_GROUP_WAVES;
        if (bytes + (size_t)pl.group_waves * wave > kLdsPerCu)
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

        // one arena for every table; variables start in their stored form
        // (a clock x = v0 as the reset stamp -v0)
        const smc::ModelImage &im = m->img;
        std::vector<double> var_start(im.var_init);
        for (size_t v = 0; v < var_start.size() && v < im.var_clock.size(); ++v)
            if (im.var_clock[v]) var_start[v] = -var_start[v];
        struct Part { const void *src; size_t bytes; size_t off; };
        std::vector<Part> parts = {
            {im.locs.data(), im.locs.size() * sizeof(DevLoc), 0},
            {im.invs.data(), im.invs.size() * sizeof(DevInv), 0},
            {im.edges.data(), im.edges.size() * sizeof(DevEdge), 0},
            {im.guards.data(), im.guards.size() * sizeof(DevGuard), 0},
            {im.upds.data(), im.upds.size() * sizeof(DevUpd), 0},
            {im.consts.data(), im.consts.size() * sizeof(double), 0},
            {im.code.data(), im.code.size() * sizeof(DevInsn), 0},
            {var_start.data(), var_start.size() * sizeof(double), 0},
            {im.comp_init_loc.data(), im.comp_init_loc.size() * sizeof(int32_t), 0},
            {im.var_clock.data(), im.var_clock.size() * sizeof(uint8_t), 0},
            {im.comp_first_loc.data(), im.comp_first_loc.size() * sizeof(int32_t), 0},
            {im.loc_edge0.data(), im.loc_edge0.size() * sizeof(int32_t), 0},
        };
        size_t total = 0;
        for (auto &p : parts) { p.off = total; total = align_up(total + p.bytes, 256); }
        std::vector<uint8_t> host(total, 0);
        for (auto &p : parts) if (p.bytes) memcpy(host.data() + p.off, p.src, p.bytes);
        hip_check(hipMalloc(&m->arena, total), "hipMalloc model arena");
        hip_check(hipMemcpy(m->arena, host.data(), total, hipMemcpyHostToDevice), "upload model");
        auto at = [&](int i) { return (const void *)((const uint8_t *)m->arena + parts[i].off); };
        m->table_bytes = (uint32_t)parts[7].off;  // locs .. code: what the run kernel reads every step
        m->small_off = (uint32_t)parts[1].off;    // invs .. code
        m->small_bytes = (uint32_t)(parts[7].off - parts[1].off);
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
        d.var_clock = (decltype(d.var_clock))at(9);
        d.comp_first_loc = (decltype(d.comp_first_loc))at(10);
        d.loc_edge0 = (decltype(d.loc_edge0))at(11);
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
                              use ? use->tables : 0, m->table_bytes, m->small_bytes, use ? use->waves : 0,
                              use ? use->waves_per_cu : 0, n_runs, n_runs, trace, m->cu_count);
        m->runs.ensure(n_runs);
        pl.total_chunks = (n_runs + pl.chunk_runs - 1) / pl.chunk_runs;
        int waves = pl.waves;
        if ((uint64_t)waves > pl.total_chunks)
            waves = (int)((pl.total_chunks + pl.group_waves - 1) / pl.group_waves * pl.group_waves);

        m->h_inst.m = m->dm;
        m->h_inst.p = make_params(q, m->img, seed, run_begin, n_runs, m->runs, outputs && outputs->time, trace);
        m->h_inst.chunk_first = 0;
        m->h_inst.lanes = pl.S;
        m->h_inst.chunk_runs = pl.chunk_runs;
        LaunchParams lp{};
        lp.inst = m->d_inst;
        lp.n_inst = 1;
        lp.var_stride = pl.stride;
        lp.total_chunks = pl.total_chunks;
        lp.counter = m->d_counter;
        lp.round = nullptr;
        lp.tables = pl.tab == 1   ? (decltype(lp.tables))m->arena
                    : pl.tab == 2 ? (decltype(lp.tables))((const uint8_t *)m->arena + m->small_off)
                                  : nullptr;
        lp.table_mode = pl.tab;
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
                              use ? use->tables : 0, m->table_bytes, m->small_bytes, use ? use->waves : 0,
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
        m->h_inst.lanes = pl.S;
        m->h_inst.chunk_runs = pl.chunk_runs;
        LaunchParams lp{};
        lp.inst = d_inst;
        lp.n_inst = 1;
        lp.var_stride = pl.stride;
        lp.total_chunks = pl.total_chunks;
        lp.counter = d_counter;
        lp.round = nullptr;
        lp.tables = pl.tab == 1   ? (decltype(lp.tables))m->arena
                    : pl.tab == 2 ? (decltype(lp.tables))((const uint8_t *)m->arena + m->small_off)
                                  : nullptr;
        lp.table_mode = pl.tab;
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

}  // extern "C"

namespace {

// Persisted autotuner choices: one line per (model, query, device, library
// build, sample size) key in a text file — SMC_TUNE_CACHE names it (default
// $XDG_CACHE_HOME/cusmc_amd/tune.txt or ~/.cache/cusmc_amd/tune.txt;
// SMC_TUNE_CACHE=off disables it).
uint64_t fnv1a(uint64_t h, const void *p, size_t n) {
    const uint8_t *b = (const uint8_t *)p;
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    return h;
}
template <class T>
uint64_t fnv_vec(uint64_t h, const std::vector<T> &v) {
    return fnv1a(h, v.data(), v.size() * sizeof(T));
}

std::string tune_cache_path() {
    const char *e = std::getenv("SMC_TUNE_CACHE");
    if (e && *e) return (std::string(e) == "off" || std::string(e) == "0") ? std::string() : std::string(e);
    const char *x = std::getenv("XDG_CACHE_HOME");
    const char *home = std::getenv("HOME");
    if (x && *x) return std::string(x) + "/cusmc_amd/tune.txt";
    if (home && *home) return std::string(home) + "/.cache/cusmc_amd/tune.txt";
    return std::string();
}

uint64_t tune_key(const smc_model *m, const smc_query *q, uint64_t sample_runs) {
    const smc::ModelImage &im = m->img;
    uint64_t h = 1469598103934665603ull;
    h = fnv_vec(h, im.locs); h = fnv_vec(h, im.invs); h = fnv_vec(h, im.edges); h = fnv_vec(h, im.guards);
    h = fnv_vec(h, im.upds); h = fnv_vec(h, im.consts); h = fnv_vec(h, im.code); h = fnv_vec(h, im.var_init);
    h = fnv_vec(h, im.comp_init_loc);
    const int32_t qi[4] = {q->kind, q->var, q->is_max, q->is_estimate};
    const double qd[2] = {q->threshold, q->time_bound};
    h = fnv1a(h, qi, sizeof qi); h = fnv1a(h, qd, sizeof qd);
    h = fnv1a(h, &sample_runs, sizeof sample_runs);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, m->device) == hipSuccess) {
        h = fnv1a(h, prop.gcnArchName, strnlen(prop.gcnArchName, sizeof prop.gcnArchName));
        h = fnv1a(h, &prop.multiProcessorCount, sizeof prop.multiProcessorCount);
    }
    static const char build[] = __DATE__ " " __TIME__;  // kernels of another build may rank differently
    return fnv1a(h, build, sizeof build);
}

bool tune_lookup(const std::string &path, uint64_t key, smc_launch *out) {
    if (path.empty()) return false;
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) return false;
    unsigned long long k;
    int S, wpc, occ, tab;
    bool found = false;
    while (std::fscanf(f, "%llx %d %d %d %d", &k, &S, &wpc, &occ, &tab) == 5)
        if (k == key) { *out = smc_launch{S, 0, wpc, occ, tab}; found = true; }
    std::fclose(f);
    return found;
}

void tune_store(const std::string &path, uint64_t key, const smc_launch &c) {
    if (path.empty()) return;
    const size_t slash = path.rfind('/');
    if (slash != std::string::npos) {
        std::error_code ec;
        std::filesystem::create_directories(path.substr(0, slash), ec);
    }
    if (FILE *f = std::fopen(path.c_str(), "a")) {
        std::fprintf(f, "%016llx %d %d %d %d\n", (unsigned long long)key, c.lanes_per_run, c.waves_per_cu, c.occupancy,
                     c.tables);
        std::fclose(f);
    }
}

}  // namespace

extern "C" {

int smc_autotune(smc_model *m, const smc_query *q, uint64_t seed, uint64_t sample_runs, smc_launch *best) {
    if (!m || !q) return fail(SMC_E_INVALID, "null argument");
    return guarded([&]() -> int {
        if (sample_runs == 0) sample_runs = 65536;
        validate_query(q, m->img.n_vars);
        const std::string cache = tune_cache_path();
        const uint64_t key = tune_key(m, q, sample_runs);
        smc_launch cached{};
        if (tune_lookup(cache, key, &cached)) {
            m->tuned = cached;
            m->has_tuned = true;
            if (best) *best = cached;
            return SMC_OK;
        }
        const int P = m->img.n_leaves;
        // lanes per run x register policy x table placement x waves per CU
        // (all the occupancy allows, or half of it: fewer waves contend less
        // for the caches and LDS)
        std::vector<smc_launch> cands;
        for (int S = P < 64 ? P : 64; S >= 1; S >>= 1) {
            if (P / S > 16) break;
            for (int occ : {0, 3})
                for (int tab : {0, 1, 2}) {
                    cands.push_back(smc_launch{S, 0, 0, occ, tab});
                    int blocks = 0;
                    try {
                        Plan pl = plan_launch(P, m->img.n_vars, S, occ, tab, m->table_bytes, m->small_bytes, 0, 0, sample_runs,
                                              sample_runs, false, m->cu_count);
                        hip_check(smc::sim_occupancy(pl.C, false, pl.occ, pl.tab, pl.group_waves, pl.lds, &blocks),
                                  "occupancy query");
                        const int per_cu = blocks * pl.group_waves;
                        if (per_cu >= 2 * pl.group_waves) cands.push_back(smc_launch{S, 0, per_cu / 2, occ, tab});
                    } catch (const std::exception &) {
                    }
                }
        }
        // each candidate runs at least twice the runs its grid holds at once,
        // so the sample measures the steady state rather than the tail
        double best_cost = 1e300;
        smc_launch pick = cands.front();
        for (const smc_launch &c : cands) {
            uint64_t n = sample_runs;
            try {
                Plan pl = plan_launch(P, m->img.n_vars, c.lanes_per_run, c.occupancy, c.tables, m->table_bytes, m->small_bytes, 0,
                                      c.waves_per_cu, sample_runs, sample_runs, false, m->cu_count);
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
        tune_store(cache, key, pick);
        return SMC_OK;
    });
}

// ------------------------------------------------------------ batches

}  // extern "C"

struct smc_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<smc_batch_item> items;
    std::vector<RunBuffers> bufs;
    int C = 8, occ = 0, waves = 1, stride = 1;  // the one launch: slots per lane, register policy
    size_t lds = 0;
    InstDesc *d_all = nullptr;
    unsigned long long *d_counter = nullptr;
    ReducePartial *d_partials = nullptr, *d_acc = nullptr;
    uint64_t *d_round = nullptr;
    uint32_t blocks = 1;
    LaunchParams lp{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    ~smc_batch() {
        hipSetDevice(device);
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        for (auto &b : bufs) b.release();
        for (void *p : {(void *)d_all, (void *)d_counter, (void *)d_partials, (void *)d_acc, (void *)d_round})
            if (p) hipFree(p);
        for (hipEvent_t e : {ev0, ev1}) if (e) hipEventDestroy(e);
        if (stream) hipStreamDestroy(stream);
    }
};

namespace {

// The instance records of one batch launch with C slots per lane: lanes per
// run S_i = next_pow2(components_i) / C (at least 1), chunks of k R_i runs
// (R_i = 64 / S_i runs per wave) sized so the queue holds several chunks per
// wave; returns the LDS bytes one wave needs (the largest instance) and the
// total chunk count.  `runs` overrides every instance's n_runs (tuning
// samples).
struct BatchShape {
    std::vector<InstDesc> inst;
    size_t wave_lds = 0;
    uint64_t chunks = 0;
};

BatchShape batch_shape(smc_batch *b, int C, int waves, uint64_t runs_override) {
    BatchShape sh;
    uint64_t total = 0;
    for (auto &it : b->items) total += runs_override ? std::min<uint64_t>(runs_override, it.n_runs) : it.n_runs;
    const uint64_t per_wave = (total + (uint64_t)waves - 1) / (uint64_t)waves;
    for (size_t i = 0; i < b->items.size(); ++i) {
        const smc_batch_item &it = b->items[i];
        const int P = it.model->img.n_leaves;
        if (P > 64 * C) throw std::runtime_error("internal: instance wider than the launch's packing");
        const int S = P / C >= 1 ? P / C : 1;
        const int R = 64 / S;
        int levels = 0;
        while ((1 << levels) < S) ++levels;
        const size_t wave = ((size_t)R * b->stride * 8 + (size_t)2 * levels * 64 * 8 + 15) & ~(size_t)15;
        sh.wave_lds = std::max(sh.wave_lds, wave);
        const uint64_t n = runs_override ? std::min<uint64_t>(runs_override, it.n_runs) : it.n_runs;
        uint64_t k = (per_wave + 4ull * R - 1) / (4ull * R);
        k = std::max<uint64_t>(1, std::min<uint64_t>(k, 64));
        const uint64_t chunk = std::max<uint64_t>(1, std::min<uint64_t>(k * R, n));
        InstDesc d{};
        d.m = it.model->dm;
        d.p = make_params(&it.query, it.model->img, it.seed, it.run_begin, n, b->bufs[i], false, false);
        d.chunk_first = sh.chunks;
        d.lanes = S;
        d.chunk_runs = (uint32_t)chunk;
        sh.chunks += (n + chunk - 1) / chunk;
        sh.inst.push_back(d);
    }
    if (sh.wave_lds > 60000) throw std::runtime_error("limit: variables of the packed runs exceed the LDS budget");
    return sh;
}

// Waves of one launch at C slots per lane and register policy occ: every
// CU filled to its occupancy (or to waves_per_cu when that is lower).
int batch_waves(int C, int occ, size_t wave_lds, int waves_per_cu, int cu_count) {
    int blocks = 0;
    hip_check(smc::sim_occupancy(C, false, occ, false, 1, wave_lds, &blocks), "occupancy query");
    int per_cu = blocks;
    if (waves_per_cu > 0 && waves_per_cu < per_cu) per_cu = waves_per_cu;
    return std::max(1, per_cu) * cu_count;
}

}  // namespace

extern "C" {

int smc_batch_create(const smc_batch_item *items, int n_items, int tune, smc_batch **out) {
    if (!items || n_items <= 0 || !out) return fail(SMC_E_INVALID, "null argument or empty batch");
    *out = nullptr;
    return guarded([&]() -> int {
        auto b = std::make_unique<smc_batch>();
        b->items.assign(items, items + n_items);
        b->device = items[0].model ? items[0].model->device : 0;
        int maxP = 1, max_vars = 1;
        for (int i = 0; i < n_items; ++i) {
            const smc_batch_item &it = items[i];
            if (!it.model) throw std::runtime_error("batch item without a model");
            if (it.model->device != b->device) throw std::runtime_error("batch items on different devices");
            validate_query(&it.query, it.model->img.n_vars);
            if (it.n_runs == 0 || it.n_runs > (1ull << 32) - 1) throw std::runtime_error("batch item n_runs out of range");
            maxP = std::max(maxP, it.model->img.n_leaves);
            max_vars = std::max(max_vars, it.model->img.n_vars);
        }
        hip_check(hipSetDevice(b->device), "hipSetDevice");
        hip_check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
        for (hipEvent_t *e : {&b->ev0, &b->ev1}) hip_check(hipEventCreate(e), "hipEventCreate");
        const int cu = items[0].model->cu_count;
        b->stride = (max_vars + 1) | 1;  // clock time, then the variables
        b->bufs.resize(n_items);
        uint64_t maxr = 0;
        for (int i = 0; i < n_items; ++i) {
            b->bufs[i].ensure(items[i].n_runs);
            maxr = std::max<uint64_t>(maxr, items[i].n_runs);
        }
        b->blocks = std::min<uint32_t>(reduce_blocks(maxr), 64);
        hip_check(hipMalloc(&b->d_all, n_items * sizeof(InstDesc)), "hipMalloc instances");
        hip_check(hipMalloc(&b->d_counter, sizeof(unsigned long long)), "hipMalloc counter");
        hip_check(hipMalloc(&b->d_partials, (size_t)n_items * b->blocks * sizeof(ReducePartial)), "hipMalloc partials");
        hip_check(hipMalloc(&b->d_acc, n_items * sizeof(ReducePartial)), "hipMalloc totals");
        hip_check(hipMalloc(&b->d_round, sizeof(uint64_t)), "hipMalloc round");
        hip_check(hipMemset(b->d_round, 0, sizeof(uint64_t)), "clear round");

        // the launch's slots per lane: the smallest that holds the widest
        // instance in 64 lanes, or (tune) the fastest of that and wider
        // packings under both register policies on a sample of every
        // instance's runs
        int Cmin = 1;
        while (64 * Cmin < maxP) Cmin <<= 1;
        if (Cmin > 16) throw std::runtime_error("limit: more than 1024 components per network");
        auto set_launch = [&](const BatchShape &sh, int C, int occ, int waves) {
            b->C = C;
            b->occ = occ;
            b->waves = (int)std::min<uint64_t>((uint64_t)waves, std::max<uint64_t>(sh.chunks, 1));
            b->lds = sh.wave_lds;
            b->lp = LaunchParams{};
            b->lp.inst = b->d_all;
            b->lp.n_inst = n_items;
            b->lp.var_stride = b->stride;
            b->lp.total_chunks = sh.chunks;
            b->lp.counter = b->d_counter;
            b->lp.round = b->d_round;
            b->lp.tables = nullptr;
            b->lp.table_mode = 0;
            b->lp.table_bytes = 0;
            b->lp.wave_lds = (uint32_t)sh.wave_lds;
            hip_check(hipMemcpy(b->d_all, sh.inst.data(), n_items * sizeof(InstDesc), hipMemcpyHostToDevice),
                      "upload instances");
        };
        int bestC = std::max(Cmin, 8), best_occ = 0;
        if (tune) {
            double best_ms = 1e300;
            uint64_t sample = 0;
            for (auto &it : b->items) sample = std::max<uint64_t>(sample, std::min<uint64_t>(it.n_runs, 2048));
            for (int C = Cmin; C <= 16; C <<= 1)
                for (int occ : {0, 3}) {
                    BatchShape probe = batch_shape(b.get(), C, 1, sample);
                    const int waves = batch_waves(C, occ, probe.wave_lds, 0, cu);
                    BatchShape sh = batch_shape(b.get(), C, waves, sample);
                    set_launch(sh, C, occ, waves);
                    float ms = 0.f;
                    for (int rep = 0; rep < 2; ++rep) {  // the first launch warms the code
                        hip_check(smc::launch_prep(b->d_counter, nullptr, 0, b->stream), "prep");
                        hip_check(hipEventRecord(b->ev0, b->stream), "event");
                        hip_check(smc::launch_sim(b->lp, C, false, occ, b->waves, 1, b->lds, b->stream), "tune launch");
                        hip_check(hipEventRecord(b->ev1, b->stream), "event");
                        hip_check(hipStreamSynchronize(b->stream), "tune");
                        hip_check(hipEventElapsedTime(&ms, b->ev0, b->ev1), "event time");
                    }
                    if (ms < best_ms) { best_ms = ms; bestC = C; best_occ = occ; }
                }
        }
        {
            BatchShape probe = batch_shape(b.get(), bestC, 1, 0);
            const int waves = batch_waves(bestC, best_occ, probe.wave_lds, 0, cu);
            BatchShape sh = batch_shape(b.get(), bestC, waves, 0);
            set_launch(sh, bestC, best_occ, waves);
        }
        hip_check(hipDeviceSynchronize(), "batch setup");

        // capture one round: the run kernel over every instance, the
        // reduction accumulates, then the counter is cleared and the round
        // advances
        hip_check(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "begin capture");
        hip_check(smc::launch_prep(b->d_counter, nullptr, 0, b->stream), "capture prep");
        hip_check(smc::launch_sim(b->lp, b->C, false, b->occ, b->waves, 1, b->lds, b->stream), "capture run kernel");
        hip_check(smc::launch_reduce(b->d_all, n_items, b->blocks, b->d_partials, b->d_acc, 1, b->stream), "capture reduce");
        hip_check(smc::launch_prep(b->d_counter, b->d_round, 1, b->stream), "capture advance");
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
            stats->groups = 1;
            stats->rounds = rounds;
        }
        return SMC_OK;
    });
}

void smc_batch_destroy(smc_batch *b) { delete b; }

}  // extern "C"
