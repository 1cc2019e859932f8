// capi.cpp — the C ABI (include/smc_capi.h).
#include "smc_capi.h"

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "compile.hpp"
#include "kernels.hpp"
#include "uppaal.hpp"

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

template <class F>
int guarded(F &&f) {
    try {
        return f();
    } catch (const HipError &e) {
        return fail(SMC_E_DEVICE, e.what());
    } catch (const std::bad_alloc &) {
        return fail(SMC_E_ALLOC, "out of host memory");
    } catch (const std::exception &e) {
        std::string m = e.what();
        if (m.rfind("limit:", 0) == 0) return fail(SMC_E_LIMIT, m);
        return fail(SMC_E_INVALID, m);
    }
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace

struct smc_parsed {
    std::unique_ptr<smc::Parsed> p;
    smc_network view{};
};

struct smc_model {
    int device = 0;
    hipStream_t stream = nullptr;
    smc::ModelImage img;
    DevModel dm{};
    void *arena = nullptr;
    int cu_count = 0;
    // per-run device buffers, grown on demand
    size_t cap = 0;
    uint8_t *d_flag = nullptr;
    double *d_value = nullptr;
    uint32_t *d_steps = nullptr, *d_trans = nullptr;
    uint64_t *d_digest = nullptr;
    double *d_time = nullptr;
    unsigned long long *d_counter = nullptr;
    ReducePartial *d_partials = nullptr;
    int n_partials = 0;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};

    ~smc_model() {
        if (device >= 0) hipSetDevice(device);
        free_runs();
        if (d_counter) hipFree(d_counter);
        if (d_partials) hipFree(d_partials);
        if (arena) hipFree(arena);
        for (auto &e : ev) if (e) hipEventDestroy(e);
        if (stream) hipStreamDestroy(stream);
    }
    void free_runs() {
        for (void *p : {(void *)d_flag, (void *)d_value, (void *)d_steps, (void *)d_trans, (void *)d_digest, (void *)d_time})
            if (p) hipFree(p);
        d_flag = nullptr; d_value = nullptr; d_steps = nullptr; d_trans = nullptr; d_digest = nullptr; d_time = nullptr;
        cap = 0;
    }
    void ensure_runs(size_t n) {
        if (n <= cap) return;
        free_runs();
        size_t c = std::max<size_t>(n, 1024);
        hip_check(hipMalloc(&d_flag, c), "hipMalloc flags");
        hip_check(hipMalloc(&d_value, c * 8), "hipMalloc values");
        hip_check(hipMalloc(&d_steps, c * 4), "hipMalloc steps");
        hip_check(hipMalloc(&d_trans, c * 4), "hipMalloc transitions");
        hip_check(hipMalloc(&d_digest, c * 8), "hipMalloc digests");
        hip_check(hipMalloc(&d_time, c * 8), "hipMalloc times");
        cap = c;
    }
};

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
        };
        size_t total = 0;
        for (auto &p : parts) { p.off = total; total = align_up(total + p.bytes, 256); }
        std::vector<uint8_t> host(total, 0);
        for (auto &p : parts) if (p.bytes) memcpy(host.data() + p.off, p.src, p.bytes);
        hip_check(hipMalloc(&m->arena, total), "hipMalloc model arena");
        hip_check(hipMemcpy(m->arena, host.data(), total, hipMemcpyHostToDevice), "upload model");
        auto at = [&](int i) { return (const void *)((const uint8_t *)m->arena + parts[i].off); };
        DevModel &d = m->dm;
        d.locs = (const DevLoc *)at(0);
        d.invs = (const DevInv *)at(1);
        d.edges = (const DevEdge *)at(2);
        d.guards = (const DevGuard *)at(3);
        d.upds = (const DevUpd *)at(4);
        d.consts = (const double *)at(5);
        d.code = (const DevInsn *)at(6);
        d.var_init = (const double *)at(7);
        d.comp_init_loc = (const int32_t *)at(8);
        d.clocks = (const uint16_t *)at(9);
        d.n_comp = im.n_comp;
        d.n_vars = im.n_vars;
        int nclk = 0;
        for (int k : im.var_kind) nclk += k == SMC_VAR_CLOCK;
        d.n_clocks = nclk;
        d.n_leaves = im.n_leaves;
        d.n_locs = (int)im.locs.size();
        d.n_edges = net->n_edges;
        hip_check(hipMalloc(&m->d_counter, sizeof(unsigned long long)), "hipMalloc counter");
        m->n_partials = 1024;
        hip_check(hipMalloc(&m->d_partials, m->n_partials * sizeof(ReducePartial)), "hipMalloc partials");
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
        if (q->kind != SMC_QUERY_REACH && q->kind != SMC_QUERY_VAR) throw std::runtime_error("unknown query kind");
        if (q->kind == SMC_QUERY_VAR && (q->var < 0 || q->var >= m->img.n_vars))
            throw std::runtime_error("query variable id out of range");
        if (!(q->time_bound >= 0.0)) throw std::runtime_error("time bound must be a non-negative number");
        if (n_runs == 0) throw std::runtime_error("n_runs must be positive");
        if (n_runs > (1ull << 32)) throw std::runtime_error("limit: at most 2^32 runs per call");
        if (run_begin > (~0ull) - n_runs) throw std::runtime_error("run range overflows");

        const int P = m->img.n_leaves;
        int S = launch && launch->lanes_per_run ? launch->lanes_per_run : (P < 64 ? P : 64);
        if (S <= 0 || S > 64 || (S & (S - 1))) throw std::runtime_error("lanes_per_run must be a power of two <= 64");
        if (S > P) S = P;
        int C = P / S;
        if (C > 16) throw std::runtime_error("limit: more than 16 components per lane; use more lanes per run");
        const int R = 64 / S;
        int stride = m->img.n_vars > 0 ? m->img.n_vars : 1;
        stride |= 1;
        size_t lds = (size_t)R * stride * sizeof(double);
        if (lds > 65536) throw std::runtime_error("limit: variables of the packed runs exceed 64 KiB of LDS");
        const bool trace = outputs && outputs->digest;

        hip_check(hipSetDevice(m->device), "hipSetDevice");
        m->ensure_runs(n_runs);
        int waves = launch && launch->waves > 0 ? launch->waves : 0;
        if (!waves) {
            int per_cu = 0;
            hip_check(smc::sim_occupancy(C, trace, lds, &per_cu), "occupancy query");
            if (launch && launch->waves_per_cu > 0) per_cu = launch->waves_per_cu;
            if (per_cu <= 0) per_cu = 1;
            waves = m->cu_count * per_cu;
        }
        uint64_t need = (n_runs + R - 1) / R;
        if ((uint64_t)waves > need) waves = (int)need;

        SimParams p{};
        p.seed = seed;
        p.run_begin = run_begin;
        p.n_runs = n_runs;
        p.bound = q->time_bound;
        p.threshold = q->threshold;
        p.reach = q->kind == SMC_QUERY_REACH;
        p.qvar = q->kind == SMC_QUERY_VAR ? q->var : -1;
        p.qinit = (p.qvar >= 0 && m->img.var_kind[p.qvar] == SMC_VAR_INT) ? m->img.var_init[p.qvar] : 0.0;
        p.is_max = q->is_max ? 1 : 0;
        p.lanes = S;
        p.var_stride = stride;
        p.max_steps = q->max_steps ? q->max_steps : SMC_DEFAULT_MAX_STEPS;
        p.flag = m->d_flag;
        p.value = m->d_value;
        p.steps = m->d_steps;
        p.trans = m->d_trans;
        p.digest = trace ? m->d_digest : nullptr;
        p.time = (outputs && outputs->time) ? m->d_time : nullptr;
        p.counter = m->d_counter;

        hip_check(hipMemsetAsync(m->d_counter, 0, sizeof(unsigned long long), m->stream), "clear counter");
        hip_check(hipEventRecord(m->ev[0], m->stream), "event");
        hip_check(smc::launch_sim(m->dm, p, C, trace, waves, lds, m->stream), "run kernel launch");
        hip_check(hipEventRecord(m->ev[1], m->stream), "event");
        int blocks = (int)std::min<uint64_t>((uint64_t)m->n_partials, (n_runs + 4095) / 4096);
        if (blocks < 1) blocks = 1;
        hip_check(smc::launch_reduce(m->d_flag, m->d_value, m->d_steps, m->d_trans, n_runs, p.max_steps, blocks,
                                     m->d_partials, m->stream), "reduce kernel launch");
        hip_check(hipEventRecord(m->ev[2], m->stream), "event");
        std::vector<ReducePartial> parts(blocks);
        hip_check(hipMemcpyAsync(parts.data(), m->d_partials, blocks * sizeof(ReducePartial),
                                 hipMemcpyDeviceToHost, m->stream), "copy partials");
        if (outputs) {
            if (outputs->flag) hip_check(hipMemcpyAsync(outputs->flag, m->d_flag, n_runs, hipMemcpyDeviceToHost, m->stream), "copy flags");
            if (outputs->value) hip_check(hipMemcpyAsync(outputs->value, m->d_value, n_runs * 8, hipMemcpyDeviceToHost, m->stream), "copy values");
            if (outputs->steps) hip_check(hipMemcpyAsync(outputs->steps, m->d_steps, n_runs * 4, hipMemcpyDeviceToHost, m->stream), "copy steps");
            if (outputs->transitions) hip_check(hipMemcpyAsync(outputs->transitions, m->d_trans, n_runs * 4, hipMemcpyDeviceToHost, m->stream), "copy transitions");
            if (outputs->digest) hip_check(hipMemcpyAsync(outputs->digest, m->d_digest, n_runs * 8, hipMemcpyDeviceToHost, m->stream), "copy digests");
            if (outputs->time) hip_check(hipMemcpyAsync(outputs->time, m->d_time, n_runs * 8, hipMemcpyDeviceToHost, m->stream), "copy times");
        }
        hip_check(hipStreamSynchronize(m->stream), "run");
        float k_ms = 0.f, r_ms = 0.f;
        hip_check(hipEventElapsedTime(&k_ms, m->ev[0], m->ev[1]), "event time");
        hip_check(hipEventElapsedTime(&r_ms, m->ev[1], m->ev[2]), "event time");

        ReducePartial tot{};
        tot.vmin = HUGE_VAL;
        tot.vmax = -HUGE_VAL;
        for (const auto &r : parts) {
            tot.sum += r.sum;
            tot.sumsq += r.sumsq;
            tot.vmin = std::fmin(tot.vmin, r.vmin);
            tot.vmax = std::fmax(tot.vmax, r.vmax);
            tot.hits += r.hits;
            tot.steps += r.steps;
            tot.trans += r.trans;
            tot.truncated += r.truncated;
        }
        const double N = (double)n_runs;
        out->n_runs = n_runs;
        out->hits = tot.hits;
        out->sum = tot.sum;
        out->sumsq = tot.sumsq;
        out->vmin = tot.vmin;
        out->vmax = tot.vmax;
        out->steps = tot.steps;
        out->transitions = tot.trans;
        out->truncated = tot.truncated;
        out->ch_eps = smc_ch_epsilon(n_runs, 0.05);
        if (q->kind == SMC_QUERY_VAR && q->is_estimate) {
            out->estimate = tot.sum / N;
            double var = n_runs > 1 ? std::max(0.0, (tot.sumsq - tot.sum * tot.sum / N) / (N - 1.0)) : 0.0;
            out->clt_half = 1.959963984540054 * std::sqrt(var / N);
        } else {
            double pr = (double)tot.hits / N;
            out->estimate = pr;
            out->clt_half = 1.959963984540054 * std::sqrt(pr * (1.0 - pr) / N);
        }
        out->kernel_ms = k_ms;
        out->reduce_ms = r_ms;
        out->lanes_per_run = S;
        out->waves = waves;
        out->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
        return SMC_OK;
    });
}

}  // extern "C"
