/*
 * smc_capi.h — C ABI of the MI355X statistical model checker.
 *
 * Conventions: C linkage, no exceptions cross this boundary, every function
 * that can fail returns 0 on success and a negative smc_status otherwise, and
 * smc_last_error() returns a thread-local message for the last failure.
 * Handles are not thread-safe; distinct handles may be used concurrently.
 * The caller owns every host array it passes; the library copies what it
 * keeps.  Results are on the host when a call returns.
 */
#ifndef SMC_CAPI_H
#define SMC_CAPI_H

#include <stddef.h>
#include <stdint.h>
#include "smc_network.h"

#ifdef __cplusplus
extern "C" {
#endif

enum smc_status {
    SMC_OK = 0,
    SMC_E_INVALID = -1,     /* malformed network, query or argument        */
    SMC_E_PARSE = -2,       /* model file could not be read or understood  */
    SMC_E_DEVICE = -3,      /* HIP runtime error or no usable GPU          */
    SMC_E_LIMIT = -4,       /* model exceeds a documented size limit       */
    SMC_E_ALLOC = -5        /* allocation failed                           */
};

/* ------------------------------------------------------------ front end */

/* An owned, parsed network plus the names the command line refers to. */
typedef struct smc_parsed smc_parsed;

/*
 * Parse a UPPAAL XML model (the reference's uppaal_xml_parser::parse,
 * automata_parser/uppaal_xml_parser.cpp:623-706, plus the goal marking of
 * smc.cu:405-436 and domain_optimization_visitor.cpp:48-49,151-167).
 * loc_query is the `-q` location query "Process.location" or NULL/"" for a
 * variable query; locations it names, and locations named "Goal", become
 * goal locations.
 */
int smc_parse_uppaal(const char *path, const char *loc_query, smc_parsed **out);
/* Same, from an in-memory document. */
int smc_parse_uppaal_string(const char *xml, const char *loc_query, smc_parsed **out);
/* Borrowed view of the parsed network, valid until smc_parsed_destroy. */
const smc_network *smc_parsed_network(const smc_parsed *p);
/* Names: component (template) c, location l, variable v; "" if none. */
const char *smc_parsed_component_name(const smc_parsed *p, int c);
const char *smc_parsed_location_name(const smc_parsed *p, int l);
const char *smc_parsed_variable_name(const smc_parsed *p, int v);
/* Number of goal locations the query marked (0 means the query matched none). */
int smc_parsed_goal_count(const smc_parsed *p);
void smc_parsed_destroy(smc_parsed *p);

/* ------------------------------------------------------------ models */

typedef struct smc_model smc_model;

/*
 * Build a simulator for a network on GPU `device`: validate it, compile its
 * expressions to the run kernel's register bytecode, and upload the tables.
 * Replaces init_shared_model_state (simulation/state/shared_model_state.cuh:
 * 219-225, shared_model_state.cu:164-689), which allocated one device buffer
 * per expression and never freed them; this model owns one arena, freed by
 * smc_model_destroy.
 */
int smc_model_create(const smc_network *net, int device, smc_model **out);
void smc_model_destroy(smc_model *m);

/* Launch shape.  Zero fields mean "choose": lanes_per_run = next power of two
 * of the component count capped at 64; waves = enough to fill the GPU. */
typedef struct smc_launch {
    int32_t lanes_per_run;     /* S: power of two <= 64                    */
    int32_t waves;             /* persistent waves in the grid              */
    int32_t waves_per_cu;      /* occupancy target used when waves == 0     */
    int32_t occupancy;         /* 0: registers unrestricted (two waves per
                                  SIMD); 3: at most 168 registers, three
                                  waves per SIMD                           */
    int32_t tables;            /* 0: model tables read from global memory;
                                  1: staged in LDS, one copy per workgroup
                                  of four waves; 2: staged in LDS but for the
                                  location records, read from global memory
                                  (smc_run only; batches read the tables
                                  from global memory)                      */
} smc_launch;

/* Per-run outputs, all optional (NULL = not wanted), each n_runs long. */
typedef struct smc_run_outputs {
    uint8_t *flag;         /* reach: goal hit; var probability: beyond threshold */
    double *value;         /* var query: the per-run max (or min)                */
    uint32_t *steps;       /* race rounds taken                                  */
    uint32_t *transitions; /* edges fired, by race winners and receivers         */
    uint64_t *digest;      /* hash of the whole trajectory (parity checks)       */
    double *time;          /* global time when the run ended                     */
} smc_run_outputs;

typedef struct smc_result {
    uint64_t n_runs;
    uint64_t hits;          /* reach, var probability: runs counted             */
    double estimate;        /* hits/n, or the mean per-run value (estimate mode) */
    double sum, sumsq;      /* of the per-run values (var query)                */
    double vmin, vmax;      /* of the per-run values (var query)                */
    double ch_eps;          /* Chernoff-Hoeffding half-width, delta = 0.05       */
    double clt_half;        /* 95% normal half-width of the estimate             */
    uint64_t steps;         /* total race rounds                                 */
    uint64_t transitions;   /* total edges fired                                 */
    uint64_t truncated;     /* runs cut off at max_steps before the time bound   */
    double kernel_ms;       /* run kernel time, HIP events on the model stream   */
    double reduce_ms;       /* reduction kernel time                             */
    double wall_ms;         /* whole call                                        */
    int32_t lanes_per_run;  /* launch actually used                              */
    int32_t waves;
} smc_result;

/*
 * Simulate runs [run_begin, run_begin + n_runs) of `q` with stream `seed`
 * and reduce them on the GPU.  Counterpart of run_statistical_model_checking
 * (smc.cu:11-281) with the kernel launch (simulation.cu:674-901) and the
 * result reduction (simulation/statistics.cuh:126-163, smc.cu:378-401).
 * Runs are independent and addressed by their global index, so any split of
 * a run range across calls or devices gives the same per-run results.
 */
int smc_run(smc_model *m, const smc_query *q, uint64_t seed,
            uint64_t run_begin, uint64_t n_runs, const smc_launch *launch,
            smc_run_outputs *outputs, smc_result *out);

/*
 * smc_run without device allocation.  No counterpart in the reference: for
 * callers that cannot allocate (code under stream capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The same runs, kernels and
 * totals as smc_run with outputs == NULL, but every intermediate buffer (the
 * per-run results the reduction reads, the reduction's partials, the work
 * counter, the instance record) lives in a fixed scratch array of about
 * 17 MiB per GPU, held until the process ends.  n_runs is at most 2^20
 * (SMC_E_LIMIT beyond).  Every call on one GPU shares that GPU's scratch:
 * calls that overlap on one GPU (from several threads or models) race.
 * Not a graph-capture entry point: it synchronises the model's stream.
 */
int smc_run_noalloc(smc_model *m, const smc_query *q, uint64_t seed,
                    uint64_t run_begin, uint64_t n_runs, const smc_launch *launch,
                    smc_result *out);

/*
 * Autotuner: time a sample of `sample_runs` runs of `q` with each candidate
 * launch shape (lanes per run from next_pow2(components) down to a
 * sixteenth of it, i.e. 1-16 components per lane; register policy; tables
 * in global memory or in LDS when they fit; waves per CU at the occupancy
 * limit or half of it) and keep the fastest in the model; smc_run uses it
 * whenever its launch argument is NULL or all zero.  The chosen shape is
 * written to *best.  Choices persist in a cache file keyed by the model,
 * the query, the device and the library build: SMC_TUNE_CACHE names it
 * (default $XDG_CACHE_HOME/cusmc_amd/tune.txt, else
 * ~/.cache/cusmc_amd/tune.txt; "off" disables it); a cached key is not
 * timed again.  Results of smc_run do not depend on the shape
 * (bit-identical per run).
 */
int smc_autotune(smc_model *m, const smc_query *q, uint64_t seed, uint64_t sample_runs, smc_launch *best);

/* ------------------------------------------------------------ batches */

/*
 * Many (model, query, run range) instances with uneven component counts in
 * one persistent launch: one run kernel over a queue of (instance, chunk of
 * runs); the launch holds C component slots per lane and each instance
 * packs a run into S_i = next_pow2(components_i) / C lanes (64 / S_i runs
 * per wave), so narrow and wide networks share the grid; one reduction
 * folds every instance's runs.  The whole round is captured once in a HIP
 * graph; each replay simulates the next n_runs runs of every instance
 * (round r covers run indices run_begin + r*n_runs ...), and totals
 * accumulate across rounds.
 * No counterpart in the reference (one model per process, smc.cu:283).
 */
typedef struct smc_batch smc_batch;

typedef struct smc_batch_item {
    smc_model *model;       /* all on one device; must outlive the batch     */
    smc_query query;
    uint64_t seed;
    uint64_t run_begin;
    uint64_t n_runs;        /* per round                                      */
} smc_batch_item;

typedef struct smc_batch_stats {
    double graph_ms;        /* HIP-event time of all replays of the last run  */
    uint64_t runs;          /* runs simulated by the last smc_batch_run       */
    uint64_t steps, transitions;
    int32_t groups;         /* run-kernel launches per round (1)              */
    int32_t rounds;
} smc_batch_stats;

/* tune != 0: time one round on a sample of every instance's runs for each
 * slots-per-lane value (from the smallest that fits the widest instance up
 * to 16) and register policy, and keep the fastest; else C = 8, or the
 * smallest that fits. */
int smc_batch_create(const smc_batch_item *items, int n_items, int tune, smc_batch **out);
/* Replay `rounds` rounds; results[i] = totals of item i over these rounds. */
int smc_batch_run(smc_batch *b, int rounds, smc_result *results, smc_batch_stats *stats);
void smc_batch_destroy(smc_batch *b);

/* ------------------------------------------------------------ importance sampling */

/*
 * Learned importance sampling for a rare-event variable-probability query
 * Pr[<=T](<> v > threshold) (max queries).  The proposal multiplies the
 * rate of every exponential component by 2^theta(location, state), theta
 * from an MLP (features: one-hot locations, scaled variables, time; two
 * hidden layers, BF16 weights on MFMA with FP32 accumulation), trained by
 * multilevel cross-entropy on trajectories the simulator generates; every
 * run carries an FP64 log2 likelihood ratio, so the estimator
 * mean(hit * 2^log2lr) is unbiased for any proposal.  No counterpart in the
 * reference, which only estimates by plain Monte Carlo.
 */
typedef struct smc_is smc_is;

typedef struct smc_is_config {
    int32_t hidden;          /* hidden width (default 128)                      */
    int32_t rounds;          /* cross-entropy rounds (default 5)                */
    uint64_t runs_per_round; /* trajectories per round (default 100000)         */
    double elite_frac;       /* fraction kept per round (default 0.1)           */
    int32_t epochs;          /* gradient steps per round (default 20)           */
    float learning_rate;     /* Adam step (default 1e-3)                        */
    float tilt_max;          /* bound on |theta| (log2 units, default 4)        */
    int32_t per_location;    /* 1: a tilt per location; 0 (default): one per
                                template location, shared by the structurally
                                identical components of the network          */
    uint64_t seed;           /* weight initialisation and sampling streams      */
} smc_is_config;

typedef struct smc_is_round {
    double level;            /* elite level of this round                       */
    uint64_t elites;
    uint64_t hits;           /* runs reaching the query threshold               */
    double estimate;         /* IS estimate from this round's runs              */
    double ms;               /* sampling + training time                        */
    double ess;              /* effective sample size of the elite weights      */
    uint64_t pairs;          /* recorded (run, macro-step) pairs trained on     */
} smc_is_round;

typedef struct smc_is_result {
    uint64_t n_runs, hits;
    double estimate;         /* mean(hit * 2^log2lr)                            */
    double rel_half_width;   /* 95% normal half-width / estimate                */
    double ms;               /* wall time of the estimation                     */
    double ess;              /* effective sample size of the hit weights        */
    uint64_t steps, transitions;
} smc_is_result;

typedef struct smc_is_outputs {  /* per-run, each n_runs long, all optional     */
    uint8_t *flag;
    double *log2lr;
    uint64_t *digest;
    uint32_t *steps;
    uint32_t *transitions;
    double *time;
} smc_is_outputs;

int smc_is_create(smc_model *m, const smc_query *q, const smc_is_config *cfg, smc_is **out);
/* Run the cross-entropy rounds; rounds[] (cfg.rounds entries) may be NULL. */
int smc_is_train(smc_is *s, smc_is_round *rounds);
/* Sample batches of batch_runs until rel_half_width <= rel_target or
 * max_runs runs. */
int smc_is_estimate(smc_is *s, uint64_t seed, uint64_t run_begin, uint64_t batch_runs, uint64_t max_runs,
                    double rel_target, smc_is_result *out);
/* Sample n_runs runs with the trained proposal (theta == NULL and
 * use_mlp != 0), with no tilt (theta == NULL, use_mlp == 0), or with the
 * fixed tilt theta[location] (log2 units).  lockstep != 0 advances by
 * macro-steps; early_stop != 0 ends a run once its outcome is settled. */
int smc_is_sample(smc_is *s, const float *theta, int use_mlp, uint64_t seed, uint64_t run_begin, uint64_t n_runs,
                  int lockstep, int early_stop, smc_is_outputs *out, smc_is_result *res);
/* Parameter count, and flat FP32 parameters [W1 b1 W2 b2 W3 b3] (W row-major
 * in x out, rows of W1 padded to features_padded, columns of W3 to
 * outputs_padded: one output per tilt group). */
int smc_is_shape(const smc_is *s, int32_t *features_padded, int32_t *hidden, int32_t *outputs_padded,
                 int32_t *locations_padded, uint64_t *n_params);
/* Tilt group of every location (the MLP output it reads); returns the
 * number of groups, or -1. */
int smc_is_groups(const smc_is *s, int32_t *group_of_location);
int smc_is_get_params(const smc_is *s, float *out);
int smc_is_set_params(smc_is *s, const float *in);
/* MLP on `rows` feature rows x (rows x features_padded, FP32, rounded to
 * BF16): theta per location (rows x locations_padded, each location reading
 * its group's output); with g != NULL also the gradient of sum(g * theta)
 * over the parameters (g is rows x locations_padded; entries at clamped or
 * untiltable outputs are ignored). */
int smc_is_mlp_eval(smc_is *s, const float *x, int rows, const float *g, float *theta, float *grad_out);
/* Replay runs[0..n_runs) (seed) under the fixed tilt theta[location], or the
 * MLP when theta == NULL, recording every macro-step (a macro-step ends with
 * the first race whose winner fires an edge) with row weights w (NULL: 1).
 * Pairs are numbered macro-step-major, run-minor; up to cap are copied out:
 * X [pairs][features_padded] bf16 bits of the features at the macro-step's
 * start, A [pairs][locations_padded] exponential wins per location, B the
 * same shape, ln2 * duration * base rate.  *pairs = number recorded. */
int smc_is_record(smc_is *s, const float *theta, uint64_t seed, const uint64_t *runs, int n_runs, const float *w,
                  int early_stop, uint16_t *X, float *A, float *B, uint64_t cap, uint64_t *pairs);
/* Gradient of the cross-entropy loss -sum_pairs w ln q(theta = MLP(X)) over
 * the recorded pairs, wrt the flat parameters: what one epoch descends. */
int smc_is_ce_gradient(smc_is *s, float *grad_out);
void smc_is_destroy(smc_is *s);

/* Interval helpers (host). */
double smc_ch_epsilon(uint64_t n, double delta);

const char *smc_last_error(void);
const char *smc_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SMC_CAPI_H */
