/*
 * oracle.h — CPU restatement of cuSMC's simulation semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker of the GPU run
 * kernel: only the tests, the smoke check and the benchmark's CPU-baseline
 * leg may load it.  The product (libcusmc) never calls it.
 *
 * It walks the expression trees of an smc_network directly (no bytecode) and
 * runs each simulation sequentially, following simulation/simulation.cu and
 * simulation/expressions.cu of cuSMC as summarised in oracle.cpp.
 */
#ifndef SMC_ORACLE_H
#define SMC_ORACLE_H

#include <stdint.h>
#include "smc_network.h"

#ifdef __cplusplus
extern "C" {
#endif

enum oracle_mode {
    ORACLE_CANONICAL = 0, /* the GPU's race sampling: same draws, same bits      */
    ORACLE_LITERAL = 1    /* cuSMC's: every component draws its own delay each step */
};

typedef struct oracle_outputs {
    uint8_t *flag;
    double *value;
    uint32_t *steps;
    uint32_t *transitions;
    uint64_t *digest;
    double *time;
} oracle_outputs;

typedef struct oracle_result {
    uint64_t n_runs;
    uint64_t hits;
    double estimate;
    double sum, sumsq;
    uint64_t steps, transitions;
    double seconds;     /* wall time of the simulation loop only */
    int32_t threads;
    int32_t reserved;
} oracle_result;

/* Simulate runs [run_begin, run_begin + n_runs) on `threads` host threads
 * (0 = all cores).  Returns 0 or -1 (see oracle_last_error). */
int oracle_run(const smc_network *net, const smc_query *q, uint64_t seed, uint64_t run_begin,
               uint64_t n_runs, int mode, int threads, oracle_outputs *outputs, oracle_result *res);

/* One run with its event log: events[k] = {step, component, edge, 0} for
 * every fired edge, times[k] = global time before the step.  Returns the
 * number of events (may exceed cap; only cap are written) or -1. */
long oracle_trace(const smc_network *net, const smc_query *q, uint64_t seed, uint64_t run,
                  int mode, int32_t *events, double *times, long cap);

/* Importance sampling (canonical race only): exponential rates multiplied
 * by 2^theta[location] (theta may be NULL: no tilt), per-run log2 likelihood
 * ratio of the base law against the tilted one; early_stop ends a
 * var-probability run once its max (min) has crossed the threshold. */
typedef struct oracle_is_outputs {
    uint8_t *flag;
    double *log2lr;
    uint64_t *digest;
    uint32_t *steps;
    uint32_t *transitions;
    double *time;
} oracle_is_outputs;

int oracle_is_run(const smc_network *net, const smc_query *q, uint64_t seed, uint64_t run_begin, uint64_t n_runs,
                  const float *theta, int early_stop, int threads, oracle_is_outputs *out);

/* Macro-steps of one tilted run (a macro-step ends with the first race whose
 * winner fires an edge): per macro-step k the bf16 features at its start
 * X[k][fpad] (one-hot locations, vars * var_scale, t / bound, zero padding)
 * and per location the exponential wins A[k][lpad] and ln2 * dt * base rate
 * B[k][lpad].  Returns the number of macro-steps (only cap written) or -1. */
long oracle_is_record(const smc_network *net, const smc_query *q, uint64_t seed, uint64_t run, const float *theta,
                      int early_stop, const float *var_scale, int fpad, int lpad, uint16_t *X, float *A, float *B,
                      long cap);

/* Algorithmic work of runs [run_begin, run_begin + n_runs) (canonical
 * race): lane-operations by the cost model in oracle.cpp (one per FP64
 * add/mul/compare; division, Philox, log2 and the other published methods at
 * their operation counts), with the race steps and transitions. */
int oracle_cost(const smc_network *net, const smc_query *q, uint64_t seed, uint64_t run_begin, uint64_t n_runs,
                int threads, uint64_t *steps, uint64_t *transitions, double *lane_ops);

/* Pairwise-tree sum and proportional descent used by the canonical race,
 * exposed for tests: returns the tree sum of rates[0..n) padded to the next
 * power of two, and writes the leaf chosen for target v * sum. */
double oracle_tree_pick(const double *rates, int n, double v, int *chosen);

/* Stream primitives and expression evaluation, exposed for tests. */
void oracle_philox(const uint32_t *ctr, const uint32_t *key, uint32_t *out);
double oracle_u01(uint32_t hi, uint32_t lo);
/* Component c's uniform from its shared draw w[4] (receiver != 0: a receiver
 * choice, else a uniform delay), and the index of that draw. */
double oracle_shared_u(const uint32_t *w, uint32_t c, int receiver);
uint32_t oracle_shared_draw(uint32_t c, int receiver);
double oracle_log2(double x);
double oracle_pow(double a, double b);
double oracle_eval_expr(const smc_network *net, int e, const double *vars);

const char *oracle_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
