// is_kernels.hpp — launch interface of the importance-sampling kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "devmodel.h"

enum { IS_ACTIVE = 1, IS_HIT = 2, IS_GOAL = 4, IS_MID = 8 };

// State of the runs of a lock-step batch, one row per run (device arrays).
struct IsRows {
    uint16_t *loc;        // [rows][n_comp]
    double *vars;         // [rows][n_vars]
    double *t, *log2lr, *qmax, *qmin;
    uint32_t *step, *trans;
    int32_t *chan, *sender;
    uint8_t *status;      // IS_* bits
    uint64_t *digest;
    uint64_t *run;        // global run index (Philox counter)
    uint32_t *job;        // index of the run in the batch (outputs, weights)
};

// Results of a batch's runs, indexed by job.
struct IsOut {
    uint8_t *status;
    double *log2lr, *score, *t;   // score: the running max (min) of the query variable
    uint32_t *step, *trans;
    uint64_t *digest;
};

struct IsParams {
    DevModel m;
    IsRows st;
    int32_t n_rows;
    int32_t lanes;
    int32_t var_stride;
    int32_t lpad;              // row length of theta and of the recorded statistics (>= n_locs)
    int32_t fpad;              // row length of features (>= n_locs + n_vars + 1)
    // query
    uint64_t seed;
    double bound, threshold, qinit;
    int32_t reach, qvar, is_max;
    uint32_t max_steps;
    int32_t stop_at_transition; // 1: lock-step macro-steps, 0: run to the end
    int32_t early_stop;         // end a var-probability run once it is decided
    int32_t trace;              // keep the digest of every fired edge
    // proposal
    const float *theta;         // log2 tilts per location, or null (no tilt)
    int32_t theta_stride;       // row stride of theta (0: one row for every run)
    const uint8_t *tiltable;    // [lpad]
    const int32_t *tilt_index;  // [lpad]: the location's column in a theta row, -1 if untiltable
    const SMC_GLOBAL uint8_t *tables;  // hot tables staged in LDS when table_bytes > 0: arena start
    uint32_t table_bytes;       // bytes staged (multiple of 16)
    uint32_t wave_lds;          // LDS bytes of one wave after the shared part (multiple of 16)
    int32_t theta_len;          // entries of a theta row (<= lpad when theta_stride == 0, else <= theta_cap)
    int32_t theta_cap;          // per-row tilt entries each wave's LDS holds room for
    uint16_t *feat;             // [rows][fpad] bf16 features of the end state, or null
    const float *var_scale;     // [n_vars]
    // recording of cross-entropy statistics (replay of elite runs)
    const int32_t *rec_pos;     // [rows] slot of this macro-step's record, -1: none; null: no recording
    uint16_t *rec_X;            // [slots][fpad] features at the macro-step's start
    float *rec_A;               // [slots][lpad] exponential wins per location
    float *rec_B;               // [slots][lpad] ln2 * dt * base rate per location
    float *rec_w;               // [slots] weight of the row
    const float *weight;        // [jobs]
    unsigned int *active_count; // rows still active after the launch
};

// The proposal network's forward pass in one launch: theta_g =
// clamp(W3^T relu(W2^T relu(W1^T x + b1) + b2) + b3), zero at untiltable
// groups.  Same products, same fp32 accumulation order and epilogues as
// three chained gemm() calls, so the outputs are bit-identical to them.
struct MlpFwdArgs {
    const uint16_t *X;                 // [rows][fpad] bf16 features
    const void *image;                 // the weights as staged in LDS (mlp_pack)
    const uint8_t *col_mask;           // [gpad]: group tiltable
    uint16_t *a1, *a2;                 // optional [rows][H] bf16 activations (for the backward pass)
    float *og;                         // [rows][ldo]: outputs g < g16, zeros g16 <= g < ldo
    int rows, fpad, H, gpad, g16, ldo; // rows % 16 == 0; g16 = groups rounded up to 16 (<= 64)
    float clamp_max;                   // > 0
};

struct GemmArgs {
    const uint16_t *A;  int lda;     // bf16
    const uint16_t *B;  int ldb;     // bf16
    float *C;           int ldc;     // fp32 out (optional)
    uint16_t *Cb;                    // bf16 copy of the output (optional, ldc)
    int M, N, K;
    int accumulate;                  // C += result
    const float *bias;               // [N]
    int relu;
    const uint16_t *mask; int ldm;   // zero where mask (bf16 bits) is 0
    float clamp_max;                 // > 0: clamp to [-c, c] ...
    const uint8_t *col_mask;         // ... and zero columns whose mask is 0
    int kchunk;                      // K range per blockIdx.z (set by the launchers)
    size_t zstride;                  // output layer stride per blockIdx.z
};

namespace smc {

// One build of the advance kernel (is_variant.hip).
typedef hipError_t (*IsAdvanceLaunch)(const IsParams &p, int grid, int group_waves, size_t lds, hipStream_t s);
IsAdvanceLaunch is_advance_c1_o0();
IsAdvanceLaunch is_advance_c2_o0();
IsAdvanceLaunch is_advance_c4_o0();
IsAdvanceLaunch is_advance_c8_o0();

hipError_t is_launch_advance(const IsParams &p, int slots, size_t lds, int max_groups, hipStream_t s);
// Reset rows [0, rows) to the initial state of runs runs[i] (or run_begin + i);
// rows >= n_valid are inactive.  feat0 is the feature row of the initial state.
hipError_t is_init_rows(const IsParams &p, const uint64_t *runs, uint64_t run_begin, int n_valid,
                        const uint16_t *feat0, hipStream_t s);
// Results of the rows that are no longer active, into out[job].
hipError_t is_collect(const IsParams &p, const IsOut &out, hipStream_t s);
// Stable compaction of the active rows of p.st (and their feature rows in
// feat) into dst / dst_feat, at the positions pos (a record-slot scan with
// base 0).
hipError_t is_compact(const IsParams &p, const int32_t *pos, const IsRows &dst, const uint16_t *feat,
                      uint16_t *dst_feat, hipStream_t s);
// Record slots for the active rows, in row order, after *base; advances *base.
hipError_t is_record_slots(const uint8_t *status, int n_rows, int32_t *pos, unsigned int *base, unsigned int cap,
                           hipStream_t s);
// Cross-entropy gradient wrt the MLP outputs (see is_kernels.hip).
hipError_t is_ce_grad(const float *theta, const float *A, const float *B, const float *w, const uint8_t *tiltable,
                      float tilt_max, int rows, int lpad, float *G, hipStream_t s);
// tl[r][l] = tg[r][group[l]] (0 for padding locations).
hipError_t expand_groups(const float *tg, int gpad, const int32_t *group, int n_locs, int lpad, int rows, float *tl,
                         hipStream_t s);
// gg[r][g] = sum of gl[r][l] over the members l of group g (fp32, and bf16
// into ggb unless it is null).
hipError_t group_sum(const float *gl, int lpad, const int32_t *member_first, const int32_t *members, int n_groups,
                     int gpad, int rows, float *gg, uint16_t *ggb, hipStream_t s);
// C (+)= op(A) op(B) on MFMA 16x16x32: M, N multiples of 64, K of 32, lda
// and ldb of 8, 16-byte aligned operands; hipErrorInvalidValue otherwise.
hipError_t gemm(const GemmArgs &g, bool ta, bool tb, hipStream_t s);
// Whether mlp_forward handles these sizes (H 64/128/256, weights in LDS).
bool mlp_forward_supported(int fpad, int H, int g16, float clamp_max);
// Bytes of the packed weight image: W1, W2, W3 transposed to [n][k] bf16
// (rows padded by 8), then b1, b2, b3 (first g16) in fp32.
size_t mlp_image_bytes(int fpad, int H, int g16);
// Pack it from the flat fp32 parameters (offsets of W1, b1, W2, b2, W3, b3);
// bf16 rounding as to_bf16, so the fused pass multiplies the same values.
hipError_t mlp_pack(const float *w, size_t o_w1, size_t o_b1, size_t o_w2, size_t o_b2, size_t o_w3, size_t o_b3,
                    int fpad, int H, int gpad, int g16, void *image, hipStream_t s);
hipError_t mlp_forward(const MlpFwdArgs &a, hipStream_t s);
// dst (+)= op(A) op(B) with K split over `splits` workgroup layers (partial
// sums in `work`, splits * M * N floats, folded in a fixed order).
hipError_t gemm_splitk(const GemmArgs &g, bool ta, bool tb, int splits, float *work, hipStream_t s);
// dst[N] (+)= column sums of X[rows][ld] (first N columns), split over rows.
hipError_t colsum(const float *X, int rows, int ld, int N, float *dst, int splits, float *work, hipStream_t s);
hipError_t to_bf16(const float *x, uint16_t *y, size_t n, hipStream_t s);
hipError_t adam_step(float *w, const float *g, float *m1, float *m2, size_t n, float lr, int t, uint16_t *wb,
                     hipStream_t s);

}  // namespace smc
