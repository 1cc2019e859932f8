// is_kernels.hip — the MLP (MFMA GEMM with fused epilogues, output gradient,
// Adam), row set-up and record-slot kernels of importance sampling, and the
// launchers of the advance kernel's builds (is_advance.h, is_variant.hip).
#include <hip/hip_runtime.h>
#include <algorithm>

#include "is_advance.h"

namespace {

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
        st.job[row] = valid ? (uint32_t)row : 0xFFFFFFFFu;
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

__global__ void is_collect_kernel(IsParams P, IsOut out) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= P.n_rows) return;
    const IsRows &st = P.st;
    const uint8_t s = st.status[row];
    if (s & IS_ACTIVE) return;
    const uint32_t j = st.job[row];
    if (j == 0xFFFFFFFFu) return;  // padding
    out.status[j] = s;
    out.log2lr[j] = st.log2lr[row];
    out.score[j] = P.is_max ? st.qmax[row] : st.qmin[row];
    out.t[j] = st.t[row];
    out.step[j] = st.step[row];
    out.trans[j] = st.trans[row];
    out.digest[j] = st.digest[row];
}

// One workgroup per source row; active rows move to pos[row].
__global__ __launch_bounds__(64) void is_compact_kernel(IsParams P, const int32_t *pos, IsRows dst,
                                                        const uint16_t *feat, uint16_t *dst_feat) {
    const int row = blockIdx.x;
    const int p = pos[row];
    if (p < 0) return;
    const int tid = threadIdx.x;
    const IsRows &st = P.st;
    const int n = P.m.n_comp, nv = P.m.n_vars;
    for (int c = tid; c < n; c += 64) dst.loc[(size_t)p * n + c] = st.loc[(size_t)row * n + c];
    for (int v = tid; v < nv; v += 64) dst.vars[(size_t)p * nv + v] = st.vars[(size_t)row * nv + v];
    if (feat)
        for (int k = tid; k < P.fpad; k += 64) dst_feat[(size_t)p * P.fpad + k] = feat[(size_t)row * P.fpad + k];
    if (tid == 0) {
        dst.t[p] = st.t[row];
        dst.log2lr[p] = st.log2lr[row];
        dst.qmax[p] = st.qmax[row];
        dst.qmin[p] = st.qmin[row];
        dst.step[p] = st.step[row];
        dst.trans[p] = st.trans[row];
        dst.chan[p] = st.chan[row];
        dst.sender[p] = st.sender[row];
        dst.status[p] = st.status[row];
        dst.digest[p] = st.digest[row];
        dst.run[p] = st.run[row];
        dst.job[p] = st.job[row];
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
// fixed member order (CSR member_first / members); one thread per (row,
// group).  The padding outputs are cleared by the launcher.
__global__ void group_sum_kernel(const float *gl, int lpad, const int32_t *member_first, const int32_t *members,
                                 int n_groups, int gpad, size_t n, float *gg, uint16_t *ggb) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int gidx = (int)(i % (size_t)n_groups);
    const size_t row = i / (size_t)n_groups;
    const float *src = gl + row * lpad;
    float s = 0.f;
    for (int k = member_first[gidx]; k < member_first[gidx + 1]; ++k) s += src[members[k]];
    gg[row * gpad + gidx] = s;
    if (ggb) ggb[row * gpad + gidx] = f32_to_bf16(s);
}

__global__ void clear_columns_kernel(float *gg, uint16_t *ggb, int from, int gpad, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int w = gpad - from;
    if (i >= n * (size_t)w) return;
    const size_t row = i / (size_t)w;
    const int col = from + (int)(i % (size_t)w);
    gg[row * gpad + col] = 0.f;
    if (ggb) ggb[row * gpad + col] = 0;
}

// Column sums over a slice of rows (blockIdx.y), 64 columns per workgroup.
__global__ __launch_bounds__(64) void colsum_kernel(const float *X, int rows, int ld, int N, int rows_per,
                                                    float *part) {
    const int col = blockIdx.x * 64 + threadIdx.x;
    if (col >= N) return;
    const int r0 = blockIdx.y * rows_per, r1 = min(rows, r0 + rows_per);
    float s = 0.f;
    // unrolled so the loads overlap; the sum stays in row order
#pragma unroll 8
    for (int r = r0; r < r1; ++r) s += X[(size_t)r * ld + col];
    part[(size_t)blockIdx.y * N + col] = s;
}

// dst[i] (+)= sum_z part[z][i], z in order.
__global__ void fold_splits_kernel(const float *part, int splits, size_t n, float *dst, int accumulate) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
#pragma unroll 8
    for (int z = 0; z < splits; ++z) s += part[(size_t)z * n + i];
    dst[i] = accumulate ? dst[i] + s : s;
}

// ------------------------------------------------------------ GEMM
//
// C[M x N] (+)= op(A)[M x K] . op(B)[K x N], bf16 inputs, fp32 accumulation
// on gfx950's MFMA 16x16x32.  A workgroup of four waves computes a 64 x 64
// tile; each wave a 32 x 32 quarter as 2 x 2 MFMA blocks; K advances 32 at a
// time through LDS.  Operand layout of v_mfma_f32_16x16x32_bf16: lane l
// holds A[l % 16][8 (l / 16) .. +7] and B[8 (l / 16) .. +7][l % 16] (eight
// bf16, one 16-byte LDS read each: A is staged [m][k], B transposed [n][k]);
// the result lane l holds rows 4 (l / 16) .. +3 of column l % 16.  M, N are
// multiples of 64, K of 32, leading dimensions and base addresses of 8
// elements (the host pads every buffer).  Each output element sums K in a
// fixed order, so a row's result does not depend on the other rows.

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Staged tile rows: 32 bf16 of K plus 8 of padding, so 16-byte fragment
// reads stay aligned and eight consecutive rows start in distinct banks.
constexpr int kGemmKP = 32 + 8;

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g) {
    __shared__ __attribute__((aligned(16))) uint16_t As[64][kGemmKP];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[64][kGemmKP];
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
    // A and B tiles of 64 x 32, 2048 elements each: every thread moves eight
    // consecutive elements of each with one 16-byte load; the next tile's
    // loads are issued before the current tile's MFMAs
    const int ta_k = tid >> 3, ta_i = 8 * (tid & 7);  // A stored [K][M]
    const int na_i = tid >> 2, na_k = 8 * (tid & 3);  // A stored [M][K]
    const int tb_j = tid >> 2, tb_k = 8 * (tid & 3);  // B stored [N][K]
    const int nb_k = tid >> 3, nb_j = 8 * (tid & 7);  // B stored [K][N]
    auto load_a = [&](int k0) -> uint4 {
        return TA ? *(const uint4 *)(g.A + (size_t)(k0 + ta_k) * g.lda + (m0 + ta_i))
                  : *(const uint4 *)(g.A + (size_t)(m0 + na_i) * g.lda + (k0 + na_k));
    };
    auto load_b = [&](int k0) -> uint4 {
        return TB ? *(const uint4 *)(g.B + (size_t)(n0 + tb_j) * g.ldb + (k0 + tb_k))
                  : *(const uint4 *)(g.B + (size_t)(k0 + nb_k) * g.ldb + (n0 + nb_j));
    };
    // eight elements of one k, consecutive in m (or n), into a [m][k] tile
    auto scatter = [](uint16_t (*T)[kGemmKP], int i0, int k, uint4 v) {
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            T[i0 + 2 * q][k] = (uint16_t)w[q];
            T[i0 + 2 * q + 1][k] = (uint16_t)(w[q] >> 16);
        }
    };
    uint4 ra = kz < kend ? load_a(kz) : uint4{0u, 0u, 0u, 0u};
    uint4 rb = kz < kend ? load_b(kz) : uint4{0u, 0u, 0u, 0u};
    for (int k0 = kz; k0 < kend; k0 += 32) {
        if (TA) scatter(As, ta_i, ta_k, ra);
        else *(uint4 *)&As[na_i][na_k] = ra;
        if (TB) *(uint4 *)&Bs[tb_j][tb_k] = rb;
        else scatter(Bs, nb_j, nb_k, rb);
        __syncthreads();
        if (k0 + 32 < kend) {
            ra = load_a(k0 + 32);
            rb = load_b(k0 + 32);
        }
        const int kb = 8 * (l >> 4);
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const bf16x8 fa = *(const bf16x8 *)&As[wr + a * 16 + (l & 15)][kb];
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const bf16x8 fb = *(const bf16x8 *)&Bs[wc + b * 16 + (l & 15)][kb];
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa, fb, acc[a][b], 0, 0, 0);
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

// ------------------------------------------------------------ fused forward
//
// Persistent workgroups of eight waves.  The workgroup stages the three
// weight matrices transposed ([n][k], rows padded by 8 bf16: aligned 16-byte
// fragment reads, distinct banks for eight consecutive rows) and the biases
// in LDS once; each wave then takes 16-row tiles: layer 1 reads its A
// fragments straight from the feature rows, layers 2 and 3 from the wave's
// activation tile in LDS.  MFMA 16x16x32 bf16 with fp32 accumulation, K in
// steps of 32 from 0 upwards, as gemm_kernel, so both give the same bits.

// Waves per workgroup: two per SIMD share one copy of the weights, so one
// wave's memory waits overlap the other's MFMAs.
constexpr int kMlpWaves = 8;

__host__ __device__ constexpr size_t mlp_image_size(int fpad, int H, int g16) {
    return 2 * ((size_t)H * (fpad + 8) + (size_t)H * (H + 8) + (size_t)g16 * (H + 8)) + 4 * (2 * (size_t)H + g16);
}

__host__ __device__ constexpr size_t mlp_lds_bytes(int fpad, int H, int g16) {
    return mlp_image_size(fpad, H, g16) + 2 * (size_t)kMlpWaves * 16 * (H + 8);
}

// The packed image, one thread per element: bf16 part, then fp32 biases.
__global__ void mlp_pack_kernel(const float *w, size_t o_w1, size_t o_b1, size_t o_w2, size_t o_b2, size_t o_w3,
                                size_t o_b3, int fpad, int H, int gpad, int g16, uint16_t *img) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t KP1 = fpad + 8, KP = H + 8;
    const size_t n1 = (size_t)H * KP1, n2 = (size_t)H * KP, n3 = (size_t)g16 * KP;
    if (i < n1) {
        const size_t n = i / KP1, k = i % KP1;
        img[i] = k < (size_t)fpad ? f32_to_bf16(w[o_w1 + k * H + n]) : 0;
    } else if (i < n1 + n2) {
        const size_t j = i - n1, n = j / KP, k = j % KP;
        img[i] = k < (size_t)H ? f32_to_bf16(w[o_w2 + k * H + n]) : 0;
    } else if (i < n1 + n2 + n3) {
        const size_t j = i - n1 - n2, n = j / KP, k = j % KP;
        img[i] = k < (size_t)H ? f32_to_bf16(w[o_w3 + k * gpad + n]) : 0;
    } else {
        const size_t j = i - n1 - n2 - n3;
        float *b = (float *)(img + n1 + n2 + n3);
        if (j < (size_t)H) b[j] = w[o_b1 + j];
        else if (j < 2 * (size_t)H) b[j] = w[o_b2 + (j - H)];
        else if (j < 2 * (size_t)H + g16) b[j] = w[o_b3 + (j - 2 * H)];
    }
}

constexpr int kMlpMaxK = 16;  // layer-1 k-steps of 32 held in registers: fpad <= 512

template <int H>
__global__ __launch_bounds__(64 * kMlpWaves) void mlp_forward_kernel(MlpFwdArgs a) {
    extern __shared__ __attribute__((aligned(16))) uint16_t smem[];
    constexpr int NT = H / 16, KP = H + 8;
    const int KP1 = a.fpad + 8, g16 = a.g16;
    uint16_t *w1t = smem;
    uint16_t *w2t = w1t + (size_t)H * KP1;
    uint16_t *w3t = w2t + (size_t)H * KP;
    float *b1s = (float *)(w3t + (size_t)g16 * KP);
    float *b2s = b1s + H;
    float *b3s = b2s + H;
    const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
    uint16_t *hb = (uint16_t *)(b3s + g16) + (size_t)wave * 16 * KP;

    copy_to_lds((const SMC_GLOBAL uint4 *)a.image, (SMC_LDS uint4 *)(SMC_LDS uint16_t *)smem,
                (uint32_t)(mlp_image_size(a.fpad, H, g16) / 16));
    __syncthreads();

    const int ar = l & 15, kq = 8 * (l >> 4);  // A/B fragments: row (column) and first k
    const int cr = 4 * (l >> 4), cc = l & 15;  // C fragment: rows cr..cr+3, column cc
    const int tiles = a.rows / 16;
    for (int t = blockIdx.x * kMlpWaves + wave; t < tiles; t += gridDim.x * kMlpWaves) {
        const int r0 = t * 16;
        f32x4 acc[NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
        // layer 1: the tile's A fragments are loaded first, all at once (the
        // loads overlap instead of each k-step waiting on memory)
        const uint16_t *xrow = a.X + (size_t)(r0 + ar) * a.fpad + kq;
        const int KT = a.fpad / 32;
        bf16x8 xa[kMlpMaxK];
#pragma unroll
        for (int kt = 0; kt < kMlpMaxK; ++kt)
            if (kt < KT) xa[kt] = *(const bf16x8 *)(xrow + kt * 32);
#pragma unroll
        for (int kt = 0; kt < kMlpMaxK; ++kt)
            if (kt < KT) {
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) {
                    const bf16x8 fb = *(const bf16x8 *)(w1t + (size_t)(nt * 16 + ar) * KP1 + kt * 32 + kq);
                    acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xa[kt], fb, acc[nt], 0, 0, 0);
                }
            }
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int col = nt * 16 + cc;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float v = acc[nt][i] + b1s[col];
                v = v > 0.f ? v : 0.f;
                const uint16_t h = f32_to_bf16(v);
                hb[(cr + i) * KP + col] = h;
                if (a.a1) a.a1[(size_t)(r0 + cr + i) * H + col] = h;
            }
            acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
        wave_sync();
        // layer 2
#pragma unroll 2
        for (int k0 = 0; k0 < H; k0 += 32) {
            const bf16x8 fa = *(const bf16x8 *)(hb + ar * KP + k0 + kq);
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const bf16x8 fb = *(const bf16x8 *)(w2t + (nt * 16 + ar) * KP + k0 + kq);
                acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa, fb, acc[nt], 0, 0, 0);
            }
        }
        wave_sync();
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int col = nt * 16 + cc;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float v = acc[nt][i] + b2s[col];
                v = v > 0.f ? v : 0.f;
                const uint16_t h = f32_to_bf16(v);
                hb[(cr + i) * KP + col] = h;
                if (a.a2) a.a2[(size_t)(r0 + cr + i) * H + col] = h;
            }
        }
        wave_sync();
        // layer 3 (at most four 16-column tiles)
        f32x4 o[4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) o[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
        for (int k0 = 0; k0 < H; k0 += 32) {
            const bf16x8 fa = *(const bf16x8 *)(hb + ar * KP + k0 + kq);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
                if (nt * 16 < g16) {
                    const bf16x8 fb = *(const bf16x8 *)(w3t + (nt * 16 + ar) * KP + k0 + kq);
                    o[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa, fb, o[nt], 0, 0, 0);
                }
        }
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
            if (nt * 16 < g16) {
                const int col = nt * 16 + cc;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    float v = o[nt][i] + b3s[col];
                    v = v > a.clamp_max ? a.clamp_max : (v < -a.clamp_max ? -a.clamp_max : v);
                    if (!a.col_mask[col]) v = 0.f;
                    a.og[(size_t)(r0 + cr + i) * a.ldo + col] = v;
                }
            }
        for (int c = g16 + l; c < a.ldo; c += 64)
            for (int r = 0; r < 16; ++r) a.og[(size_t)(r0 + r) * a.ldo + c] = 0.f;
        wave_sync();  // the next tile overwrites hb
    }
}

}  // namespace

namespace smc {

hipError_t is_launch_advance(const IsParams &p, int slots, size_t lds, int max_groups, hipStream_t s) {
    const int R = 64 / p.lanes;
    const int row_groups = (p.n_rows + R - 1) / R;
    if (row_groups <= 0) return hipSuccess;
    // staged tables: workgroups of four waves, at most max_groups of them
    // (persistent: each wave loops over row groups); else one wave per
    // workgroup and one row group per wave
    const int W = p.table_bytes ? 4 : 1;
    int grid = (row_groups + W - 1) / W;
    if (p.table_bytes && max_groups > 0 && grid > max_groups) grid = max_groups;
    IsAdvanceLaunch f = nullptr;
    switch (slots) {
    case 1: f = is_advance_c1_o0(); break;
    case 2: f = is_advance_c2_o0(); break;
    case 4: f = is_advance_c4_o0(); break;
    case 8: f = is_advance_c8_o0(); break;
    default: return hipErrorInvalidValue;
    }
    return f(p, grid, W, lds, s);
}

static hipError_t gemm_launch(const GemmArgs &g, bool ta, bool tb, int zsplit, hipStream_t s) {
    dim3 grid(g.N / 64, g.M / 64, zsplit);
    if (!ta && !tb) hipLaunchKernelGGL((gemm_kernel<false, false>), grid, dim3(256), 0, s, g);
    else if (ta && !tb) hipLaunchKernelGGL((gemm_kernel<true, false>), grid, dim3(256), 0, s, g);
    else if (!ta && tb) hipLaunchKernelGGL((gemm_kernel<false, true>), grid, dim3(256), 0, s, g);
    else hipLaunchKernelGGL((gemm_kernel<true, true>), grid, dim3(256), 0, s, g);
    return hipGetLastError();
}

static int max_lds_per_block() {
    static int v = -1;
    if (v < 0) {
        int dev = 0, x = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&x, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess)
            v = x;
        else
            v = 0;
    }
    return v;
}

bool mlp_forward_supported(int fpad, int H, int g16, float clamp_max) {
    if (!(H == 64 || H == 128 || H == 256) || fpad <= 0 || fpad % 32 || fpad > 32 * kMlpMaxK || g16 <= 0 ||
        g16 % 16 || g16 > 64 ||
        !(clamp_max > 0.f))
        return false;
    return mlp_lds_bytes(fpad, H, g16) <= (size_t)max_lds_per_block();
}

size_t mlp_image_bytes(int fpad, int H, int g16) { return mlp_image_size(fpad, H, g16); }

hipError_t mlp_pack(const float *w, size_t o_w1, size_t o_b1, size_t o_w2, size_t o_b2, size_t o_w3, size_t o_b3,
                    int fpad, int H, int gpad, int g16, void *image, hipStream_t s) {
    const size_t n = (size_t)H * (fpad + 8) + (size_t)H * (H + 8) + (size_t)g16 * (H + 8) + 2 * (size_t)H + g16;
    hipLaunchKernelGGL(mlp_pack_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, w, o_w1, o_b1, o_w2, o_b2,
                       o_w3, o_b3, fpad, H, gpad, g16, (uint16_t *)image);
    return hipGetLastError();
}

hipError_t mlp_forward(const MlpFwdArgs &a, hipStream_t s) {
    if (a.rows <= 0 || a.rows % 16 || a.g16 > a.gpad || a.ldo < a.g16 || !a.og || !a.X || !a.image ||
        !mlp_forward_supported(a.fpad, a.H, a.g16, a.clamp_max))
        return hipErrorInvalidValue;
    const size_t lds = mlp_lds_bytes(a.fpad, a.H, a.g16);
    const void *fn = a.H == 64 ? (const void *)mlp_forward_kernel<64>
                   : a.H == 128 ? (const void *)mlp_forward_kernel<128> : (const void *)mlp_forward_kernel<256>;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    int dev = 0, cus = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
    const int tiles = a.rows / 16;
    const int grid = std::max(1, std::min((tiles + kMlpWaves - 1) / kMlpWaves, cus));
    const dim3 block(64 * kMlpWaves);
    if (a.H == 64) hipLaunchKernelGGL(mlp_forward_kernel<64>, dim3(grid), block, lds, s, a);
    else if (a.H == 128) hipLaunchKernelGGL(mlp_forward_kernel<128>, dim3(grid), block, lds, s, a);
    else hipLaunchKernelGGL(mlp_forward_kernel<256>, dim3(grid), block, lds, s, a);
    return hipGetLastError();
}

hipError_t gemm(const GemmArgs &g0, bool ta, bool tb, hipStream_t s) {
    if (g0.M % 64 || g0.N % 64 || g0.K % 32 || g0.M <= 0 || g0.N <= 0 || g0.K <= 0 || g0.lda % 8 || g0.ldb % 8)
        return hipErrorInvalidValue;
    GemmArgs g = g0;
    g.kchunk = g.K;
    g.zstride = 0;
    return gemm_launch(g, ta, tb, 1, s);
}

hipError_t gemm_splitk(const GemmArgs &g0, bool ta, bool tb, int splits, float *work, hipStream_t s) {
    if (g0.M % 64 || g0.N % 64 || g0.K % 32 || g0.M <= 0 || g0.N <= 0 || g0.K <= 0 || g0.lda % 8 || g0.ldb % 8 ||
        !g0.C)
        return hipErrorInvalidValue;
    if (g0.bias || g0.relu || g0.mask || g0.clamp_max > 0.f || g0.Cb || g0.ldc != g0.N) return hipErrorInvalidValue;
    if (splits <= 1) return gemm(g0, ta, tb, s);
    int kchunk = (g0.K + splits - 1) / splits;
    kchunk = (kchunk + 31) / 32 * 32;
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
    if (rows <= 0) return hipSuccess;
    const size_t n = (size_t)rows * n_groups;
    if (n)
        hipLaunchKernelGGL(group_sum_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, gl, lpad, member_first,
                           members, n_groups, gpad, n, gg, ggb);
    if (gpad > n_groups) {
        const size_t z = (size_t)rows * (gpad - n_groups);
        hipLaunchKernelGGL(clear_columns_kernel, dim3((unsigned)((z + 255) / 256)), dim3(256), 0, s, gg, ggb, n_groups,
                           gpad, (size_t)rows);
    }
    return hipGetLastError();
}

hipError_t is_init_rows(const IsParams &p, const uint64_t *runs, uint64_t run_begin, int n_valid,
                        const uint16_t *feat0, hipStream_t s) {
    if (p.n_rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(is_init_kernel, dim3(p.n_rows), dim3(64), 0, s, p, runs, run_begin, n_valid, feat0);
    return hipGetLastError();
}

hipError_t is_collect(const IsParams &p, const IsOut &out, hipStream_t s) {
    if (p.n_rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(is_collect_kernel, dim3((unsigned)((p.n_rows + 255) / 256)), dim3(256), 0, s, p, out);
    return hipGetLastError();
}

hipError_t is_compact(const IsParams &p, const int32_t *pos, const IsRows &dst, const uint16_t *feat,
                      uint16_t *dst_feat, hipStream_t s) {
    if (p.n_rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(is_compact_kernel, dim3(p.n_rows), dim3(64), 0, s, p, pos, dst, feat, dst_feat);
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
