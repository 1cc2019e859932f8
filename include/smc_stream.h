/*
 * smc_stream.h — the random stream of the simulator, shared bit for bit by
 * the GPU run kernel and every host implementation of the step semantics.
 *
 * cuSMC draws from per-(run, component) curand XORWOW states
 * (simulation.cu:57,515,529,758); those states are shared by every block
 * (smc.cu:228), so its streams are not reproducible.  Here they are
 * replaced by a counter-based generator: Philox4x32-10 (Salmon, Moraes, Dror,
 * Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11), keyed by the
 * seed and addressed by (step, purpose, component, run).  Any number of
 * draws can be made in any order on any device and a host reimplementation
 * gets the same bits.
 *
 * Everything here uses only integer operations and IEEE-754 double
 * add/sub/mul/div/fma, written as explicit calls, so that a host compiler and
 * the gfx950 compiler produce identical results (compile with
 * -ffp-contract=off on both sides).
 */
#ifndef SMC_STREAM_H
#define SMC_STREAM_H

#include <stdint.h>
#include <string.h>
#include <math.h>

#if defined(__HIPCC__)
#define SMC_HD __host__ __device__ __forceinline__
#else
#define SMC_HD static inline
#endif

#ifdef __cplusplus
extern "C++" {
#endif

/* ---------------------------------------------------------------- Philox */

struct smc_u32x4 { uint32_t v[4]; };

SMC_HD void smc_mulhilo32(uint32_t a, uint32_t b, uint32_t *hi, uint32_t *lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

/* Philox4x32 with 10 rounds; constants from the SC'11 paper. */
SMC_HD smc_u32x4 smc_philox4x32_10(smc_u32x4 ctr, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t c0 = ctr.v[0], c1 = ctr.v[1], c2 = ctr.v[2], c3 = ctr.v[3];
#if defined(__HIPCC__)
#pragma unroll
#endif
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        smc_mulhilo32(M0, c0, &hi0, &lo0);
        smc_mulhilo32(M1, c2, &hi1, &lo1);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += W0; k1 += W1;
    }
    smc_u32x4 out;
    out.v[0] = c0; out.v[1] = c1; out.v[2] = c2; out.v[3] = c3;
    return out;
}

/* --------------------------------------------------------- stream layout */

/* Purposes of a draw.  A draw is addressed by
 *   counter = (step, purpose << 24 | index, run_lo, run_hi),
 *   key     = (seed_lo, seed_hi).
 * Within one step each (purpose, index) is drawn at most once, and each of
 * its words serves one component, so no two uses of one run share bits.
 * Neighbouring components share a draw where one needs fewer than 128 bits:
 * the uniform delays of components 2g and 2g+1 are draw g of SMC_P_UNIFORM
 * (words 0-1 and 2-3), the receiver choices of components 4g .. 4g+3 are
 * draw g of SMC_P_RECEIVER (word c & 3), so a lane holding several
 * components makes one call where it would make two or four. */
enum smc_purpose {
    SMC_P_RACE = 0,     /* per run and step: exponential race (words 0-1: the
                           race delay, words 2-3: which exponential component) */
    SMC_P_UNIFORM = 1,  /* per pair of components: delay under a clock upper
                           bound (53 bits each)                              */
    SMC_P_WINNER = 2,   /* per component: weighted edge choice of the winner  */
    SMC_P_RECEIVER = 3, /* per four components: weighted choice of a broadcast
                           receiver (32 bits each)                           */
    SMC_P_TILT = 4,     /* importance sampling: spare stream                  */
    SMC_P_DEAD = 5      /* per run and step: first null event after the time
                           bound of the components that cannot fire        */
};

SMC_HD smc_u32x4 smc_draw(uint64_t seed, uint64_t run, uint32_t step,
                          uint32_t purpose, uint32_t comp) {
    smc_u32x4 c;
    c.v[0] = step;
    c.v[1] = (purpose << 24) | (comp & 0xFFFFFFu);
    c.v[2] = (uint32_t)run;
    c.v[3] = (uint32_t)(run >> 32);
    return smc_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

/* Uniform double on (0, 1]: 53 random bits k, u = (k + 1) * 2^-53 (exact). */
SMC_HD double smc_u01(uint32_t hi, uint32_t lo) {
    uint64_t k = ((uint64_t)hi << 21) | (uint64_t)(lo >> 11);
    return (double)(k + 1u) * 1.1102230246251565404236316680908203125e-16;
}

/* Uniform double on (0, 1] from one word: u = (k + 1) * 2^-32 (exact).  A
 * weighted choice among edges of integer weights takes floor(W u), so each
 * edge's probability is off its weight share by less than 2^-32 (cuSMC's
 * curand_uniform draws carry 24 bits). */
SMC_HD double smc_u01_32(uint32_t k) {
    return ((double)k + 1.0) * 2.3283064365386962890625e-10;
}

/* The draw and words of component c's uniform delay and receiver choice. */
SMC_HD uint32_t smc_uniform_draw(uint32_t c) { return c >> 1; }
SMC_HD double smc_uniform_u(smc_u32x4 w, uint32_t c) {
    return (c & 1u) ? smc_u01(w.v[2], w.v[3]) : smc_u01(w.v[0], w.v[1]);
}
SMC_HD uint32_t smc_receiver_draw(uint32_t c) { return c >> 2; }
SMC_HD double smc_receiver_u(smc_u32x4 w, uint32_t c) {
    const uint32_t a = (c & 1u) ? w.v[1] : w.v[0];
    const uint32_t b = (c & 1u) ? w.v[3] : w.v[2];
    return smc_u01_32((c & 2u) ? b : a);
}

/* ------------------------------------------------------------- log2 */

SMC_HD uint64_t smc_bits(double x) { uint64_t b; memcpy(&b, &x, 8); return b; }
SMC_HD double smc_from_bits(uint64_t b) { double x; memcpy(&x, &b, 8); return x; }

/*
 * log2(x) for finite normal x > 0.
 * Argument reduction x = m * 2^e with m in (sqrt(1/2), sqrt(2)], then
 * ln(m) = 2 atanh(s), s = (m-1)/(m+1), |s| < 0.1716, summed as the series
 * 2 s (1 + s^2/3 + s^4/5 + ... + s^22/23) (Abramowitz & Stegun 4.1.27), whose
 * truncation error is below 2^-60 relative on that interval.  Accuracy is a
 * few ulp; what matters is that host and device produce the same bits.
 */
SMC_HD double smc_log2(double x) {
    uint64_t b = smc_bits(x);
    int e = (int)((b >> 52) & 0x7FF) - 1023;
    double m = smc_from_bits((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
    double s = (m - 1.0) / (m + 1.0);
    double z = s * s;
    double p = 1.0 / 23.0;
    p = fma(p, z, 1.0 / 21.0);
    p = fma(p, z, 1.0 / 19.0);
    p = fma(p, z, 1.0 / 17.0);
    p = fma(p, z, 1.0 / 15.0);
    p = fma(p, z, 1.0 / 13.0);
    p = fma(p, z, 1.0 / 11.0);
    p = fma(p, z, 1.0 / 9.0);
    p = fma(p, z, 1.0 / 7.0);
    p = fma(p, z, 1.0 / 5.0);
    p = fma(p, z, 1.0 / 3.0);
    p = fma(p, z, 1.0);
    double ln_m = (2.0 * s) * p;
    return fma(ln_m, 1.4426950408889634074, (double)e);
}

/* 2^y for |y| < 1022, used by the power operator for non-integer exponents.
 * y = n + f with f in [-1/2, 1/2]; 2^f = exp(f ln2) by its Taylor series to
 * degree 15 (remainder < 2^-60). */
SMC_HD double smc_exp2(double y) {
    if (y > 1023.0) return HUGE_VAL;
    if (y < -1022.0) return 0.0;
    double n = floor(y + 0.5);
    double t = (y - n) * 0.6931471805599453094;
    double p = 1.0 / 1307674368000.0;
    p = fma(p, t, 1.0 / 87178291200.0);
    p = fma(p, t, 1.0 / 6227020800.0);
    p = fma(p, t, 1.0 / 479001600.0);
    p = fma(p, t, 1.0 / 39916800.0);
    p = fma(p, t, 1.0 / 3628800.0);
    p = fma(p, t, 1.0 / 362880.0);
    p = fma(p, t, 1.0 / 40320.0);
    p = fma(p, t, 1.0 / 5040.0);
    p = fma(p, t, 1.0 / 720.0);
    p = fma(p, t, 1.0 / 120.0);
    p = fma(p, t, 1.0 / 24.0);
    p = fma(p, t, 1.0 / 6.0);
    p = fma(p, t, 0.5);
    p = fma(p, t, 1.0);
    p = fma(p, t, 1.0);
    int ni = (int)n;
    return p * smc_from_bits((uint64_t)(ni + 1023) << 52);
}

/* Power with a reproducible result: binary exponentiation for integer
 * exponents |b| <= 1024, otherwise 2^(b log2 a) for a > 0, NaN for a < 0
 * with a non-integer exponent, and the IEEE special cases for a == 0. */
#if defined(__HIP_DEVICE_COMPILE__)
/* Out of line on the GPU: models rarely raise to powers, and inlined into
 * every expression site its temporaries set the kernels' register peak. */
__attribute__((noinline)) __device__ double smc_pow(double a, double b) {
#else
SMC_HD double smc_pow(double a, double b) {
#endif
    if (b == floor(b) && fabs(b) <= 1024.0) {
        long n = (long)(b < 0 ? -b : b);
        double r = 1.0, x = a;
        while (n) { if (n & 1) r = r * x; x = x * x; n >>= 1; }
        return b < 0 ? 1.0 / r : r;
    }
    if (a > 0.0) return smc_exp2(b * smc_log2(a));
    if (a == 0.0) return b > 0 ? 0.0 : HUGE_VAL;
    return NAN;
}

/* Bits of a double for the trace digest: every NaN maps to one pattern
 * (hosts and GPUs produce NaNs with different sign and payload bits). */
SMC_HD uint64_t smc_digest_bits(double x) {
    return x != x ? 0x7FF8000000000000ull : smc_bits(x);
}

/* Mixing step of the per-run trace digest (splitmix64 finaliser). */
SMC_HD uint64_t smc_mix(uint64_t h, uint64_t x) {
    uint64_t z = h ^ (x + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2));
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

#ifdef __cplusplus
}
#endif

#endif /* SMC_STREAM_H */
