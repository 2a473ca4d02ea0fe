/*
 * Deterministic exp(x) for |x| <= 700, bit-identical on the host (gcc) and on
 * gfx950 (hipcc): only IEEE-754 double +, -, *, exact power-of-two scaling and
 * integer conversions are used, each correctly rounded, in a fixed order with
 * floating-point contraction disabled.  It decides Metropolis acceptance
 * (u < exp(delta / T)) in the annealing chains, so that a CPU replay of a
 * chain takes exactly the same decisions as the device.
 *
 * Method (Cody & Waite range reduction, Taylor polynomial):
 *   x = k ln2 + r, k = nearest integer to x / ln2, |r| <= ln2 / 2,
 *   ln2 split as LN2_HI (exactly representable, trailing zero bits) + LN2_LO;
 *   exp(r) = sum_{i=0}^{13} r^i / i! by Horner (truncation error < 2^-53
 *   for |r| <= 0.35);  exp(x) = 2^k exp(r).
 * The absolute accuracy is about 1 ulp, which does not matter here: what the
 * chains need is the same value on both sides.
 */
#ifndef BM_DETEXP_H
#define BM_DETEXP_H

#include <stdint.h>

#if defined(__HIPCC__)
#define BM_DX_HD __host__ __device__
#else
#define BM_DX_HD static
#endif

#if defined(__clang__)
#define BM_NO_CONTRACT _Pragma("clang fp contract(off)")
#else
#define BM_NO_CONTRACT
#endif

BM_DX_HD inline double bm_pow2i(int k) {
    /* 2^k for -1022 <= k <= 1023 by building the exponent field */
    union {
        uint64_t u;
        double d;
    } c;
    c.u = (uint64_t)(k + 1023) << 52;
    return c.d;
}

BM_DX_HD inline double bm_det_exp(double x) {
    BM_NO_CONTRACT
    if (x != x) return 1.0;                          /* NaN: treated as exp(0) */
    if (x < -700.0) return 0.0;
    if (x > 700.0) return 1.0e304;                   /* finite stand-in for overflow */
    const double INV_LN2 = 1.4426950408889634;
    const double LN2_HI = 6.93147180369123816490e-01; /* 0x3FE62E42FEE00000 */
    const double LN2_LO = 1.90821492927058770002e-10; /* 0x3DEA39EF35793C76 */
    const double t = x * INV_LN2;
    const int k = (int)(t < 0.0 ? t - 0.5 : t + 0.5); /* nearest, halves away from 0 */
    const double kd = (double)k;
    const double r = (x - kd * LN2_HI) - kd * LN2_LO;
    /* Horner for sum r^i / i!, i = 0..13 */
    double p = 1.0 / 6227020800.0;                   /* 1/13! */
    p = p * r + 1.0 / 479001600.0;                   /* 1/12! */
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    if (k < -1022) return 0.0;
    return p * bm_pow2i(k);
}

#endif
