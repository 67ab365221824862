/*
 * orc_math.h -- TEST INFRASTRUCTURE ONLY.
 *
 * Portable FP64 exp, log, pow and erfc built from IEEE +, -, *, /, sqrt, floor,
 * frexp and ldexp only, so that a CPU build and a GPU build that evaluate them
 * with the same operation sequence (and no fused multiply-add) give identical
 * bits.  The oracle is built twice: with the host libm (the faithful reading of
 * the reference, which calls the Fortran runtime's intrinsics) and with these
 * functions (-DORC_PORTABLE_MATH), which the HIP kernel's copy reproduces bit
 * for bit.
 *
 * Methods:
 *   exp   Cody & Waite (1980) argument reduction x = k ln2 + r with ln2 split in
 *         two (the high part has zero low bits, so k*ln2_hi is exact), then the
 *         Taylor series of e^r - 1 to degree 14 (|r| <= 0.347, truncation
 *         < 1e-18) and a final exact scaling by 2^k.
 *   log   x = m 2^e with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s),
 *         s = (m-1)/(m+1), |s| <= 0.1716, series to s^23 (truncation < 1e-18).
 *   pow   exp(y log x) for x > 0 (the scheme raises positive pressure ratios).
 *   erfc  x < 2: 1 - erf(x) with erf from its everywhere-positive series
 *         erf(x) = 2x/sqrt(pi) e^{-x^2} sum_n (2x^2)^n / (1*3*...*(2n+1));
 *         x >= 2: the Laplace continued fraction
 *         erfc(x) = e^{-x^2}/sqrt(pi) / (x + (1/2)/(x + 1/(x + (3/2)/(x + ...)))),
 *         evaluated backwards from depth 60.  Relative error below 1e-13.
 */
#ifndef ORC_MATH_H
#define ORC_MATH_H

#include <math.h>

static inline double orc_pm_exp(double x)
{
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    const double inv_ln2 = 1.44269504088896338700e+00;
    if (x != x) return x;
    if (x > 709.782712893384) return HUGE_VAL;
    if (x < -745.1332191019412) return 0.0;
    const double kd = floor(x * inv_ln2 + 0.5);
    const double r = (x - kd * ln2_hi) - kd * ln2_lo;
    double p = 1.0 / 87178291200.0;          /* 1/14! */
    p = p * r + 1.0 / 6227020800.0;          /* 1/13! */
    p = p * r + 1.0 / 479001600.0;
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
    const double em1 = r + (r * r) * p;
    return ldexp(1.0 + em1, (int)kd);
}

static inline double orc_pm_log(double x)
{
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    if (x != x) return x;
    if (x <= 0.0) return x == 0.0 ? -HUGE_VAL : (x - x) / 0.0;
    if (x == HUGE_VAL) return x;
    int e;
    double m = frexp(x, &e);
    if (m < 0.70710678118654752440) {
        m = m + m;
        e = e - 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double p = 1.0 / 23.0;
    p = p * z + 1.0 / 21.0;
    p = p * z + 1.0 / 19.0;
    p = p * z + 1.0 / 17.0;
    p = p * z + 1.0 / 15.0;
    p = p * z + 1.0 / 13.0;
    p = p * z + 1.0 / 11.0;
    p = p * z + 1.0 / 9.0;
    p = p * z + 1.0 / 7.0;
    p = p * z + 1.0 / 5.0;
    p = p * z + 1.0 / 3.0;
    const double logm = (s + s) + (s + s) * (z * p);
    const double ed = (double)e;
    return ed * ln2_hi + (ed * ln2_lo + logm);
}

static inline double orc_pm_pow(double x, double y)
{
    if (x == 0.0) return y > 0.0 ? 0.0 : (y == 0.0 ? 1.0 : HUGE_VAL);
    if (y == 0.0) return 1.0;
    return orc_pm_exp(y * orc_pm_log(x));
}

static inline double orc_pm_erfc(double x)
{
    const double two_over_sqrtpi = 1.12837916709551257390;
    const double inv_sqrtpi = 0.56418958354775628695;
    if (x != x) return x;
    if (x < 0.0) return 2.0 - orc_pm_erfc(-x);
    if (x < 2.0) {
        const double t = 2.0 * (x * x);
        double term = 1.0, sum = 1.0;
        for (int n = 1; n <= 48; ++n) {
            term = term * t / (double)(2 * n + 1);
            sum = sum + term;
        }
        return 1.0 - two_over_sqrtpi * x * orc_pm_exp(-(x * x)) * sum;
    }
    if (x > 27.3) return 0.0;
    double f = x;
    for (int n = 60; n >= 1; --n) f = x + (0.5 * (double)n) / f;
    return orc_pm_exp(-(x * x)) * inv_sqrtpi / f;
}

#endif
