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
 *   erfc  W. J. Cody's rational Chebyshev approximations (Math. Comp. 23 (1969)
 *         631-637; the CALERF routine of Cody's SPECFUN package, 1990), which is
 *         what CESM's error_function module evaluates for the erfc the scheme calls
 *         (uwshcu.F90:5, 2116, 2145, 4964).  Three intervals of |x|:
 *           |x| <= 0.46875      erfc = 1 - x R1(x^2),
 *           0.46875 < |x| <= 4  erfc = e^{-x^2} R2(|x|),
 *           4 < |x| < 26.543    erfc = e^{-x^2} (1/sqrt(pi) - z R3(z)) / |x|, z = 1/x^2,
 *         each R a ratio of polynomials evaluated by Cody's Horner recurrences, and
 *         e^{-x^2} formed as e^{-s^2} e^{-(|x|-s)(|x|+s)} with s = |x| truncated to a
 *         multiple of 1/16 (both factors' arguments exact or nearly so).  For x < 0
 *         outside the first interval erfc = 2 - erfc(|x|).  orc_cody_erfc takes the
 *         exp to use: the portable flavour passes orc_pm_exp, the libm flavour the
 *         host's exp.
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

/* Cody (1969), Tables for erf on [0, 0.5], erfc on [0.46875, 4] and erfc on
 * [4, inf): numerator and denominator coefficients of R1, R2, R3 in the order the
 * CALERF recurrences consume them. */
static const double orc_cody_a[5] = {3.16112374387056560e00, 1.13864154151050156e02, 3.77485237685302021e02,
                                     3.20937758913846947e03, 1.85777706184603153e-1};
static const double orc_cody_b[4] = {2.36012909523441209e01, 2.44024637934444173e02, 1.28261652607737228e03,
                                     2.84423683343917062e03};
static const double orc_cody_c[9] = {5.64188496988670089e-1, 8.88314979438837594e00, 6.61191906371416295e01,
                                     2.98635138197400131e02, 8.81952221241769090e02, 1.71204761263407058e03,
                                     2.05107837782607147e03, 1.23033935479799725e03, 2.15311535474403846e-8};
static const double orc_cody_d[8] = {1.57449261107098347e01, 1.17693950891312499e02, 5.37181101862009858e02,
                                     1.62138957456669019e03, 3.29079923573345963e03, 4.36261909014324716e03,
                                     3.43936767414372164e03, 1.23033935480374942e03};
static const double orc_cody_p[6] = {3.05326634961232344e-1, 3.60344899949804439e-1, 1.25781726111229246e-1,
                                     1.60837851487422766e-2, 6.58749161529837803e-4, 1.63153871373020978e-2};
static const double orc_cody_q[5] = {2.56852019228982242e00, 1.87295284992346047e00, 5.27905102951428412e-1,
                                     6.05183413124413191e-2, 2.33520497626869185e-3};

static inline double orc_cody_erfc(double x, double (*ex)(double))
{
    if (x != x) return x;
    const double y = fabs(x);
    if (y <= 0.46875) {
        /* erf(x) = x R1(x^2); below 1.11e-16 the x^2 terms vanish */
        const double z = y > 1.11e-16 ? y * y : 0.0;
        double num = orc_cody_a[4] * z, den = z;
        for (int i = 0; i < 3; ++i) {
            num = (num + orc_cody_a[i]) * z;
            den = (den + orc_cody_b[i]) * z;
        }
        return 1.0 - x * (num + orc_cody_a[3]) / (den + orc_cody_b[3]);
    }
    double r;
    if (y <= 4.0) {
        double num = orc_cody_c[8] * y, den = y;
        for (int i = 0; i < 7; ++i) {
            num = (num + orc_cody_c[i]) * y;
            den = (den + orc_cody_d[i]) * y;
        }
        r = (num + orc_cody_c[7]) / (den + orc_cody_d[7]);
    } else if (y >= 26.543) {
        r = 0.0;                         /* below the smallest double: no scaling */
    } else {
        const double z = 1.0 / (y * y);
        double num = orc_cody_p[5] * z, den = z;
        for (int i = 0; i < 4; ++i) {
            num = (num + orc_cody_p[i]) * z;
            den = (den + orc_cody_q[i]) * z;
        }
        r = z * (num + orc_cody_p[4]) / (den + orc_cody_q[4]);
        r = (5.6418958354775628695e-1 - r) / y;
    }
    if (y < 26.543) {
        const double s = floor(y * 16.0) / 16.0;
        const double del = (y - s) * (y + s);
        r = ex(-s * s) * ex(-del) * r;
    }
    return x < 0.0 ? 2.0 - r : r;
}

static inline double orc_pm_erfc(double x) { return orc_cody_erfc(x, orc_pm_exp); }
static inline double orc_libm_erfc(double x) { return orc_cody_erfc(x, exp); }

#endif
