// uwshcu_math.h -- FP64 exp, log, pow, erfc from IEEE basic operations only.
//
// The column physics branches on values that a one-ulp change in a transcendental
// function can push across a threshold (for example the choice between the two
// critical mixing fractions, uwshcu.F90:2543-2547, or the Newton guard in
// compute_ppen, uwshcu.F90:5015-5019), so two libms that differ in the last bit
// give outputs that differ by up to 1e-3 in a few columns out of 10^4.  The FP64
// path therefore evaluates its transcendental functions with the sequences below,
// which use only correctly rounded +, -, *, /, floor, frexp and ldexp and are
// compiled without fused multiply-add: a CPU build and this GPU build give
// identical bits.  Accuracy: exp <= 1 ulp, log and pow <= 2 ulp, erfc <= 4 ulp.
//
// Methods:
//   exp   Cody-Waite reduction x = k ln2 + r (ln2 split so that k*ln2_hi is exact),
//         degree-14 Taylor series of e^r - 1 on |r| <= 0.347, exact scaling by 2^k.
//   log   x = m 2^e, m in [sqrt(1/2), sqrt(2)); log m = 2 atanh((m-1)/(m+1)) by its
//         odd series to degree 23.
//   pow   exp(y log x) for x > 0.
//   erfc  W. J. Cody's rational Chebyshev approximations (Math. Comp. 23 (1969)
//         631-637; CALERF of Cody's SPECFUN package), the algorithm of CESM's
//         error_function module that the scheme calls (uwshcu.F90:5, 2116, 2145,
//         4964): |x| <= 0.46875: 1 - x R1(x^2); 0.46875 < |x| <= 4: e^{-x^2} R2(|x|);
//         4 < |x| < 26.543: e^{-x^2} (1/sqrt(pi) - z R3(z)) / |x| with z = 1/x^2;
//         beyond, 0.  e^{-x^2} = e^{-s^2} e^{-(|x|-s)(|x|+s)}, s = |x| truncated to a
//         multiple of 1/16.  Negative x outside the first interval: 2 - erfc(|x|).
#ifndef UWSHCU_MATH_H
#define UWSHCU_MATH_H

#include <hip/hip_runtime.h>

#ifndef UWSHCU_HD
#define UWSHCU_HD __host__ __device__
#endif

namespace uwshcu {

UWSHCU_HD inline double pm_exp(double x) {
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    const double inv_ln2 = 1.44269504088896338700e+00;
    if (x != x) return x;
    if (x > 709.782712893384) return HUGE_VAL;
    if (x < -745.1332191019412) return 0.0;
    const double kd = floor(x * inv_ln2 + 0.5);
    const double r = (x - kd * ln2_hi) - kd * ln2_lo;
    double p = 1.0 / 87178291200.0;
    p = p * r + 1.0 / 6227020800.0;
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

UWSHCU_HD inline double pm_log(double x) {
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

UWSHCU_HD inline double pm_pow(double x, double y) {
    if (x == 0.0) return y > 0.0 ? 0.0 : (y == 0.0 ? 1.0 : HUGE_VAL);
    if (y == 0.0) return 1.0;
    return pm_exp(y * pm_log(x));
}

// Coefficients of Cody's R1 (erf, first interval), R2 and R3 (erfc, second and third
// intervals): numerators a, c, p and denominators b, d, q, in the order his Horner
// recurrences consume them.
struct CodyErfc {
    static constexpr double a[5] = {3.16112374387056560e00, 1.13864154151050156e02, 3.77485237685302021e02,
                                    3.20937758913846947e03, 1.85777706184603153e-1};
    static constexpr double b[4] = {2.36012909523441209e01, 2.44024637934444173e02, 1.28261652607737228e03,
                                    2.84423683343917062e03};
    static constexpr double c[9] = {5.64188496988670089e-1, 8.88314979438837594e00, 6.61191906371416295e01,
                                    2.98635138197400131e02, 8.81952221241769090e02, 1.71204761263407058e03,
                                    2.05107837782607147e03, 1.23033935479799725e03, 2.15311535474403846e-8};
    static constexpr double d[8] = {1.57449261107098347e01, 1.17693950891312499e02, 5.37181101862009858e02,
                                    1.62138957456669019e03, 3.29079923573345963e03, 4.36261909014324716e03,
                                    3.43936767414372164e03, 1.23033935480374942e03};
    static constexpr double p[6] = {3.05326634961232344e-1, 3.60344899949804439e-1, 1.25781726111229246e-1,
                                    1.60837851487422766e-2, 6.58749161529837803e-4, 1.63153871373020978e-2};
    static constexpr double q[5] = {2.56852019228982242e00, 1.87295284992346047e00, 5.27905102951428412e-1,
                                    6.05183413124413191e-2, 2.33520497626869185e-3};
};

UWSHCU_HD inline double pm_erfc(double x) {
    using K = CodyErfc;
    if (x != x) return x;
    const double y = fabs(x);
    if (y <= 0.46875) {
        const double z = y > 1.11e-16 ? y * y : 0.0;
        double num = K::a[4] * z, den = z;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            num = (num + K::a[i]) * z;
            den = (den + K::b[i]) * z;
        }
        return 1.0 - x * (num + K::a[3]) / (den + K::b[3]);
    }
    double r;
    if (y <= 4.0) {
        double num = K::c[8] * y, den = y;
#pragma unroll
        for (int i = 0; i < 7; ++i) {
            num = (num + K::c[i]) * y;
            den = (den + K::d[i]) * y;
        }
        r = (num + K::c[7]) / (den + K::d[7]);
    } else if (y >= 26.543) {
        r = 0.0;
    } else {
        const double z = 1.0 / (y * y);
        double num = K::p[5] * z, den = z;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            num = (num + K::p[i]) * z;
            den = (den + K::q[i]) * z;
        }
        r = z * (num + K::p[4]) / (den + K::q[4]);
        r = (5.6418958354775628695e-1 - r) / y;
    }
    if (y < 26.543) {
        const double s = floor(y * 16.0) / 16.0;
        const double del = (y - s) * (y + s);
        r = pm_exp(-s * s) * pm_exp(-del) * r;
    }
    return x < 0.0 ? 2.0 - r : r;
}

// Overloads used by the templated column code: the FP64 path gets the portable
// functions above, the FP32 bulk of the mixed path the device single-precision libm.
UWSHCU_HD inline double uexp(double x) { return pm_exp(x); }
UWSHCU_HD inline double ulog(double x) { return pm_log(x); }
UWSHCU_HD inline double upow(double x, double y) { return pm_pow(x, y); }
UWSHCU_HD inline double uerfc(double x) { return pm_erfc(x); }
UWSHCU_HD inline float uexp(float x) { return expf(x); }
UWSHCU_HD inline float ulog(float x) { return logf(x); }
UWSHCU_HD inline float upow(float x, float y) { return powf(x, y); }
UWSHCU_HD inline float uerfc(float x) { return erfcf(x); }

// a / b, correctly rounded, from y = 1.0 / b (itself correctly rounded): q0 = a*y is
// within one ulp of a/b, the remainder a - b*q0 is exact in one fused multiply-add,
// and q0 + remainder*y rounds to a/b (Markstein, IBM J. Res. Dev. 34(1), 1990, the
// theorem behind fused-multiply-add division).  So a divisor that recurs pays for one
// division, and each quotient costs three dependent operations instead of a full
// division sequence, with identical bits.  Valid for finite a, finite nonzero b and
// |a| above 2^-900 or zero (no underflow in the remainder); a = -0 returns +0.
UWSHCU_HD inline double div_rcp(double a, double b, double y) {
    const double q0 = a * y;
    return fma(fma(-b, q0, a), y, q0);
}

// a / b on the GPU by the operations of the hardware's IEEE FP64 division without its
// range scaling (v_div_scale) and special-value fix-up (v_div_fixup): reciprocal
// estimate, two Newton steps, quotient, exact remainder and the final fused
// correction.  When neither operand nor the quotient is near the ends of the
// exponent range (|exponents| below 900, no zero, infinite or NaN divisor) the
// scaling steps are identities and the fix-up returns its input, so the result is the
// correctly rounded quotient, the bits of `a / b`; two dependent steps shorter.  Host
// builds use the operator itself.
UWSHCU_HD inline double div_fast(double a, double b) {
#if defined(__HIP_DEVICE_COMPILE__)
    double y = __builtin_amdgcn_rcp(b);
    double e = fma(-b, y, 1.0);
    y = fma(y, e, y);
    e = fma(-b, y, 1.0);
    y = fma(y, e, y);
    const double q = a * y;
    return fma(fma(-b, q, a), y, q);
#else
    return a / b;
#endif
}

}  // namespace uwshcu
#endif
