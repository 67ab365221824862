/* divcheck.c -- TEST HARNESS: checks on the host that the quotient from a correctly
 * rounded reciprocal plus one fused correction (div_rcp in uwshcu_math.h),
 *     q0 = a*y,  q = fma(fma(-b, q0, a), y, q0),  y = 1.0/b,
 * equals the IEEE quotient a/b, over random operands and hard cases: divisor and
 * dividend significands next to all-ones, short divisor significands, short dividends
 * and quotients next to representable numbers.  Prints the number of pairs tested and
 * the number that differ.  Usage: divcheck <pairs per thread-block> */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static inline uint64_t next(uint64_t* s) {
    uint64_t x = *s;
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    *s = x;
    return x;
}
static inline double make(uint64_t bits52, int e) {
    const uint64_t u = ((uint64_t)(1023 + e) << 52) | (bits52 & ((1ull << 52) - 1));
    double d;
    memcpy(&d, &u, 8);
    return d;
}
static inline double with_bits(double d, uint64_t set, uint64_t clear) {
    uint64_t u;
    memcpy(&u, &d, 8);
    u = (u | set) & ~clear;
    memcpy(&d, &u, 8);
    return d;
}

int main(int argc, char** argv) {
    const long per = argc > 1 ? atol(argv[1]) : 1000000;
    long long bad = 0, n = 0;
#pragma omp parallel for reduction(+ : bad, n) schedule(static)
    for (int t = 0; t < 64; ++t) {
        uint64_t s = 0x9E3779B97F4A7C15ull * (uint64_t)(t + 1);
        for (long i = 0; i < per; ++i) {
            const uint64_t r1 = next(&s), r2 = next(&s), r3 = next(&s);
            double a = make(r1, (int)(r3 % 600) - 300), b = make(r2, (int)((r3 >> 10) % 600) - 300);
            switch ((r3 >> 20) & 7) {
                case 1: b = with_bits(b, (1ull << 52) - 1, r2 & 0xff); break;
                case 2: a = with_bits(a, (1ull << 52) - 1, r1 & 0xff); break;
                case 3: a = nextafter(make(r1, 0) * b, (r1 & 1) ? INFINITY : -INFINITY); break;
                case 4: b = make(r2 & 0xffff000000000ull, (int)((r3 >> 10) % 600) - 300); break;
                case 5: a = make(r1 & 0xffffffff00000ull, (int)(r3 % 600) - 300); break;
                default: break;
            }
            if (r3 & (1ull << 40)) a = -a;
            if (r3 & (1ull << 41)) b = -b;
            const double y = 1.0 / b;
            const double q0 = a * y;
            const double q = fma(fma(-b, q0, a), y, q0);
            const double e = a / b;
            ++n;
            if (memcmp(&q, &e, 8)) {
                ++bad;
                if (bad < 4) printf("differs: %a / %a: %a, expected %a\n", a, b, q, e);
            }
        }
    }
    printf("%lld %lld\n", n, bad);
    return 0;
}
