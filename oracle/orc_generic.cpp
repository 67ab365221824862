/*
 * orc_generic.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * The oracle (uwshcu_oracle.c, portable-math flavour) compiled as C++ with its FP64
 * type replaced by a number class, so that the same restatement, with the same
 * operations in the same order and therefore the same values and branch decisions,
 * also yields:
 *
 *   -DORC_COUNT  (liboracle_count.so)  the number of FP64 operations each column
 *                performs: add/subtract, multiply, divide, square root, comparisons,
 *                and the floor/frexp/ldexp steps of the portable exp/log/pow/erfc
 *                (which are thereby counted as the basic operations they consist of);
 *   -DORC_DUAL   (liboracle_dual.so)   forward-mode derivatives: every value carries a
 *                tangent (a dual number v + d eps, eps^2 = 0), so a run seeded with a
 *                direction in the inputs and the tunables returns the outputs' tangents,
 *                the Jacobian-vector product of the scheme along that direction, exact up
 *                to rounding, with the branch pattern of the FP64 run.
 *
 * Neither is used by the product: the counts give the algorithmic work per column that
 * the benchmark divides by kernel time, and the tangents check the adjoint.
 */
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#if defined(ORC_COUNT) == defined(ORC_DUAL)
#error "build with exactly one of -DORC_COUNT or -DORC_DUAL"
#endif

#ifdef ORC_COUNT
// Operation counters of the calling thread, by kind.
enum { OC_ADD, OC_MUL, OC_DIV, OC_SQRT, OC_CMP, OC_OTHER, OC_N };
static thread_local unsigned long long g_ops[OC_N];
#define OC(k) (++g_ops[(k)])
#else
#define OC(k) ((void)0)
#endif

struct Num {
    double v;
#ifdef ORC_DUAL
    double d;
    Num(double a = 0.0) : v(a), d(0.0) {}
    Num(double a, double b) : v(a), d(b) {}
#else
    Num(double a = 0.0) : v(a) {}
#endif
    Num(int a) : Num((double)a) {}
    Num(long a) : Num((double)a) {}
    Num(unsigned long a) : Num((double)a) {}
    explicit operator int() const { return (int)v; }
    explicit operator long() const { return (long)v; }
    explicit operator double() const { return v; }
};
#ifdef ORC_DUAL
#define MK(val, der) Num((val), (der))
#define DER(x) (x).d
#else
#define MK(val, der) Num(val)
#define DER(x) 0.0
#endif

inline Num operator+(Num a, Num b) { OC(OC_ADD); return MK(a.v + b.v, DER(a) + DER(b)); }
inline Num operator-(Num a, Num b) { OC(OC_ADD); return MK(a.v - b.v, DER(a) - DER(b)); }
inline Num operator-(Num a) { return MK(-a.v, -DER(a)); }
inline Num operator+(Num a) { return a; }
inline Num operator*(Num a, Num b) { OC(OC_MUL); return MK(a.v * b.v, DER(a) * b.v + a.v * DER(b)); }
inline Num operator/(Num a, Num b) {
    OC(OC_DIV);
    const double q = a.v / b.v;
    return MK(q, (DER(a) - q * DER(b)) / b.v);
}
inline Num& operator+=(Num& a, Num b) { return a = a + b; }
inline Num& operator-=(Num& a, Num b) { return a = a - b; }
inline Num& operator*=(Num& a, Num b) { return a = a * b; }
inline Num& operator/=(Num& a, Num b) { return a = a / b; }
inline bool operator<(Num a, Num b) { OC(OC_CMP); return a.v < b.v; }
inline bool operator>(Num a, Num b) { OC(OC_CMP); return a.v > b.v; }
inline bool operator<=(Num a, Num b) { OC(OC_CMP); return a.v <= b.v; }
inline bool operator>=(Num a, Num b) { OC(OC_CMP); return a.v >= b.v; }
inline bool operator==(Num a, Num b) { OC(OC_CMP); return a.v == b.v; }
inline bool operator!=(Num a, Num b) { OC(OC_CMP); return a.v != b.v; }

inline Num fabs(Num a) { return a.v < 0.0 ? -a : a; }
inline Num sqrt(Num a) {
    OC(OC_SQRT);
    const double s = std::sqrt(a.v);
    return MK(s, s > 0.0 ? DER(a) / (2.0 * s) : 0.0);
}
inline Num fmin(Num a, Num b) { OC(OC_CMP); return (b.v < a.v) ? b : a; }
inline Num fmax(Num a, Num b) { OC(OC_CMP); return (b.v > a.v) ? b : a; }
inline Num copysign(Num a, Num b) { return (std::signbit(a.v) != std::signbit(b.v)) ? -a : a; }
inline Num floor(Num a) { OC(OC_OTHER); return MK(std::floor(a.v), 0.0); }
inline Num trunc(Num a) { OC(OC_OTHER); return MK(std::trunc(a.v), 0.0); }
inline Num frexp(Num a, int* e) {
    OC(OC_OTHER);
    const double m = std::frexp(a.v, e);
    return MK(m, DER(a) * std::ldexp(1.0, -*e));
}
inline Num ldexp(Num a, int e) { OC(OC_OTHER); return MK(std::ldexp(a.v, e), std::ldexp(DER(a), e)); }
// Only the Goff-Gratch table builder (gffgch, once per run) calls these; the column
// code's transcendental functions are the portable ones of orc_math.h.
inline Num log10(Num a) { return MK(std::log10(a.v), DER(a) / (a.v * std::log(10.0))); }
inline Num log(Num a) { return MK(std::log(a.v), DER(a) / a.v); }
inline Num exp(Num a) { const double e = std::exp(a.v); return MK(e, e * DER(a)); }
inline Num pow(Num a, Num b) {
    const double p = std::pow(a.v, b.v);
    return MK(p, p * (DER(b) * std::log(a.v) + b.v * DER(a) / a.v));
}

#define ORC_PORTABLE_MATH 1
#define double Num
#include "uwshcu_oracle.c"
#undef double

/* ------------------------------------------------------------------------- */
#ifdef ORC_COUNT
extern "C" {
// Operation counts of one column: ops[0..5] = add/sub, mul, div, sqrt, compare,
// floor/frexp/ldexp.  Runs orc_compute_uwshcu_inv on column `col` of a chunk whose
// arrays have leading dimension mix (the oracle's own layout) and returns its status.
int orc_count_column(const orc_consts* c, int mix, int mkx, int ncnst, double_t dt, int col, const double_t* const* in,
                     double_t* const* out, int* diag, unsigned long long* ops);
}

int orc_count_column(const orc_consts* c, int mix, int mkx, int ncnst, double_t dt, int col, const double_t* const* in,
                     double_t* const* out, int* diag, unsigned long long* ops) {
    // in[]: ps0 zs0 p0 z0 dp0 u0 v0 qv0 ql0 qi0 t0 s0 tr0 tke cldfrc concldfrc pblh cush dpdry0
    // out[]: umf slflx qtflx flxprc1 flxsnow1 qvten qlten qiten sten uten vten trten qrten qsten
    //        precip snow evapc cufrc qcu qlu qiu cbmf qc rliq cnt cnb wu
    // Num has the layout of a double here, so the caller's arrays are used in place.
    auto I = [&](int k) { return reinterpret_cast<const Num*>(in[k]) + col; };
    auto O = [&](int k) { return reinterpret_cast<Num*>(out[k]) + col; };
    std::memset(g_ops, 0, sizeof(g_ops));
    int bad = -1;
    const int rc = orc_compute_uwshcu_inv(c, mix, mkx, 1, ncnst, Num(dt), I(0), I(1), I(2), I(3), I(4), I(5), I(6),
        I(7), I(8), I(9), I(10), I(11), I(12), I(13), I(14), I(15), I(16), const_cast<Num*>(I(17)), O(0), O(1),
        O(2), O(3), O(4), O(5), O(6), O(7), O(8), O(9), O(10), O(11), O(12), O(13), O(14), O(15), O(16), O(17),
        O(18), O(19), O(20), O(21), O(22), O(23), O(24), O(25), I(18), O(26), diag, &bad);
    for (int k = 0; k < OC_N; ++k) ops[k] = g_ops[k];
    return rc;
}
static_assert(sizeof(Num) == sizeof(double_t), "the counting number must have the layout of a double");
#endif

/* ------------------------------------------------------------------------- */
#ifdef ORC_DUAL
#include <vector>
extern "C" {
// Jacobian-vector product of one column.  in[k] / tin[k] (tin[k] nullable: zero
// tangent) follow orc_count_column's input order and the oracle's layout with leading
// dimension mix; par[3] = rpen, rle, rkm with tangents dpar[3].  Writes the outputs'
// values to out[k] and their tangents to tout[k] (same layout, output order of
// orc_count_column) for column `col`, and its diagnostics to diag.  Returns the status.
int orc_dual_jvp(int mix, int mkx, int ncnst, double_t dt, int col, const double_t* par, const double_t* dpar,
                 const double_t* const* in, const double_t* const* tin, double_t* const* out,
                 double_t* const* tout, int* diag);
}

int orc_dual_jvp(int mix, int mkx, int ncnst, double_t dt, int col, const double_t* par, const double_t* dpar,
                 const double_t* const* in, const double_t* const* tin, double_t* const* out,
                 double_t* const* tout, int* diag) {
    enum { NIN = 19, NOUT = 27 };
    // rows of each field: interfaces mkx+1, layers mkx, constituents mkx*ncnst, scalars 1
    const int kif = mkx + 1;
    const int rin[NIN] = {kif, kif, mkx, mkx, mkx, mkx, mkx, mkx, mkx, mkx, mkx, mkx, mkx * ncnst, kif, mkx, mkx,
                          1, 1, mkx};
    const int rout[NOUT] = {kif, kif, kif, kif, kif, mkx, mkx, mkx, mkx, mkx, mkx, mkx * ncnst, mkx, mkx, 1, 1,
                            mkx, mkx, mkx, mkx, mkx, 1, mkx, 1, 1, 1, kif};
    std::vector<std::vector<Num>> vi(NIN), vo(NOUT);
    for (int k = 0; k < NIN; ++k) {
        vi[k].resize(rin[k] > 0 ? rin[k] : 1);
        for (int r = 0; r < rin[k]; ++r)
            vi[k][r] = Num(in[k][(size_t)r * mix + col], tin[k] ? tin[k][(size_t)r * mix + col] : 0.0);
    }
    for (int k = 0; k < NOUT; ++k) vo[k].assign(rout[k] > 0 ? rout[k] : 1, Num(0.0));
    orc_consts c;
    orc_default_consts(&c);
    c.rpen = Num(par[0], dpar[0]);
    c.rle = Num(par[1], dpar[1]);
    c.rkm = Num(par[2], dpar[2]);
    auto I = [&](int k) { return vi[k].data(); };
    auto O = [&](int k) { return vo[k].data(); };
    int bad = -1;
    const int rc = orc_compute_uwshcu_inv(&c, 1, mkx, 1, ncnst, Num(dt), I(0), I(1), I(2), I(3), I(4), I(5), I(6),
        I(7), I(8), I(9), I(10), I(11), I(12), I(13), I(14), I(15), I(16), I(17), O(0), O(1), O(2), O(3), O(4), O(5),
        O(6), O(7), O(8), O(9), O(10), O(11), O(12), O(13), O(14), O(15), O(16), O(17), O(18), O(19), O(20), O(21),
        O(22), O(23), O(24), O(25), I(18), O(26), diag, &bad);
    for (int k = 0; k < NOUT; ++k)
        for (int r = 0; r < rout[k]; ++r) {
            out[k][(size_t)r * mix + col] = vo[k][r].v;
            tout[k][(size_t)r * mix + col] = vo[k][r].d;
        }
    return rc;
}
#endif
