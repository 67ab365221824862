// uwshcu_thermo.h -- device thermodynamic primitives of the UW shallow-cumulus scheme.
//
// Each function follows the reference routine cited beside it.  The saturation
// vapour pressure table lives in LDS (250 doubles per workgroup); every lookup is a
// two-point gather from it.
#ifndef UWSHCU_THERMO_H
#define UWSHCU_THERMO_H

#include <hip/hip_runtime.h>
#include "uwshcu_types.h"
#include "uwshcu_math.h"
#include "uwshcu_ad.h"

// Every primitive is callable from host code too: the test suite builds the column
// code for the CPU to check its logic against the oracle before it runs on a GPU.
#define UWSHCU_HD __host__ __device__
#ifndef UWSHCU_OUTLINE
#define UWSHCU_OUTLINE __attribute__((noinline))
#endif

namespace uwshcu {

// Record the first fatal condition (code, column) of a launch.
UWSHCU_HD inline void record_error(int* err, int code, int col) {
    int expected = 0;
    if (__atomic_compare_exchange_n(&err[0], &expected, code, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) err[1] = col;
}

// The FP64 saturation adjustment's internal state at its end, for the adjoint's
// implicit-function derivative: the ice fraction, effective latent heat, the last two
// Newton iterates and the saturation values at the last.
struct SatState {
    bool sat;                       // false: unsaturated, no Newton steps taken
    double tc, nu, leff, temps, tprev, es, qs;
};

// FAST selects the saturation adjustment of the fast path (conden_d below): the same
// equation solved to convergence instead of by the reference's ten fixed steps.
template <class D, bool FAST = false>
struct ThermoT {
    const uwshcu_consts* c;   // kernel-argument copy
    const double* estbl;      // LDS copy of the table (FP64 values)

    UWSHCU_HD inline D exnf(D p) const { return upow(p / c->p00, c->rovcp); }  // uwshcu.F90:50-54

    // wv_saturation.F90:64-80
    UWSHCU_HD inline D estblf(D td) const {
        D e = fmax(fmin(td, c->tmax), c->tmin);
        D d = e - c->tmin;
        int i = (int)val(d);                   // int(e-tmin)+1, 1-based -> i (0-based)
        const double ai = (double)i;           // aint(e-tmin) for e >= tmin
        return (c->tmin + ai - e + 1.0) * estbl[i] - (c->tmin + ai - e) * estbl[i + 1];
    }

    // vqsatd with len=1, es and qs only (wv_saturation.F90:530-546)
    UWSHCU_HD inline void qsat_es_qs(D t, D p, D& es, D& qs) const {
        const double omeps = 1.0 - c->epsqs;
        es = estblf(t);
        qs = c->epsqs * es / (p - omeps * es);
        qs = fmin(1.0, qs);
        if (qs < 0.0) { qs = 1.0; es = p; }
    }

    // vqsatd with len=1 including gam = (L/cp) dqs/dT (wv_saturation.F90:482-609)
    UWSHCU_HD inline void qsat_full(D t, D p, D& es, D& qs, D& gam) const {
        const double omeps = 1.0 - c->epsqs;
        qsat_es_qs(t, p, es, qs);
        D hlatsb, hlatvp, hltalt, desdt;
        if (!c->icephs || c->ttrice == 0.0) {
            hlatvp = c->hlatv - 2369.0 * (t - c->tmelt);
            hlatsb = c->icephs ? c->hlatv + c->hlatf : c->hlatv;
            hltalt = (t < c->tmelt) ? hlatsb : hlatvp;
            desdt = hltalt * es / (c->rgasv * t * t);
        } else {
            const double trinv = 1.0 / c->ttrice;
            const D tc = t - c->tmelt;
            const bool lflg = (tc >= -c->ttrice && tc < 0.0);
            const D weight = fmin(-tc * trinv, 1.0);
            hlatsb = c->hlatv + weight * c->hlatf;
            hlatvp = c->hlatv - 2369.0 * tc;
            hltalt = (t < c->tmelt) ? hlatsb : hlatvp;
            D tterm = 0.0;
            if (lflg) tterm = c->pcf[0] + tc * (c->pcf[1] + tc * (c->pcf[2] + tc * (c->pcf[3] + tc * c->pcf[4])));
            desdt = hltalt * es / (c->rgasv * t * t) + tterm * trinv;
        }
        gam = hltalt * qs * p * desdt / (c->cpws * es * (p - omeps * es));
        if (qs == 1.0) gam = 0.0;
    }

    // Saturation adjustment, uwshcu.F90:4718-4785.  Outputs th, qv, ql, qi; returns id_check.
    UWSHCU_HD __attribute__((always_inline)) int conden(D p, D thl, D qt, D& th, D& qv, D& ql, D& qi) const {
        return conden_x(p, exnf(p), thl, qt, th, qv, ql, qi);
    }
    // The same with the Exner function of p supplied by the caller (ex must equal
    // exnf(p); callers pass the set-A values of the model levels).  The FP64 form is
    // expanded at each call site: an out-of-line call would pass the outputs through
    // scratch memory and reload the constants at every call.  The recording (AD) form
    // stays one out-of-line copy.
    UWSHCU_HD __attribute__((always_inline)) int conden_x(D p, D ex, D thl, D qt, D& th, D& qv, D& ql,
                                                         D& qi) const {
        if constexpr (is_ad<D>::value) return conden_ad(p, ex, thl, qt, th, qv, ql, qi);
        else return conden_d(p, ex, thl, qt, th, qv, ql, qi);
    }
    UWSHCU_HD UWSHCU_OUTLINE int conden_ad(D p, D ex, D thl, D qt, D& th, D& qv, D& ql, D& qi) const {
#ifndef UWSHCU_AD_TAPE_CONDEN
        int id;
        if (conden_ift(p, thl, qt, th, qv, ql, qi, id)) return id;
#endif
        const D tc = thl * ex;
        const D nu = fmax(fmin((268.0 - tc) / 20.0, 1.0), 0.0);
        const D leff = (1.0 - nu) * c->xlv + nu * c->xls;
        D temps = tc, es, qs;
        qsat_es_qs(temps, p, es, qs);
        D rvls = qs;
        if (qs >= qt) {
            qv = qt;
            ql = 0.0;
            qi = 0.0;
            th = tc / ex;
            return 0;
        }
        for (int it = 0; it < 10; ++it) {
            temps = temps + ((tc - temps) * c->cp / leff + qt - rvls) /
                            (c->cp / leff + c->ep2 * leff * rvls / c->r / temps / temps);
            qsat_es_qs(temps, p, es, qs);
            rvls = qs;
        }
        const D qc = fmax(qt - qs, 0.0);
        qv = qt - qc;
        ql = qc * (1.0 - nu);
        qi = nu * qc;
        th = temps / ex;
        return (fabs((temps - (leff / c->cp) * qc) - tc) >= 1.0) ? 1 : 0;
    }

    // conden_x for D = double: the same operations, each quotient computed so that it
    // has the bits of the plain expression's.
    //  - Divisions by leff, r and (twice) temps go through correctly rounded
    //    reciprocals (div_rcp, three dependent operations each).
    //  - The others, and the reciprocals, use div_fast (no range scaling; every
    //    operand is a temperature, pressure, latent heat or mixing ratio).
    //  - The Newton loop keeps the two table entries of the last lookup in registers
    //    and reads the table again only when the index moves, which after the first
    //    steps it rarely does.
    UWSHCU_HD __attribute__((always_inline)) int conden_d(double p, double ex, double thl, double qt, double& th,
                                                         double& qv, double& ql, double& qi,
                                                         SatState* st = nullptr) const {
        const double omeps = 1.0 - c->epsqs;
        int ic = -1;
        double tlo = 0.0, thi = 0.0;
        auto qsat_c = [&](double t, double& es_, double& qs_) {   // qsat_es_qs with the cached entries
            const double e = fmax(fmin(t, c->tmax), c->tmin);
            const int i = (int)(e - c->tmin);
            if (i != ic) {
                ic = i;
                tlo = estbl[i];
                thi = estbl[i + 1];
            }
            const double ai = (double)i;
            es_ = (c->tmin + ai - e + 1.0) * tlo - (c->tmin + ai - e) * thi;
            qs_ = div_fast(c->epsqs * es_, p - omeps * es_);
            qs_ = fmin(1.0, qs_);
            if (qs_ < 0.0) { qs_ = 1.0; es_ = p; }
        };
        const double tc = thl * ex;
        const double nu = fmax(fmin(div_fast(268.0 - tc, 20.0), 1.0), 0.0);
        const double leff = (1.0 - nu) * c->xlv + nu * c->xls;
        double temps = tc, es, qs;
        qsat_c(temps, es, qs);
        double rvls = qs;
        if (qs >= qt) {
            qv = qt;
            ql = 0.0;
            qi = 0.0;
            th = div_fast(tc, ex);
            if (st) st->sat = false;
            return 0;
        }
        const double rleff = div_fast(1.0, leff), rr = div_fast(1.0, c->r);
        const double cpl = div_rcp(c->cp, leff, rleff);
        double tprev = temps;
        if constexpr (FAST) {
            // The fixed point of the reference's iteration is the root of
            //     F(T) = (tc - T) cp/leff + qt - qs(T, p),
            // with qs from the piecewise-linear table.  Newton's method with F's own
            // derivative, -cp/leff - dqs/dT, where dqs/dT = epsqs p (des/dT) / (p - (1-epsqs) es)^2
            // and des/dT the table segment's slope (zero where T is clamped or qs capped),
            // converges quadratically inside a segment; the loop stops once a step moves
            // T by at most 1e-10 of itself (at most ten steps, as the reference).
            for (int it = 0; it < 10; ++it) {
                tprev = temps;
                const double dedt = (temps > c->tmin && temps < c->tmax) ? thi - tlo : 0.0;
                const double den = p - omeps * es;
                const double dqs = (qs < 1.0) ? div_fast(c->epsqs * p * dedt, den * den) : 0.0;
                const double f = div_rcp((tc - temps) * c->cp, leff, rleff) + qt - qs;
                temps = temps + div_fast(f, cpl + dqs);
                qsat_c(temps, es, qs);
                if (fabs(temps - tprev) <= 1.e-10 * temps) break;
            }
            rvls = qs;
        } else {
            for (int it = 0; it < 10; ++it) {
                tprev = temps;
                const double rt = div_fast(1.0, temps);
                const double num = div_rcp((tc - temps) * c->cp, leff, rleff) + qt - rvls;
                const double d3 = div_rcp(div_rcp(div_rcp(c->ep2 * leff * rvls, c->r, rr), temps, rt), temps, rt);
                temps = temps + div_fast(num, cpl + d3);
                qsat_c(temps, es, qs);
                rvls = qs;
            }
        }
        const double qc = fmax(qt - qs, 0.0);
        qv = qt - qc;
        ql = qc * (1.0 - nu);
        qi = nu * qc;
        th = div_fast(temps, ex);
        if (st) *st = SatState{true, tc, nu, leff, temps, tprev, es, qs};
        return (fabs((temps - div_fast(leff, c->cp) * qc) - tc) >= 1.0) ? 1 : 0;
    }

    // Adjoint form of conden (D = AD): the values are the FP64 routine's, computed by
    // the same operations; the derivatives of the saturated branch come from the
    // implicit function theorem at the Newton point T of
    //     F(T) = (tc - T) cp/leff + qt - qs(T, p) = 0,   tc = thl (p/p00)^(R/cp),
    // which the ten Newton steps reach to round-off.  Each output is then one or two
    // tape nodes instead of the ~300 the ten steps would record.  When the last step
    // still moved T by more than 1e-12 relative (the iteration hopping between two
    // segments of the saturation table), returns false and the caller records the
    // ten steps themselves.
    UWSHCU_HD bool conden_ift(const D& p, const D& thl, const D& qt, D& th, D& qv, D& ql, D& qi, int& id) const {
        const double pv = val(p), thlv = val(thl), qtv = val(qt);
        const double ex = exnf_d(pv);
        // the values: the FP64 routine itself (the forward kernels' bits)
        SatState st;
        double thv, qvv, qlv, qiv;
        const int idv = ThermoT<double>{c, estbl}.conden_d(pv, ex, thlv, qtv, thv, qvv, qlv, qiv, &st);
        if (!st.sat) {
            // unsaturated: qv = qt, no condensate, th = (thl ex)/ex
            const D exa = exnf(p);
            const D tca = thl * exa;
            qv = qt;
            ql = 0.0;
            qi = 0.0;
            th = tca / exa;
            id = 0;
            return true;
        }
        const double tc = st.tc, nu = st.nu, leff = st.leff, temps = st.temps, es = st.es, qs = st.qs;
        if (fabs(temps - st.tprev) > 1.e-12 * temps) return false;
        const double qc = fmax(qtv - qs, 0.0);
        id = idv;
        // partial derivatives
        const double ex_p = c->rovcp * ex / pv;
        const double tc_thl = ex, tc_p = thlv * ex_p;
        const double r20 = (268.0 - tc) / 20.0;
        const double nu_tc = (r20 >= 0.0 && r20 <= 1.0) ? -1.0 / 20.0 : 0.0;
        const double leff_tc = (c->xls - c->xlv) * nu_tc;
        const double omeps = 1.0 - c->epsqs, den = pv - omeps * es;
        double qs_es = c->epsqs * pv / (den * den), qs_p = -c->epsqs * es / (den * den);
        if (c->epsqs * es / den > 1.0) qs_es = qs_p = 0.0;          // qs clipped at 1
        const double e = fmax(fmin(temps, c->tmax), c->tmin);
        const int i = (int)(e - c->tmin);
        const double es_t = (temps >= c->tmin && temps < c->tmax) ? estbl[i + 1] - estbl[i] : 0.0;
        const double qs_t = qs_es * es_t;
        const double f_t = -c->cp / leff - qs_t;
        const double f_tc = c->cp / leff - (tc - temps) * c->cp * leff_tc / (leff * leff);
        const double t_thl = -f_tc * tc_thl / f_t, t_qt = -1.0 / f_t, t_p = -(f_tc * tc_p - qs_p) / f_t;
        const double s_thl = qs_t * t_thl, s_qt = qs_t * t_qt, s_p = qs_t * t_p + qs_p;
        const bool wet = qtv - qs >= 0.0;
        const double qc_thl = wet ? -s_thl : 0.0, qc_qt = wet ? 1.0 - s_qt : 0.0, qc_p = wet ? -s_p : 0.0;
        const double nu_thl = nu_tc * tc_thl, nu_p = nu_tc * tc_p;
        auto emit = [&](double v, double d_thl, double d_qt, double d_p) {
            const D part = ad_node(0.0, thl, d_thl, qt, d_qt);
            return ad_node(v, part, 1.0, p, d_p);
        };
        qv = emit(qvv, -qc_thl, 1.0 - qc_qt, -qc_p);
        ql = emit(qlv, qc_thl * (1.0 - nu) - qc * nu_thl, qc_qt * (1.0 - nu), qc_p * (1.0 - nu) - qc * nu_p);
        qi = emit(qiv, nu * qc_thl + qc * nu_thl, nu * qc_qt, nu * qc_p + qc * nu_p);
        th = emit(thv, t_thl / ex, t_qt / ex, t_p / ex - temps * ex_p / (ex * ex));
        return true;
    }
    // FP64 pieces of the above (the D = double forms of exnf and qsat_es_qs)
    UWSHCU_HD double exnf_d(double p) const { return upow(p / c->p00, c->rovcp); }
    UWSHCU_HD void qsat_d(double t, double p, double& es, double& qs) const {
        const double omeps = 1.0 - c->epsqs;
        const double e = fmax(fmin(t, c->tmax), c->tmin);
        const double d = e - c->tmin;
        const int i = (int)d;
        const double ai = (double)i;
        es = (c->tmin + ai - e + 1.0) * estbl[i] - (c->tmin + ai - e) * estbl[i + 1];
        qs = c->epsqs * es / (p - omeps * es);
        qs = fmin(1.0, qs);
        if (qs < 0.0) { qs = 1.0; es = p; }
    }

    // uwshcu.F90:4863-4929
    UWSHCU_HD __attribute__((always_inline)) D qsinvert(D qt, D thl, D psfc) const {
        const double psmin = 100.0 * 100.0, dpsmax = 1.0;
        D es, qs, gam;
        const D Ti = thl * upow(psfc / c->p00, c->rovcp);
        qsat_full(Ti, psfc, es, qs, gam);
        const D rhi = qt / qs;
        if (rhi <= 0.01) return psmin;
        const D TLCL = 55.0 + 1.0 / (1.0 / (Ti - 55.0) - ulog(rhi) / 2840.0);
        const D PiLCL = TLCL / thl;
        D ps = c->p00 * upow(PiLCL, 1.0 / c->rovcp);
        for (int it = 0; it < 10; ++it) {
            const D Pis = upow(ps / c->p00, c->rovcp);
            const D Ts = thl * Pis;
            qsat_full(Ts, ps, es, qs, gam);
            const D err = qt - qs;
            const D nu = fmax(fmin((268.0 - Ts) / 20.0, 1.0), 0.0);
            const D leff = (1.0 - nu) * c->xlv + nu * c->xls;
            const D dlnqsdT = gam * (c->cp / leff) / qs;
            const D dTdPis = thl;
            const D dPisdps = c->rovcp * Pis / ps;
            const D dlnqsdps = -1.0 / (ps - (1.0 - c->ep2) * es);
            const D derrdps = -qs * (dlnqsdT * dTdPis * dPisdps + dlnqsdps);
            const D dps = -err / derrdps;
            ps = ps + dps;
            if (ps < 0.0) return psmin;
            if (fabs(dps) <= dpsmax) return ps;
        }
        return psmin;
    }

    // uwshcu.F90:4704-4715
    UWSHCU_HD inline D single_cin(D pbot, D thv0bot, D ptop, D thv0top,
                                                 D thvubot, D thvutop) const {
        return ((1.0 - thvubot / thv0bot) + (1.0 - thvutop / thv0top)) * (pbot - ptop) /
               (pbot / (c->r * thv0bot * exnf(pbot)) + ptop / (c->r * thv0top * exnf(ptop)));
    }

    // uwshcu.F90:4672-4702
    UWSHCU_HD inline void getbuoy(D pbot, D thv0bot, D ptop, D thv0top,
                                            D thvubot, D thvutop, D& plfc, D& cin) const {
        if (thvubot > thv0bot && thvutop > thv0top) {
            plfc = pbot;
            return;
        }
        const D den = (pbot / (c->r * thv0bot * exnf(pbot)) + ptop / (c->r * thv0top * exnf(ptop)));
        if (thvubot <= thv0bot && thvutop <= thv0top) {
            cin = cin - ((thvubot / thv0bot - 1.0) + (thvutop / thv0top - 1.0)) * (pbot - ptop) / den;
        } else if (thvubot > thv0bot && thvutop <= thv0top) {
            const D frc = (thvutop / thv0top - 1.0) / ((thvutop / thv0top - 1.0) - (thvubot / thv0bot - 1.0));
            cin = cin - (thvutop / thv0top - 1.0) * ((ptop + frc * (pbot - ptop)) - ptop) / den;
        } else {
            const D frc = (thvubot / thv0bot - 1.0) / ((thvubot / thv0bot - 1.0) - (thvutop / thv0top - 1.0));
            plfc = pbot - frc * (pbot - ptop);
            cin = cin - (thvubot / thv0bot - 1.0) * (pbot - plfc) / den;
        }
    }
};
using Thermo = ThermoT<double>;
using ThermoFast = ThermoT<double, true>;

// uwshcu.F90:4787-4829
template <class R>
UWSHCU_HD inline void roots(R a, R b, R c, R& r1, R& r2, int& status) {
    status = 0;
    if (a == R(0)) {
        if (b == R(0)) status = 1;
        else r1 = -c / b;
        r2 = r1;
    } else if (b == R(0)) {
        if (a * c > R(0)) status = 2;
        else r1 = sqrt(-c / a);
        r2 = -r1;
    } else if ((b * b - R(4) * a * c) < R(0)) {
        status = 3;
    } else {
        const R q = R(-0.5) * (b + copysign(R(1), b) * sqrt(b * b - R(4) * a * c));
        r1 = q / a;
        r2 = c / q;
    }
}

// uwshcu.F90:4931-4950
template <class D>
UWSHCU_HD inline __attribute__((always_inline)) D compute_alpha(D del_CIN, D ke) {
    D x0 = 0.0;
    for (int it = 0; it < 10; ++it) {
        const D e = uexp(-x0 * ke * del_CIN);
        x0 = x0 - (e - x0) / (-ke * del_CIN * e - 1.0);
    }
    return x0;
}

// uwshcu.F90:4952-4979 (erfc is the device libm's)
template <class R>
UWSHCU_HD inline __attribute__((always_inline)) R compute_mumin2(R mulcl, R rmaxfrac, R mulow) {
    R x0 = mulow;
    for (int it = 0; it < 10; ++it) {
        const R ex = uexp(-(x0 * x0));
        const R ef = uerfc(x0);
        const R exf = ex / ef;
        const R a = ex / R(2) / rmaxfrac;
        const R b = mulcl * R(2.5066) / R(2);
        const R f = R(0.5) * (exf * exf) - R(0.5) * (a * a) - (b * b);
        const R fs = (R(2) * (exf * exf)) * (exf / sqrt(R(3.141592)) - x0) + (R(0.5) * x0 * (ex * ex)) / (rmaxfrac * rmaxfrac);
        x0 = x0 - f / fs;
    }
    return x0;
}

// uwshcu.F90:4981-5028
template <class R>
UWSHCU_HD inline R compute_ppen(R wtwb, R D, R bogbot, R bogtop, R rho0j, R dpen) {
    const R SB = (bogtop - bogbot) / dpen;
    const R s00 = bogbot / rho0j - D * wtwb;
    R x0;
    if (D * dpen < R(1.e-8)) {
        if (s00 >= R(0)) x0 = dpen;
        else x0 = fmax(R(0), fmin(dpen, R(-0.5) * wtwb / s00));
    } else {
        x0 = (s00 >= R(0)) ? dpen : R(0);
        for (int it = 0; it < 5; ++it) {
            const R e = uexp(R(-2) * D * x0);
            const R f = e * (wtwb - (bogbot - SB / (R(2) * D)) / (D * rho0j)) + (SB * x0 + bogbot - SB / (R(2) * D)) / (D * rho0j);
            R fs = R(-2) * D * e * (wtwb - (bogbot - SB / (R(2) * D)) / (D * rho0j)) + (SB) / (D * rho0j);
            if (fs >= R(0)) fs = fmax(fs, R(1.e-10));
            else fs = fmin(fs, R(-1.e-10));
            x0 = x0 - f / fs;
        }
    }
    return -fmax(R(0), fmin(dpen, x0));
}

// Wet-bulb point of findsp (uwshcu.F90:5246-5450).  A point's sequence of updates
// depends only on its own inputs (converged points are masked, and the chunk-wide
// early exit fires only when no unconverged point remains), so each (column,level)
// is solved independently.  Returns the doit code: 0 not converged, 1 skipped,
// 2 converged, 4 too cold; enerr is the relative enthalpy error of converged points.
template <class D, bool F>
UWSHCU_HD inline int findsp_point(const ThermoT<D, F>& T, D q, D t, D p, D& qsp_out, D& enerr) {
    const uwshcu_consts* c = T.c;
    const double omeps = 1.0 - c->epsqs, trinv = 1.0 / c->ttrice;
    const double a1 = 7.5 * ulog(10.0), rair = 287.04, c3 = rair * a1 / c->cpws;
    const double dttol = 1.e-4, dqtol = 1.e-4, tt0 = 273.15;
    const D tlim = fmin(fmax(t, 173.0), 373.0);
    D es = T.estblf(tlim);
    const D denom = p - omeps * es;
    D qs = c->epsqs * es / denom;
    D tsp, qsp, enin, enout = 1.0;
    int doit = 0;
    if (p > 5.0 * es && qs > 0.0 && qs < 0.5) {
        qs = fmin(c->epsqs * es / denom, 1.0);
        const D tc = tlim - tt0;
        const D weight = fmin(-tc * trinv, 1.0);
        const D hlatsb = c->hlatv + weight * c->hlatf;
        const D hlatvp = c->hlatv - 2369.0 * tc;
        const D hltalt = (tlim < tt0) ? hlatsb : hlatvp;
        enin = c->cpws * tlim + hltalt * q;
        const D tmp = q - qs;
        const D c1 = hltalt * c3;
        const D c2 = (tlim + 36.0) * (tlim + 36.0);
        const D r1b = c2 / (c2 + c1 * qs);
        const D qvd = r1b * tmp;
        tsp = tlim + ((hltalt / c->cpws) * qvd);
        es = T.estblf(tsp);
        qsp = fmin(c->epsqs * es / (p - omeps * es), 1.0);
    } else {
        doit = 1;
        tsp = tlim;
        qsp = q;
        enin = 1.0;
    }
    for (int l = 0; l < 8 && doit == 0; ++l) {
        es = T.estblf(tsp);
        qs = fmin(c->epsqs * es / (p - omeps * es), 1.0);
        const D tc = tsp - tt0;
        const bool lflg = (tc >= -c->ttrice && tc < 0.0);
        const D weight = fmin(-tc * trinv, 1.0);
        const D hlatsb = c->hlatv + weight * c->hlatf;
        const D hlatvp = c->hlatv - 2369.0 * tc;
        const D hltalt = (tsp < tt0) ? hlatsb : hlatvp;
        D tterm = 0.0;
        if (lflg) tterm = c->pcf[0] + tc * (c->pcf[1] + tc * (c->pcf[2] + tc * (c->pcf[3] + tc * c->pcf[4])));
        const D desdt = hltalt * es / (c->rgasv * tsp * tsp) + tterm * trinv;
        const D dqsdt = (c->epsqs + omeps * qs) / (p - omeps * es) * desdt;
        const D g = enin - (c->cpws * tsp + hltalt * qsp);
        const D dgdt = -(c->cpws + hltalt * dqsdt);
        const D t1 = tsp - g / dgdt;
        const D dtr = fabs(t1 - tsp) / t1;
        tsp = fmax(t1, c->tmin);
        es = T.estblf(tsp);
        const D q1 = fmin(c->epsqs * es / (p - omeps * es), 1.0);
        const D dq = fabs(q1 - qsp) / fmax(q1, 1.e-12);
        qsp = q1;
        if (dtr < dttol && dq < dqtol) doit = 2;
        enout = c->cpws * tsp + hltalt * qsp;
        if (tsp < 174.16) doit = 4;
    }
    qsp_out = qsp;
    enerr = (doit == 2) ? fabs((enin - enout) / (enin + enout)) : 0.0;
    return doit;
}

}  // namespace uwshcu
#endif
