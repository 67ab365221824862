/*
 * uwshcu_oracle.c -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * Serial FP64 restatement of the UW shallow-cumulus scheme as the reference
 * Fortran computes it:
 *   compute_uwshcu_inv            uwshcu.F90:251-476
 *   compute_uwshcu (column loop)  uwshcu.F90:478-4664
 *   compute_zwshcu_inv / compute_zwshcu, the modified scheme:
 *                                 zwshcu.F90:336-628, 630-4969 (its differences
 *                                 from compute_uwshcu are marked where they apply)
 *   helpers                       uwshcu.F90:4672-5454
 *   estblf/gestbl/vqsatd/fqsatd   wv_saturation.F90:64-204,482-609,734-751
 *   gffgch                        gffgch.F90:1-127, parameters esinti.F90:39-51
 *
 * Arrays keep the reference's index ranges so each statement can be read
 * against its Fortran line: interface arrays are indexed 0..mkx, layer arrays
 * 1..mkx (slot 0 unused), constituents 1..ntr.  Expressions keep the
 * reference's operand order.  Built with -ffp-contract=off so that no
 * multiply-add is fused.
 *
 * Number-constituent semantics: compute_uwshcu reads NUMLIQ/NUMICE
 * (uwshcu.F90:1076-1077).  When the caller passes fewer constituents than
 * those indices, the missing constituents are treated as present and zero
 * (padded), and only trten(:,:,1:ncnst) is returned.
 *
 * Pinning status of each function is stated in tests/test_oracle_*.py and in
 * the repository design notes.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "uwshcu_oracle.h"

/* Transcendental functions of the column physics: the host libm by default, or
 * the portable FP64 versions of orc_math.h (-DORC_PORTABLE_MATH) that the HIP
 * kernel reproduces bit for bit.  erfc is Cody's rational approximation in both
 * flavours (the algorithm of CESM's error_function module), evaluated with the
 * flavour's exp.  The saturation table builder always uses the host libm (the
 * table is data shared by both sides). */
#include "orc_math.h"
#ifdef ORC_PORTABLE_MATH
#define XEXP orc_pm_exp
#define XLOG orc_pm_log
#define XPOW orc_pm_pow
#define XERFC orc_pm_erfc
#else
#define XEXP exp
#define XLOG log
#define XPOW pow
#define XERFC orc_libm_erfc
#endif

#define KMAX 200            /* mkx must be < KMAX */
#define NT   (ORC_NCNST_MAX + 1)

static inline double dmin(double a, double b) { return a < b ? a : b; }
static inline double dmax(double a, double b) { return a > b ? a : b; }

/* Stage probe (test infrastructure): when a buffer is set, the column code records
 * the internal quantities of the listed stages there (layout in uwshcu_oracle.h),
 * each overwritten by later passes, so that after a one-column call the record holds
 * the values of the pass whose results the column kept last. */
#ifdef __cplusplus
#define ORC_TLS thread_local   /* orc_generic.cpp compiles this file as C++ */
#else
#define ORC_TLS _Thread_local
#endif
static ORC_TLS double *g_prb = NULL;
void orc_set_probe(double *buf) { g_prb = buf; }
int orc_probe_layout(int *nscal, int *kmax, int *nfield)
{
    *nscal = ORC_P_NSCAL;
    *kmax = ORC_PRB_KMAX;
    *nfield = ORC_PL_NFIELD;
    return ORC_P_NSCAL + ORC_PRB_KMAX * ORC_PL_NFIELD;
}
#define PRB(i) g_prb[(i)]
#define PRBL(f, k) g_prb[ORC_P_NSCAL + (size_t)(f) * ORC_PRB_KMAX + (size_t)(k)]
#define PROBE(stmt) do { if (g_prb) { stmt; } } while (0)

/* ------------------------------------------------------------------------- */
/* Constants                                                                  */
/* ------------------------------------------------------------------------- */

void orc_default_consts(orc_consts *c)
{
    /* CESM shr_const_mod values, combined in the same order as shr_const_mod. */
    const double boltz = 1.38065e-23, avogad = 6.02214e26;
    const double rgas = avogad * boltz;
    const double mwdair = 28.966, mwwv = 18.016;
    const double rdair = rgas / mwdair, rwv = rgas / mwwv;
    memset(c, 0, sizeof(*c));
    c->xlv = 2.501e6;
    c->xlf = 3.337e5;
    c->xls = c->xlv + c->xlf;               /* uwshcu.F90:232 */
    c->cp = 1.00464e3;
    c->zvir = rwv / rdair - 1.0;
    c->r = rdair;
    c->g = 9.80616;
    c->ep2 = mwwv / mwdair;
    c->p00 = 1.e5;                          /* uwshcu.F90:238 */
    c->rovcp = c->r / c->cp;                /* uwshcu.F90:239 */
    c->rpen = 10.0;
    c->rle = 0.1;                           /* uwshcu.F90:1036 */
    c->rkm = 14.0;                          /* uwshcu.F90:1038 */
    c->epsqs = mwwv / mwdair;
    c->rgasv = rwv;
    c->hlatv = 2.501e6;
    c->hlatf = 3.337e5;
    c->cpws = 1.00464e3;
    c->tmelt = 273.15;
    /* constituents: Q (qmin 1e-12, initindx.F90:75), CLDLIQ, CLDICE, NUMLIQ,
     * NUMICE (qmin 0, microp_driver.F90:112-120); all 'wet'. */
    for (int m = 0; m < NT; ++m) { c->qmin[m] = 0.0; c->wet[m] = 1; }
    c->qmin[1] = 1.e-12;
    c->ixnumliq = 4;
    c->ixnumice = 5;
    /* esinti.F90:44-47 */
    orc_gestbl(c, 173.16, 375.16, 20.0, 1);
}

/* gffgch.F90:1-127 (Goff & Gratch 1946), itype < 0: -itype degree transition */
void orc_gffgch(const orc_consts *c, double t, double *es, int itype)
{
    double tr, e1, e2, f, f1, f2, f3, f4, f5, ps, ts, eswtr = 0.0, t0, term1, term2, term3, weight;
    if (itype < 0) { tr = fabs((double)itype); itype = 1; }
    else tr = 0.0;
    if (t < (c->tmelt - tr) && itype == 1) goto ice;
    ps = 1013.246;
    ts = 373.16;
    e1 = 11.344 * (1.0 - t / ts);
    e2 = -3.49149 * (ts / t - 1.0);
    f1 = -7.90298 * (ts / t - 1.0);
    f2 = 5.02808 * log10(ts / t);
    f3 = -1.3816 * (pow(10.0, e1) - 1.0) / 10000000.0;
    f4 = 8.1328 * (pow(10.0, e2) - 1.0) / 1000.0;
    f5 = log10(ps);
    f = f1 + f2 + f3 + f4 + f5;
    *es = pow(10.0, f) * 100.0;
    eswtr = *es;
    if (t >= c->tmelt || itype == 0) return;
ice:
    t0 = c->tmelt;
    term1 = 2.01889049 / (t0 / t);
    term2 = 3.56654 * log(t0 / t);
    term3 = 20.947031 * (t0 / t);
    *es = 575.185606e10 * exp(-(term1 + term2 + term3));
    if (t < (c->tmelt - tr)) return;
    weight = dmin((c->tmelt - t) / tr, 1.0);
    *es = weight * *es + (1.0 - weight) * eswtr;
}

/* wv_saturation.F90:82-204 (gestbl) */
void orc_gestbl(orc_consts *c, double tmn, double tmx, double trice, int ip)
{
    int itype, n, lentbl;
    double t;
    c->tmin = tmn;
    c->tmax = tmx;
    c->ttrice = trice;
    c->icephs = ip;
    lentbl = (int)(c->tmax - c->tmin + 2.000001);
    if (c->icephs) itype = (c->ttrice != 0.0) ? (int)(-c->ttrice) : 1;
    else itype = 0;
    t = c->tmin - 1.0;
    for (n = 1; n <= lentbl && n <= 250; ++n) {
        t = t + 1.0;
        orc_gffgch(c, t, &c->estbl[n - 1], itype);
    }
    for (n = lentbl + 1; n <= 250; ++n) c->estbl[n - 1] = -99999.0;
    c->pcf[0] = 5.04469588506e-01;
    c->pcf[1] = -5.47288442819e+00;
    c->pcf[2] = -3.67471858735e-01;
    c->pcf[3] = -8.95963532403e-03;
    c->pcf[4] = -7.78053686625e-05;
}

/* wv_saturation.F90:64-80 */
double orc_estblf(const orc_consts *c, double td)
{
    double e = dmax(dmin(td, c->tmax), c->tmin);
    int i = (int)(e - c->tmin) + 1;
    double ai = trunc(e - c->tmin);
    return (c->tmin + ai - e + 1.0) * c->estbl[i - 1] - (c->tmin + ai - e) * c->estbl[i];
}

/* fqsatd -> vqsatd with len=1, wv_saturation.F90:482-609,734-751 */
void orc_qsat(const orc_consts *c, double t, double p, double *es_out, double *qs_out, double *gam_out)
{
    double omeps = 1.0 - c->epsqs;
    double es = orc_estblf(c, t);
    double qs = c->epsqs * es / (p - omeps * es);
    double trinv, tc, weight, hlatsb, hlatvp, hltalt, tterm, desdt, gam;
    int lflg;
    qs = dmin(1.0, qs);
    if (qs < 0.0) { qs = 1.0; es = p; }
    if (!c->icephs || c->ttrice == 0.0) {
        hlatvp = c->hlatv - 2369.0 * (t - c->tmelt);
        hlatsb = c->icephs ? c->hlatv + c->hlatf : c->hlatv;
        hltalt = (t < c->tmelt) ? hlatsb : hlatvp;
        desdt = hltalt * es / (c->rgasv * t * t);
        gam = hltalt * qs * p * desdt / (c->cpws * es * (p - omeps * es));
        if (qs == 1.0) gam = 0.0;
    } else {
        trinv = 1.0 / c->ttrice;
        tc = t - c->tmelt;
        lflg = (tc >= -c->ttrice && tc < 0.0);
        weight = dmin(-tc * trinv, 1.0);
        hlatsb = c->hlatv + weight * c->hlatf;
        hlatvp = c->hlatv - 2369.0 * tc;
        hltalt = (t < c->tmelt) ? hlatsb : hlatvp;
        if (lflg) tterm = c->pcf[0] + tc * (c->pcf[1] + tc * (c->pcf[2] + tc * (c->pcf[3] + tc * c->pcf[4])));
        else tterm = 0.0;
        desdt = hltalt * es / (c->rgasv * t * t) + tterm * trinv;
        gam = hltalt * qs * p * desdt / (c->cpws * es * (p - omeps * es));
        if (qs == 1.0) gam = 0.0;
    }
    *es_out = es;
    *qs_out = qs;
    *gam_out = gam;
}

/* uwshcu.F90:50-54 */
static inline double exnf(const orc_consts *c, double pressure) { return XPOW(pressure / c->p00, c->rovcp); }

/* uwshcu.F90:4718-4785 */
void orc_conden(const orc_consts *c, double p, double thl, double qt,
                double *th, double *qv, double *ql, double *qi, double *rvls, int *id_check)
{
    double tc, temps, leff, nu, qc, es, qs, gam;
    int it;
    tc = thl * exnf(c, p);
    nu = dmax(dmin((268.0 - tc) / 20.0, 1.0), 0.0);
    leff = (1.0 - nu) * c->xlv + nu * c->xls;
    temps = tc;
    orc_qsat(c, temps, p, &es, &qs, &gam);
    *rvls = qs;
    if (qs >= qt) {
        *id_check = 0;
        *qv = qt;
        qc = 0.0;
        *ql = 0.0;
        *qi = 0.0;
        *th = tc / exnf(c, p);
    } else {
        for (it = 1; it <= 10; ++it) {
            temps = temps + ((tc - temps) * c->cp / leff + qt - *rvls) /
                            (c->cp / leff + c->ep2 * leff * *rvls / c->r / temps / temps);
            orc_qsat(c, temps, p, &es, &qs, &gam);
            *rvls = qs;
        }
        qc = dmax(qt - qs, 0.0);
        *qv = qt - qc;
        *ql = qc * (1.0 - nu);
        *qi = nu * qc;
        *th = temps / exnf(c, p);
        *id_check = (fabs((temps - (leff / c->cp) * qc) - tc) >= 1.0) ? 1 : 0;
    }
}

/* uwshcu.F90:4787-4829 */
void orc_roots(double a, double b, double c, double *r1, double *r2, int *status)
{
    double q;
    *status = 0;
    if (a == 0.0) {
        if (b == 0.0) *status = 1;
        else *r1 = -c / b;
        *r2 = *r1;
    } else {
        if (b == 0.0) {
            if (a * c > 0.0) *status = 2;
            else *r1 = sqrt(-c / a);
            *r2 = -*r1;
        } else {
            if ((b * b - 4.0 * a * c) < 0.0) *status = 3;
            else {
                q = -0.5 * (b + copysign(1.0, b) * sqrt(b * b - 4.0 * a * c));
                *r1 = q / a;
                *r2 = c / q;
            }
        }
    }
}

/* uwshcu.F90:4831-4861; field/p0/out indexed 1..mkx */
void orc_slope(int mkx, const double *field, const double *p0, double *out)
{
    double below = (field[2] - field[1]) / (p0[2] - p0[1]), above;
    for (int k = 2; k <= mkx; ++k) {
        above = (field[k] - field[k - 1]) / (p0[k] - p0[k - 1]);
        if (above > 0.0) out[k - 1] = dmax(0.0, dmin(above, below));
        else out[k - 1] = dmin(0.0, dmax(above, below));
        below = above;
    }
    out[mkx] = out[mkx - 1];
}

/* uwshcu.F90:4863-4929 */
double orc_qsinvert(const orc_consts *c, double qt, double thl, double psfc)
{
    double ps, Pis, Ts, err, dlnqsdT, dTdPis, dPisdps, dlnqsdps, derrdps, dps;
    double Ti, rhi, TLCL, PiLCL, psmin = 100.0 * 100.0, dpsmax = 1.0, es, qs, gam, leff, nu;
    Ti = thl * XPOW(psfc / c->p00, c->rovcp);
    orc_qsat(c, Ti, psfc, &es, &qs, &gam);
    rhi = qt / qs;
    if (rhi <= 0.01) return psmin;
    TLCL = 55.0 + 1.0 / (1.0 / (Ti - 55.0) - XLOG(rhi) / 2840.0);
    PiLCL = TLCL / thl;
    ps = c->p00 * XPOW(PiLCL, 1.0 / c->rovcp);
    for (int i = 1; i <= 10; ++i) {
        Pis = XPOW(ps / c->p00, c->rovcp);
        Ts = thl * Pis;
        orc_qsat(c, Ts, ps, &es, &qs, &gam);
        err = qt - qs;
        nu = dmax(dmin((268.0 - Ts) / 20.0, 1.0), 0.0);
        leff = (1.0 - nu) * c->xlv + nu * c->xls;
        dlnqsdT = gam * (c->cp / leff) / qs;
        dTdPis = thl;
        dPisdps = c->rovcp * Pis / ps;
        dlnqsdps = -1.0 / (ps - (1.0 - c->ep2) * es);
        derrdps = -qs * (dlnqsdT * dTdPis * dPisdps + dlnqsdps);
        dps = -err / derrdps;
        ps = ps + dps;
        if (ps < 0.0) return psmin;
        if (fabs(dps) <= dpsmax) return ps;
    }
    return psmin;
}

/* uwshcu.F90:4931-4950 */
double orc_compute_alpha(double del_CIN, double ke)
{
    double x0 = 0.0, x1;
    for (int it = 1; it <= 10; ++it) {
        x1 = x0 - (XEXP(-x0 * ke * del_CIN) - x0) / (-ke * del_CIN * XEXP(-x0 * ke * del_CIN) - 1.0);
        x0 = x1;
    }
    return x0;
}

double orc_xexp(double x) { return XEXP(x); }
double orc_xlog(double x) { return XLOG(x); }
double orc_xpow(double x, double y) { return XPOW(x, y); }
double orc_xerfc(double x) { return XERFC(x); }

/* uwshcu.F90:4952-4979; erfc is Cody's (orc_math.h), the algorithm of the CAM
 * error_function module, which is not part of the reference tree. */
double orc_compute_mumin2(double mulcl, double rmaxfrac, double mulow)
{
    double x0 = mulow, x1, ex, ef, exf, f, fs;
    for (int it = 1; it <= 10; ++it) {
        ex = XEXP(-(x0 * x0));
        ef = XERFC(x0);
        exf = ex / ef;
        f = 0.5 * (exf * exf) - 0.5 * ((ex / 2.0 / rmaxfrac) * (ex / 2.0 / rmaxfrac)) -
            ((mulcl * 2.5066 / 2.0) * (mulcl * 2.5066 / 2.0));
        fs = (2.0 * (exf * exf)) * (exf / sqrt(3.141592) - x0) + (0.5 * x0 * (ex * ex)) / (rmaxfrac * rmaxfrac);
        x1 = x0 - f / fs;
        x0 = x1;
    }
    return x0;
}

/* uwshcu.F90:4981-5028 */
double orc_compute_ppen(double wtwb, double D, double bogbot, double bogtop, double rho0j, double dpen)
{
    double x0, x1, f, fs, SB, s00;
    SB = (bogtop - bogbot) / dpen;
    s00 = bogbot / rho0j - D * wtwb;
    if (D * dpen < 1.e-8) {
        if (s00 >= 0.0) x0 = dpen;
        else x0 = dmax(0.0, dmin(dpen, -0.5 * wtwb / s00));
    } else {
        if (s00 >= 0.0) x0 = dpen;
        else x0 = 0.0;
        for (int it = 1; it <= 5; ++it) {
            f = XEXP(-2.0 * D * x0) * (wtwb - (bogbot - SB / (2.0 * D)) / (D * rho0j)) +
                (SB * x0 + bogbot - SB / (2.0 * D)) / (D * rho0j);
            fs = -2.0 * D * XEXP(-2.0 * D * x0) * (wtwb - (bogbot - SB / (2.0 * D)) / (D * rho0j)) +
                 (SB) / (D * rho0j);
            if (fs >= 0.0) fs = dmax(fs, 1.e-10);
            else fs = dmin(fs, -1.e-10);
            x1 = x0 - f / fs;
            x0 = x1;
        }
    }
    return -dmax(0.0, dmin(dpen, x0));
}

/* uwshcu.F90:5030-5094; ps0 and xflx indexed 0..mkx */
void orc_fluxbelowinv(const orc_consts *c, double cbmf, const double *ps0, int mkx, int kinv, double dt,
                      double xsrc, double xmean, double xtopin, double xbotin, double *xflx)
{
    double rcbmf, rpeff, dp, rr, pinv, xtop, xbot, xtop_ori, xbot_ori;
    for (int k = 0; k <= mkx; ++k) xflx[k] = 0.0;
    dp = ps0[kinv - 1] - ps0[kinv];
    if (fabs(xbotin - xtopin) <= 1.e-13) {
        xbot = xbotin - 1.e-13;
        xtop = xtopin + 1.e-13;
    } else {
        xbot = xbotin;
        xtop = xtopin;
    }
    xtop_ori = xtop;
    xbot_ori = xbot;
    rcbmf = (cbmf * c->g * dt) / dp;
    rpeff = (xmean - xtop) / (xbot - xtop);
    rpeff = dmin(dmax(0.0, rpeff), 1.0);
    if (rpeff == 0.0 || rpeff == 1.0) {
        xbot = xmean;
        xtop = xmean;
    }
    rr = rpeff / rcbmf;
    pinv = ps0[kinv - 1] - rpeff * dp;
    for (int k = 0; k <= kinv - 1; ++k)
        xflx[k] = cbmf * (xsrc - xbot) * (ps0[0] - ps0[k]) / (ps0[0] - pinv);
    if (rr <= 1.0) xflx[kinv - 1] = xflx[kinv - 1] - (1.0 - rr) * cbmf * (xtop_ori - xbot_ori);
}

/* uwshcu.F90:5096-5162; arrays indexed 1..mkx */
void orc_positive_moisture_single(double xlv, double xls, int mkx, double dt, double qvmin, double qlmin,
                                  double qimin, const double *dp, double *qv, double *ql, double *qi,
                                  double *s, double *qvten, double *qlten, double *qiten, double *sten)
{
    double dql, dqi, dqv = 0.0, sum, aa, dum;
    for (int k = mkx; k >= 1; --k) {
        dql = dmax(0.0, 1.0 * qlmin - ql[k]);
        dqi = dmax(0.0, 1.0 * qimin - qi[k]);
        qlten[k] = qlten[k] + dql / dt;
        qiten[k] = qiten[k] + dqi / dt;
        qvten[k] = qvten[k] - (dql + dqi) / dt;
        sten[k] = sten[k] + xlv * (dql / dt) + xls * (dqi / dt);
        ql[k] = ql[k] + dql;
        qi[k] = qi[k] + dqi;
        qv[k] = qv[k] - dql - dqi;
        s[k] = s[k] + xlv * dql + xls * dqi;
        dqv = dmax(0.0, 1.0 * qvmin - qv[k]);
        qvten[k] = qvten[k] + dqv / dt;
        qv[k] = qv[k] + dqv;
        if (k != 1) {
            qv[k - 1] = qv[k - 1] - dqv * dp[k] / dp[k - 1];
            qvten[k - 1] = qvten[k - 1] - dqv * dp[k] / dp[k - 1] / dt;
        }
        qv[k] = dmax(qv[k], qvmin);
        ql[k] = dmax(ql[k], qlmin);
        qi[k] = dmax(qi[k], qimin);
    }
    if (dqv > 1.e-20) {
        sum = 0.0;
        for (int k = 1; k <= mkx; ++k)
            if (qv[k] > 2.0 * qvmin) sum = sum + qv[k] * dp[k];
        aa = dqv * dp[1] / dmax(1.e-20, sum);
        if (aa < 0.5) {
            for (int k = 1; k <= mkx; ++k) {
                if (qv[k] > 2.0 * qvmin) {
                    dum = aa * qv[k];
                    qv[k] = qv[k] - dum;
                    qvten[k] = qvten[k] - dum / dt;
                }
            }
        }
    }
}

/* uwshcu.F90:4704-4715 */
double orc_single_cin(const orc_consts *c, double pbot, double thv0bot, double ptop, double thv0top,
                      double thvubot, double thvutop)
{
    return ((1.0 - thvubot / thv0bot) + (1.0 - thvutop / thv0top)) * (pbot - ptop) /
           (pbot / (c->r * thv0bot * exnf(c, pbot)) + ptop / (c->r * thv0top * exnf(c, ptop)));
}

/* uwshcu.F90:4672-4702 */
void orc_getbuoy(const orc_consts *c, double pbot, double thv0bot, double ptop, double thv0top,
                 double thvubot, double thvutop, double *plfc, double *cin)
{
    double frc;
    if (thvubot > thv0bot && thvutop > thv0top) {
        *plfc = pbot;
        return;
    } else if (thvubot <= thv0bot && thvutop <= thv0top) {
        *cin = *cin - ((thvubot / thv0bot - 1.0) + (thvutop / thv0top - 1.0)) * (pbot - ptop) /
                          (pbot / (c->r * thv0bot * exnf(c, pbot)) + ptop / (c->r * thv0top * exnf(c, ptop)));
    } else if (thvubot > thv0bot && thvutop <= thv0top) {
        frc = (thvutop / thv0top - 1.0) / ((thvutop / thv0top - 1.0) - (thvubot / thv0bot - 1.0));
        *cin = *cin - (thvutop / thv0top - 1.0) * ((ptop + frc * (pbot - ptop)) - ptop) /
                          (pbot / (c->r * thv0bot * exnf(c, pbot)) + ptop / (c->r * thv0top * exnf(c, ptop)));
    } else {
        frc = (thvubot / thv0bot - 1.0) / ((thvubot / thv0bot - 1.0) - (thvutop / thv0top - 1.0));
        *plfc = pbot - frc * (pbot - ptop);
        *cin = *cin - (thvubot / thv0bot - 1.0) * (pbot - *plfc) /
                          (pbot / (c->r * thv0bot * exnf(c, pbot)) + ptop / (c->r * thv0top * exnf(c, ptop)));
    }
}

/* findsp for a single point, uwshcu.F90:5246-5450.  The reference iterates a
 * whole level of a chunk together; a point's sequence of updates depends only
 * on its own inputs (converged points are masked out, and the chunk-wide early
 * exit fires only when no unconverged point is left), so per point is exact. */
int orc_findsp_point(const orc_consts *c, double q, double t, double p, double *tsp_out, double *qsp_out,
                     double *enerr)
{
    const double omeps = 1.0 - c->epsqs, trinv = 1.0 / c->ttrice;
    const double a1 = 7.5 * XLOG(10.0), rair = 287.04, c3 = rair * a1 / c->cpws;
    const double dttol = 1.e-4, dqtol = 1.e-4, tt0 = 273.15;
    double tlim, es, denom, qs, tc, weight, hlatsb, hlatvp, hltalt, enin, enout = 1.0, tmp, c1, c2, r1b, qvd;
    double tsp, qsp, tterm, desdt, dqsdt, gg, dgdt, t1, dt, q1, dq;
    int doit, lflg;
    tlim = dmin(dmax(t, 173.0), 373.0);
    es = orc_estblf(c, tlim);
    denom = p - omeps * es;
    qs = c->epsqs * es / denom;
    doit = 0;
    if (p > 5.0 * es && qs > 0.0 && qs < 0.5) {
        qs = dmin(c->epsqs * es / denom, 1.0);
        tc = tlim - tt0;
        weight = dmin(-tc * trinv, 1.0);
        hlatsb = c->hlatv + weight * c->hlatf;
        hlatvp = c->hlatv - 2369.0 * tc;
        hltalt = (tlim < tt0) ? hlatsb : hlatvp;
        enin = c->cpws * tlim + hltalt * q;
        tmp = q - qs;
        c1 = hltalt * c3;
        c2 = (tlim + 36.0) * (tlim + 36.0);
        r1b = c2 / (c2 + c1 * qs);
        qvd = r1b * tmp;
        tsp = tlim + ((hltalt / c->cpws) * qvd);
        es = orc_estblf(c, tsp);
        qsp = dmin(c->epsqs * es / (p - omeps * es), 1.0);
    } else {
        doit = 1;
        tsp = tlim;
        qsp = q;
        enin = 1.0;
    }
    for (int l = 1; l <= 8 && doit == 0; ++l) {
        es = orc_estblf(c, tsp);
        qs = dmin(c->epsqs * es / (p - omeps * es), 1.0);
        tc = tsp - tt0;
        lflg = (tc >= -c->ttrice && tc < 0.0);
        weight = dmin(-tc * trinv, 1.0);
        hlatsb = c->hlatv + weight * c->hlatf;
        hlatvp = c->hlatv - 2369.0 * tc;
        hltalt = (tsp < tt0) ? hlatsb : hlatvp;
        if (lflg) tterm = c->pcf[0] + tc * (c->pcf[1] + tc * (c->pcf[2] + tc * (c->pcf[3] + tc * c->pcf[4])));
        else tterm = 0.0;
        desdt = hltalt * es / (c->rgasv * tsp * tsp) + tterm * trinv;
        dqsdt = (c->epsqs + omeps * qs) / (p - omeps * es) * desdt;
        gg = enin - (c->cpws * tsp + hltalt * qsp);
        dgdt = -(c->cpws + hltalt * dqsdt);
        t1 = tsp - gg / dgdt;
        dt = fabs(t1 - tsp) / t1;
        tsp = dmax(t1, c->tmin);
        es = orc_estblf(c, tsp);
        q1 = dmin(c->epsqs * es / (p - omeps * es), 1.0);
        dq = fabs(q1 - qsp) / dmax(q1, 1.e-12);
        qsp = q1;
        if (dt < dttol && dq < dqtol) doit = 2;
        enout = c->cpws * tsp + hltalt * qsp;
        if (tsp < 174.16) doit = 4;
    }
    *tsp_out = tsp;
    *qsp_out = qsp;
    *enerr = (doit == 2) ? fabs((enin - enout) / (enin + enout)) : 0.0;
    return doit;
}

/* ------------------------------------------------------------------------- */
/* One column of compute_uwshcu (uwshcu.F90:1218-4549)                         */
/* ------------------------------------------------------------------------- */

typedef struct colws {
    /* inputs (UW bottom-up order) */
    double ps0[KMAX], zs0[KMAX], p0[KMAX], z0[KMAX], dp0[KMAX], dpdry0[KMAX], u0[KMAX], v0[KMAX], tke[KMAX];
    double qv0[KMAX], ql0[KMAX], qi0[KMAX], t0[KMAX], s0[KMAX], qw0[KMAX];
    double tr0[NT][KMAX];
    double pblh, cush;
    /* compute_zwshcu only: omega (layers) and the capped tpert */
    double omega[KMAX], tpert;
    /* environment */
    double qt0[KMAX], thl0[KMAX], thvl0[KMAX], ssqt0[KMAX], ssthl0[KMAX], ssu0[KMAX], ssv0[KMAX];
    double thv0bot[KMAX], thv0top[KMAX], thvl0bot[KMAX], thvl0top[KMAX], exn0[KMAX], exns0[KMAX];
    double sstr0[NT][KMAX];
    double qv0_star[KMAX], ql0_star[KMAX], qi0_star[KMAX], s0_star[KMAX];
    /* convection */
    double umf[KMAX], emf[KMAX], qvten[KMAX], qlten[KMAX], qiten[KMAX], sten[KMAX], uten[KMAX], vten[KMAX];
    double qrten[KMAX], qsten[KMAX], evapc[KMAX], slflx[KMAX], qtflx[KMAX], uflx[KMAX], vflx[KMAX];
    double cufrc[KMAX], qcu[KMAX], qlu[KMAX], qiu[KMAX], dwten[KMAX], diten[KMAX], fer[KMAX], fdr[KMAX];
    double uf[KMAX], vf[KMAX], qc[KMAX], qc_l[KMAX], qc_i[KMAX], qtten[KMAX], slten[KMAX], ufrc[KMAX];
    double trten[NT][KMAX], trflx[NT][KMAX], trflx_d[KMAX], trflx_u[KMAX];
    double uemf[KMAX], comsub[KMAX], qlten_sink[KMAX], qiten_sink[KMAX], nlten_sink[KMAX], niten_sink[KMAX];
    double wu[KMAX], thlu[KMAX], qtu[KMAX], uu[KMAX], vu[KMAX], thvu[KMAX], rei[KMAX];
    double tru[NT][KMAX];
    double thlu_emf[KMAX], qtu_emf[KMAX], uu_emf[KMAX], vu_emf[KMAX];
    double tru_emf[NT][KMAX];
    double flxrain[KMAX], flxsnow[KMAX], ntraprd[KMAX], ntsnprd[KMAX];
    double xflx[KMAX];
    /* iteration-1 saves (_s) */
    double qv0_s[KMAX], ql0_s[KMAX], qi0_s[KMAX], s0_s[KMAX], u0_s[KMAX], v0_s[KMAX], t0_s[KMAX], qt0_s[KMAX];
    double qvten_s[KMAX], qlten_s[KMAX], qiten_s[KMAX], qrten_s[KMAX], qsten_s[KMAX], sten_s[KMAX];
    double evapc_s[KMAX], uten_s[KMAX], vten_s[KMAX], cufrc_s[KMAX], qcu_s[KMAX], qlu_s[KMAX], qiu_s[KMAX];
    double qc_s[KMAX], umf_s[KMAX], slflx_s[KMAX], qtflx_s[KMAX];
    double tr0_s[NT][KMAX], trten_s[NT][KMAX];
    double cush_s, precip_s, snow_s, rliq_s, cbmf_s, cnt_s, cnb_s;
    /* originals (_o) */
    double qv0_o[KMAX], ql0_o[KMAX], qi0_o[KMAX], t0_o[KMAX], s0_o[KMAX], u0_o[KMAX], v0_o[KMAX];
    double qt0_o[KMAX], thl0_o[KMAX], thvl0_o[KMAX], ssthl0_o[KMAX], ssqt0_o[KMAX], thv0bot_o[KMAX];
    double thv0top_o[KMAX], thvl0bot_o[KMAX], thvl0top_o[KMAX], ssu0_o[KMAX], ssv0_o[KMAX];
    double tr0_o[NT][KMAX], sstr0_o[NT][KMAX];
    /* outputs (UW order) */
    double o_umf[KMAX], o_wu[KMAX], o_slflx[KMAX], o_qtflx[KMAX], o_flxprc1[KMAX], o_flxsnow1[KMAX];
    double o_qvten[KMAX], o_qlten[KMAX], o_qiten[KMAX], o_sten[KMAX], o_uten[KMAX], o_vten[KMAX];
    double o_qrten[KMAX], o_qsten[KMAX], o_evapc[KMAX], o_cufrc[KMAX], o_qcu[KMAX], o_qlu[KMAX], o_qiu[KMAX];
    double o_qc[KMAX];
    double o_trten[NT][KMAX];
    double o_precip, o_snow, o_cbmf, o_rliq, o_cnt, o_cnb;
    int diag[ORC_NDIAG];
} colws;

/* exit reasons stored in diag[0] */
enum { EX_NONE = 0, EX_CONDEN = 1, EX_KINV1 = 2, EX_KLCLMKX = 3, EX_KLFCMKX = 4, EX_MU3 = 5, EX_WTW = 6,
       EX_UFRC = 7, EX_WU = 8, EX_CUFILTER = 9, EX_SUBCONDEN = 10, EX_KBUPCONDEN = 11 };

#define THV(thj, qvj, qlj, qij) ((thj) * (1.0 + c->zvir * (qvj) - (qlj) - (qij)))

/* zmh_ramp (mzfunctions_mod.F90:222-248): 0 below x0, 1 above x0 + xint, linear
 * in between.  The floor on xint is the single-precision literal 1.e-20. */
static double zmh_ramp(double x, double x0, double xint)
{
    double x2;
    if (x < x0) x2 = 0.0;
    else if (x > x0 + xint) x2 = 1.0;
    else x2 = (x - x0) / dmax(xint, (double)1.e-20f);
    return dmin(x2, 1.0);
}

/* One column of compute_uwshcu, or of compute_zwshcu when zw is not NULL.
 * Returns 0 or -3 (mu > mumin2). */
static int column(const orc_consts *c, colws *W, int mkx, int ntr, double dt, const orc_zw *zw)
{
    /* uwshcu.F90:1040-1068 parameters */
    const double rkfre = 1.0, rmaxfrac = 0.10, mumin1 = 0.906, rbuoy = 1.0, rdrag = 1.0;
    const double epsvarw = 5.e-4, PGFc = 0.7, criqc = 0.7e-3, frc_rasn = 1.0, kevp = 2.e-6;
    const double xlv = c->xlv, xls = c->xls, cp = c->cp, zvir = c->zvir, r = c->r, g = c->g;
    const double rle = c->rle, rkm = c->rkm, rpen = c->rpen;
    const int ixnumliq = c->ixnumliq, ixnumice = c->ixnumice;

    double *ps0 = W->ps0, *zs0 = W->zs0, *p0 = W->p0, *z0 = W->z0, *dp0 = W->dp0, *u0 = W->u0, *v0 = W->v0;
    double *tke = W->tke, *qv0 = W->qv0, *ql0 = W->ql0, *qi0 = W->qi0, *t0 = W->t0, *s0 = W->s0;
    double *qt0 = W->qt0, *thl0 = W->thl0, *thvl0 = W->thvl0, *ssqt0 = W->ssqt0, *ssthl0 = W->ssthl0;
    double *ssu0 = W->ssu0, *ssv0 = W->ssv0, *thv0bot = W->thv0bot, *thv0top = W->thv0top;
    double *thvl0bot = W->thvl0bot, *thvl0top = W->thvl0top, *exn0 = W->exn0, *exns0 = W->exns0;
    double *umf = W->umf, *emf = W->emf, *qvten = W->qvten, *qlten = W->qlten, *qiten = W->qiten;
    double *sten = W->sten, *uten = W->uten, *vten = W->vten, *qrten = W->qrten, *qsten = W->qsten;
    double *evapc = W->evapc, *slflx = W->slflx, *qtflx = W->qtflx, *uflx = W->uflx, *vflx = W->vflx;
    double *cufrc = W->cufrc, *qcu = W->qcu, *qlu = W->qlu, *qiu = W->qiu, *dwten = W->dwten;
    double *diten = W->diten, *fer = W->fer, *fdr = W->fdr, *uf = W->uf, *vf = W->vf, *qc = W->qc;
    double *qc_l = W->qc_l, *qc_i = W->qc_i, *qtten = W->qtten, *slten = W->slten, *ufrc = W->ufrc;
    double *uemf = W->uemf, *comsub = W->comsub, *qlten_sink = W->qlten_sink, *qiten_sink = W->qiten_sink;
    double *nlten_sink = W->nlten_sink, *niten_sink = W->niten_sink, *wu = W->wu, *thlu = W->thlu;
    double *qtu = W->qtu, *uu = W->uu, *vu = W->vu, *thvu = W->thvu, *rei = W->rei;
    double *thlu_emf = W->thlu_emf, *qtu_emf = W->qtu_emf, *uu_emf = W->uu_emf, *vu_emf = W->vu_emf;
    double *flxrain = W->flxrain, *flxsnow = W->flxsnow, *ntraprd = W->ntraprd, *ntsnprd = W->ntsnprd;
    double *xflx = W->xflx, *trflx_d = W->trflx_d, *trflx_u = W->trflx_u;

    int k, m, kk, iter, iter_scaleh, iter_xc, id_check, status, km1, kp1;
    int klcl = 0, kinv = 0, krel = 0, klfc = 0, kbup = 0, kpen = 0, id_exit = 0, forcedCu;
    int kinv_o = 0, klcl_o = 0, klfc_o = 0, exit_code = EX_NONE, exit_iter = 0, path = 0;
    int limit_ufrc = 0, limit_ppen = 0, limit_emf = 0;
    double thlsrc = 0, qtsrc = 0, usrc = 0, vsrc = 0, thvlsrc = 0, uplus, vplus;
    double trsrc[NT], tre[NT], trsrc_o[NT];
    double plcl = 0, plfc = 0, prel = 0, wrel = 0;
    double ee2, ud2, wtw = 0, wtwb = 0, xc = 0, tscaleh, scaleh, cridis;
    double sigmaw, tkeavg = 0, dpsum, dpi, thvlmin = 0;
    double thlxsat, qtxsat, thvxsat, x_cu = 0, x_en = 0, thv_x0, thv_x1;
    double thj, qvj, qlj, qij, thvj, tj, thv0j = 0, rho0j = 0, rhos0j, qse;
    double cin = 0, cinlcl = 0, pe = 0, dpe = 0, exne = 0, thvebot = 0, thle = 0, qte = 0, ue = 0, ve = 0;
    double thlue, qtue, wue;
    double mu, mumin0, mumin2, mulcl, mulclstar, cbmf = 0, wcrit, winv, wlcl = 0, ufrclcl = 0;
    double exql, exqi, ppen = 0, thl0top, thl0bot, qt0bot, qt0top, thvubot = 0, thvutop = 0;
    double thlu_top = 0, qtu_top = 0, qlu_top, qiu_top, qlu_mid, qiu_mid, exntop;
    double thl0lcl, qt0lcl, thv0lcl = 0, thv0rel = 0, rho0inv, aquad, bquad, cquad, xc1, xc2;
    double excessu, excess0, xsat, xs1, xs2, bogbot = 0, bogtop = 0, delbog, drage = 0, expfac;
    double qcubelow, qlubelow = 0, qiubelow = 0, rainflx, snowflx, es, qs, gam, qsat_arg;
    double xsrc, xmean, xtop, xbot, tmp1, tmp2, tmp3;
    double thlten_sub, qtten_sub, qlten_sub, qiten_sub, nlten_sub, niten_sub, thl_prog, qt_prog;
    double qc_lm, qc_im, nc_lm, nc_im, ql_emf_kbup, qi_emf_kbup, nl_emf_kbup, ni_emf_kbup;
    double snowmlt, subsat, evprain, evpsnow, evplimit, evplimit_rain, evplimit_snow, evpint_rain, evpint_snow;
    double precip = 0, snow = 0, rliq = 0, cnt, cnb, trmin, pdelx, dum;
    double cin_i = 0, cin_f, del_CIN, ke = 0, alpha, cinlcl_i = 0;
    double plcl_o = 0, plfc_o = 0, tkeavg_o = 0, thvlmin_o = 0, qtsrc_o = 0, thvlsrc_o = 0, thlsrc_o = 0;
    double usrc_o = 0, vsrc_o = 0, thv0lcl_o = 0, cush = W->cush, pblh = W->pblh;
    int mumin2_fail = 0;

    (void)thvl0; (void)qse; (void)tj; (void)qlu_top; (void)qiu_top; (void)thvlsrc_o;
    for (m = 0; m < NT; ++m) { trsrc[m] = tre[m] = trsrc_o[m] = 0.0; }

    /* ---- S0 environment (uwshcu.F90:1255-1299) ---- */
    for (k = 1; k <= mkx; ++k) exn0[k] = XPOW(p0[k] / c->p00, c->rovcp);
    for (k = 0; k <= mkx; ++k) exns0[k] = XPOW(ps0[k] / c->p00, c->rovcp);
    for (k = 1; k <= mkx; ++k) {
        qt0[k] = (qv0[k] + ql0[k] + qi0[k]);
        thl0[k] = (t0[k] - xlv * ql0[k] / cp - xls * qi0[k] / cp) / exn0[k];
        thvl0[k] = (1.0 + zvir * qt0[k]) * thl0[k];
    }
    orc_slope(mkx, thl0, p0, ssthl0);
    orc_slope(mkx, qt0, p0, ssqt0);
    orc_slope(mkx, u0, p0, ssu0);
    orc_slope(mkx, v0, p0, ssv0);
    for (m = 1; m <= ntr; ++m) orc_slope(mkx, W->tr0[m], p0, W->sstr0[m]);
    for (k = 1; k <= mkx; ++k) {
        thl0bot = thl0[k] + ssthl0[k] * (ps0[k - 1] - p0[k]);
        qt0bot = qt0[k] + ssqt0[k] * (ps0[k - 1] - p0[k]);
        orc_conden(c, ps0[k - 1], thl0bot, qt0bot, &thj, &qvj, &qlj, &qij, &qse, &id_check);
        if (id_check == 1) { exit_code = EX_CONDEN; id_exit = 1; goto L333; }
        thv0bot[k] = THV(thj, qvj, qlj, qij);
        thvl0bot[k] = thl0bot * (1.0 + zvir * qt0bot);
        thl0top = thl0[k] + ssthl0[k] * (ps0[k] - p0[k]);
        qt0top = qt0[k] + ssqt0[k] * (ps0[k] - p0[k]);
        orc_conden(c, ps0[k], thl0top, qt0top, &thj, &qvj, &qlj, &qij, &qse, &id_check);
        if (id_check == 1) { exit_code = EX_CONDEN; id_exit = 1; goto L333; }
        thv0top[k] = THV(thj, qvj, qlj, qij);
        thvl0top[k] = thl0top * (1.0 + zvir * qt0top);
    }

    /* ---- save originals (uwshcu.F90:1306-1327) ---- */
    for (k = 1; k <= mkx; ++k) {
        W->qv0_o[k] = qv0[k]; W->ql0_o[k] = ql0[k]; W->qi0_o[k] = qi0[k]; W->t0_o[k] = t0[k];
        W->s0_o[k] = s0[k]; W->u0_o[k] = u0[k]; W->v0_o[k] = v0[k]; W->qt0_o[k] = qt0[k];
        W->thl0_o[k] = thl0[k]; W->thvl0_o[k] = thvl0[k]; W->ssthl0_o[k] = ssthl0[k];
        W->ssqt0_o[k] = ssqt0[k]; W->thv0bot_o[k] = thv0bot[k]; W->thv0top_o[k] = thv0top[k];
        W->thvl0bot_o[k] = thvl0bot[k]; W->thvl0top_o[k] = thvl0top[k]; W->ssu0_o[k] = ssu0[k];
        W->ssv0_o[k] = ssv0[k];
        for (m = 1; m <= ntr; ++m) { W->tr0_o[m][k] = W->tr0[m][k]; W->sstr0_o[m][k] = W->sstr0[m][k]; }
    }

    /* ---- initialize column arrays (uwshcu.F90:1333-1408) ---- */
#define ZERO(a) memset(W->a, 0, sizeof(double) * (size_t)(mkx + 1))
#define ZT(a) for (m = 0; m <= ntr; ++m) memset(W->a[m], 0, sizeof(double) * (size_t)(mkx + 1))
    ZERO(umf); ZERO(emf); ZERO(slflx); ZERO(qtflx); ZERO(uflx); ZERO(vflx); ZERO(qvten); ZERO(qlten);
    ZERO(qiten); ZERO(sten); ZERO(uten); ZERO(vten); ZERO(qrten); ZERO(qsten); ZERO(dwten); ZERO(diten);
    ZERO(evapc); ZERO(cufrc); ZERO(qcu); ZERO(qlu); ZERO(qiu); ZERO(fer); ZERO(fdr); ZERO(qc); ZERO(qc_l);
    ZERO(qc_i); ZERO(qtten); ZERO(slten); ZERO(ufrc); ZERO(thlu); ZERO(qtu); ZERO(uu); ZERO(vu); ZERO(wu);
    ZERO(thvu); ZERO(thlu_emf); ZERO(qtu_emf); ZERO(uu_emf); ZERO(vu_emf); ZERO(uemf); ZERO(comsub);
    ZERO(qlten_sink); ZERO(qiten_sink); ZERO(nlten_sink); ZERO(niten_sink); ZT(trflx); ZT(trten);
    ZT(tru); ZT(tru_emf); ZERO(rei); ZERO(uf); ZERO(vf);
    ZERO(flxrain); ZERO(flxsnow); ZERO(ntraprd); ZERO(ntsnprd);
    cin = 0.0; cbmf = 0.0; rliq = 0.0; cnt = (double)mkx; cnb = 0.0; precip = 0.0; snow = 0.0;

    for (iter = 1; iter <= 2; ++iter) {
        exit_iter = iter;
        tscaleh = cush;
        cush = -1.0;

        /* ---- S1 kinv (uwshcu.F90:1455-1468) ---- */
        kinv = 1;
        for (k = mkx - 1; k >= 1; --k) {
            if ((pblh + 5.0 - zs0[k]) * (pblh + 5.0 - zs0[k + 1]) < 0.0) { kinv = k + 1; break; }
        }
        if (kinv <= 1) { exit_code = EX_KINV1; id_exit = 1; goto L333; }

        /* ---- S2 PBL averages (uwshcu.F90:1484-1499) ---- */
        dpsum = 0.0;
        tkeavg = 0.0;
        thvlmin = 1000.0;
        for (k = 0; k <= kinv - 1; ++k) {
            if (k == 0) dpi = ps0[0] - p0[1];
            else if (k == (kinv - 1)) dpi = p0[kinv - 1] - ps0[kinv - 1];
            else dpi = p0[k] - p0[k + 1];
            dpsum = dpsum + dpi;
            tkeavg = tkeavg + dpi * tke[k];
            if (k != 0) thvlmin = dmin(thvlmin, dmin(thvl0bot[k], thvl0top[k]));
        }
        tkeavg = tkeavg / dpsum;

        /* ---- S3 source air (uwshcu.F90:1509-1516) ---- */
        qtsrc = qt0[1];
        thvlsrc = thvlmin;
        thlsrc = thvlsrc / (1.0 + zvir * qtsrc);
        usrc = u0[kinv - 1] + ssu0[kinv - 1] * (ps0[kinv - 1] - p0[kinv - 1]);
        vsrc = v0[kinv - 1] + ssv0[kinv - 1] * (ps0[kinv - 1] - p0[kinv - 1]);
        for (m = 1; m <= ntr; ++m) trsrc[m] = W->tr0[m][1];

        /* ---- S4 LCL (uwshcu.F90:1533-1549) ---- */
        plcl = orc_qsinvert(c, qtsrc, thlsrc, ps0[0]);
        klcl = mkx;
        for (k = 0; k <= mkx; ++k) {
            if (ps0[k] < plcl) { klcl = k; break; }
        }
        klcl = klcl > 1 ? klcl : 1;
        if (plcl < 30000.0) { exit_code = EX_KLCLMKX; id_exit = 1; goto L333; }

        /* ---- S5 thv0lcl (uwshcu.F90:1559-1567) ---- */
        thl0lcl = thl0[klcl] + ssthl0[klcl] * (plcl - p0[klcl]);
        qt0lcl = qt0[klcl] + ssqt0[klcl] * (plcl - p0[klcl]);
        orc_conden(c, plcl, thl0lcl, qt0lcl, &thj, &qvj, &qlj, &qij, &qse, &id_check);
        if (id_check == 1) { exit_code = EX_CONDEN; id_exit = 1; goto L333; }
        thv0lcl = THV(thj, qvj, qlj, qij);

        /* ---- S6 CIN (uwshcu.F90:1592-1685) ---- */
        cin = 0.0;
        cinlcl = 0.0;
        plfc = 0.0;
        klfc = mkx;
        if (klcl >= kinv) {
            for (k = kinv; k <= mkx - 1; ++k) {
                if (k < klcl) {
                    thvubot = thvlsrc;
                    thvutop = thvlsrc;
                    cin = cin + orc_single_cin(c, ps0[k - 1], thv0bot[k], ps0[k], thv0top[k], thvubot, thvutop);
                } else if (k == klcl) {
                    thvubot = thvlsrc;
                    thvutop = thvlsrc;
                    cin = cin + orc_single_cin(c, ps0[k - 1], thv0bot[k], plcl, thv0lcl, thvubot, thvutop);
                    cinlcl = dmax(cin, 0.0);
                    cin = cinlcl;
                    thvubot = thvlsrc;
                    orc_conden(c, ps0[k], thlsrc, qtsrc, &thj, &qvj, &qlj, &qij, &qse, &id_check);
                    if (id_check == 1) { exit_code = EX_CONDEN; id_exit = 1; goto L333; }
                    thvutop = THV(thj, qvj, qlj, qij);
                    orc_getbuoy(c, plcl, thv0lcl, ps0[k], thv0top[k], thvubot, thvutop, &plfc, &cin);
                    if (plfc > 0.0) { klfc = k; goto L35; }
                } else {
                    thvubot = thvutop;
                    orc_conden(c, ps0[k], thlsrc, qtsrc, &thj, &qvj, &qlj, &qij, &qse, &id_check);
                    if (id_check == 1) { exit_code = EX_CONDEN; id_exit = 1; goto L333; }
                    thvutop = THV(thj, qvj, qlj, qij);
                    orc_getbuoy(c, ps0[k - 1], thv0bot[k], ps0[k], thv0top[k], thvubot, thvutop, &plfc, &cin);
                    if (plfc > 0.0) { klfc = k; goto L35; }
                }
            }
        } else {
            cinlcl = 0.0;
            for (k = kinv; k <= mkx - 1; ++k) {
                orc_conden(c, ps0[k - 1], thlsrc, qtsrc, &thj, &qvj, &qlj, &qij, &qse, &id_check);
                if (id_check == 1) { exit_code = EX_CONDEN; id_exit = 1; goto L333; }
                thvubot = THV(thj, qvj, qlj, qij);
                orc_conden(c, ps0[k], thlsrc, qtsrc, &thj, &qvj, &qlj, &qij, &qse, &id_check);
                if (id_check == 1) { exit_code = EX_CONDEN; id_exit = 1; goto L333; }
                thvutop = THV(thj, qvj, qlj, qij);
                orc_getbuoy(c, ps0[k - 1], thv0bot[k], ps0[k], thv0top[k], thvubot, thvutop, &plfc, &cin);
                if (plfc > 0.0) { klfc = k; goto L35; }
            }
        }
    L35:
        cin = dmax(0.0, cin);
        if (klfc >= mkx) { klfc = mkx; exit_code = EX_KLFCMKX; id_exit = 1; goto L333; }

        /* ---- S7 implicit CIN (uwshcu.F90:1693-2014) ---- */
        if (iter == 1) {
            cin_i = cin;
            cinlcl_i = cinlcl;
            ke = rbuoy / (rkfre * tkeavg + epsvarw);
            kinv_o = kinv; klcl_o = klcl; klfc_o = klfc; plcl_o = plcl; plfc_o = plfc;
            tkeavg_o = tkeavg; thvlmin_o = thvlmin; qtsrc_o = qtsrc; thvlsrc_o = thvlsrc;
            thlsrc_o = thlsrc; usrc_o = usrc; vsrc_o = vsrc; thv0lcl_o = thv0lcl;
            for (m = 1; m <= ntr; ++m) trsrc_o[m] = trsrc[m];
        }
        if (iter != 1) {
            cin_f = cin;
            del_CIN = cin_f - cin_i;
            if (del_CIN > 0.0) {
                alpha = orc_compute_alpha(del_CIN, ke);
                cin = cin_i + alpha * del_CIN;
                cinlcl = cinlcl_i;
                kinv = kinv_o; klcl = klcl_o; klfc = klfc_o; plcl = plcl_o; plfc = plfc_o;
                tkeavg = tkeavg_o; thvlmin = thvlmin_o; qtsrc = qtsrc_o; thvlsrc = thvlsrc_o;
                thlsrc = thlsrc_o; usrc = usrc_o; vsrc = vsrc_o; thv0lcl = thv0lcl_o;
                for (m = 1; m <= ntr; ++m) trsrc[m] = trsrc_o[m];
                for (k = 1; k <= mkx; ++k) {
                    qv0[k] = W->qv0_o[k]; ql0[k] = W->ql0_o[k]; qi0[k] = W->qi0_o[k]; t0[k] = W->t0_o[k];
                    s0[k] = W->s0_o[k]; u0[k] = W->u0_o[k]; v0[k] = W->v0_o[k]; qt0[k] = W->qt0_o[k];
                    thl0[k] = W->thl0_o[k]; thvl0[k] = W->thvl0_o[k]; ssthl0[k] = W->ssthl0_o[k];
                    ssqt0[k] = W->ssqt0_o[k]; thv0bot[k] = W->thv0bot_o[k]; thv0top[k] = W->thv0top_o[k];
                    thvl0bot[k] = W->thvl0bot_o[k]; thvl0top[k] = W->thvl0top_o[k]; ssu0[k] = W->ssu0_o[k];
                    ssv0[k] = W->ssv0_o[k];
                    for (m = 1; m <= ntr; ++m) { W->tr0[m][k] = W->tr0_o[m][k]; W->sstr0[m][k] = W->sstr0_o[m][k]; }
                }
                ZERO(umf); ZERO(emf); ZERO(slflx); ZERO(qtflx); ZERO(uflx); ZERO(vflx); ZERO(qvten);
                ZERO(qlten); ZERO(qiten); ZERO(sten); ZERO(uten); ZERO(vten); ZERO(qrten); ZERO(qsten);
                ZERO(dwten); ZERO(diten); ZERO(evapc); ZERO(cufrc); ZERO(qcu); ZERO(qlu); ZERO(qiu);
                ZERO(fer); ZERO(fdr); ZERO(qc); ZERO(qc_l); ZERO(qc_i); ZERO(qtten); ZERO(slten);
                ZERO(ufrc); ZERO(thlu); ZERO(qtu); ZERO(uu); ZERO(vu); ZERO(wu); ZERO(thvu);
                ZERO(thlu_emf); ZERO(qtu_emf); ZERO(uu_emf); ZERO(vu_emf);
                ZT(trflx); ZT(trten); ZT(tru); ZT(tru_emf);
                precip = 0.0; snow = 0.0; rliq = 0.0; cbmf = 0.0; cnt = (double)mkx; cnb = 0.0;
            } else {
                /* keep iteration-1 outputs (uwshcu.F90:1897-2010); flxprc1/flxsnow1 are not
                 * restored there and keep their initial zero; wu comes from the live array */
                for (k = 0; k <= mkx; ++k) {
                    W->o_umf[k] = W->umf_s[k];
                    W->o_wu[k] = wu[k];
                    W->o_slflx[k] = W->slflx_s[k];
                    W->o_qtflx[k] = W->qtflx_s[k];
                    W->o_flxprc1[k] = 0.0;
                    W->o_flxsnow1[k] = 0.0;
                }
                for (k = 1; k <= mkx; ++k) {
                    W->o_qvten[k] = W->qvten_s[k]; W->o_qlten[k] = W->qlten_s[k]; W->o_qiten[k] = W->qiten_s[k];
                    W->o_sten[k] = W->sten_s[k]; W->o_uten[k] = W->uten_s[k]; W->o_vten[k] = W->vten_s[k];
                    W->o_qrten[k] = W->qrten_s[k]; W->o_qsten[k] = W->qsten_s[k]; W->o_evapc[k] = W->evapc_s[k];
                    W->o_cufrc[k] = W->cufrc_s[k]; W->o_qcu[k] = W->qcu_s[k]; W->o_qlu[k] = W->qlu_s[k];
                    W->o_qiu[k] = W->qiu_s[k]; W->o_qc[k] = W->qc_s[k];
                    for (m = 1; m <= ntr; ++m) W->o_trten[m][k] = W->trten_s[m][k];
                }
                W->o_precip = W->precip_s; W->o_snow = W->snow_s; W->cush = W->cush_s;
                W->o_cbmf = W->cbmf_s; W->o_rliq = W->rliq_s; W->o_cnt = W->cnt_s; W->o_cnb = W->cnb_s;
                path = 1;
                goto Ldone;
            }
        }

        /* ---- S8 release level (uwshcu.F90:2035-2043) ---- */
        if (klcl < kinv) {
            krel = kinv;
            prel = ps0[krel - 1];
            thv0rel = thv0bot[krel];
        } else {
            krel = klcl;
            prel = plcl;
            thv0rel = thv0lcl;
        }

        /* ---- S9 cloud-base closure (uwshcu.F90:2092-2184; zwshcu.F90:2275-2384) ---- */
        if (!zw) {
            wcrit = sqrt(2.0 * cin * rbuoy);
            sigmaw = sqrt(rkfre * tkeavg + epsvarw);
            mu = wcrit / sigmaw / 1.4142;
            if (mu >= 3.0) { exit_code = EX_MU3; id_exit = 1; goto L333; }
            rho0inv = ps0[kinv - 1] / (r * thv0top[kinv - 1] * exns0[kinv - 1]);
        } else {
            rho0inv = ps0[kinv - 1] / (r * thv0top[kinv - 1] * exns0[kinv - 1]);
            sigmaw = sqrt(rkfre * tkeavg + epsvarw);
            if (zw->pertw) {
                /* zwshcu.F90:2282-2285 */
                tkeavg = tkeavg + zmh_ramp(W->tpert, 0.0, 2.0) * 2.0;
                sigmaw = sqrt(rkfre * tkeavg + epsvarw) + zmh_ramp(W->tpert, 0.0, 1.0) * 0.1;
            }
            wcrit = sqrt(2.0 * cin * rbuoy);
            if (zw->lf) {
                /* zwshcu.F90:2296-2300: ascent at the PBL-top layer lowers wcrit */
                tmp1 = dmin(dmax(-W->omega[kinv - 1], 0.0), 3.0);
                wcrit = wcrit - tmp1 / rho0inv / 9.8;
                wcrit = dmax(wcrit, 0.0);
            }
            wcrit = wcrit * zw->fwcrit;   /* zwshcu.F90:2303 */
            mu = wcrit / sigmaw / 1.4142;
            if (mu >= 3.0) { exit_code = EX_MU3; id_exit = 1; goto L333; }
        }
        PROBE(PRB(ORC_P_MUMIN2_APPLIED) = 0);
        cbmf = (rho0inv * sigmaw / 2.5066) * XEXP(-(mu * mu));
        {
            double cbmflimit = 0.9 * dp0[kinv - 1] / g / dt;
            mumin0 = 0.0;
            if (cbmf > cbmflimit) mumin0 = sqrt(-XLOG(2.5066 * cbmflimit / rho0inv / sigmaw));
        }
        mu = dmax(dmax(mu, mumin0), mumin1);
        mulcl = sqrt(2.0 * cinlcl * rbuoy) / 1.4142 / sigmaw;
        {
            double e1 = XEXP(-(mu * mu)) / 2.5066, ef = XERFC(mu);
            mulclstar = sqrt(dmax(0.0, 2.0 * (e1 * e1) * (1.0 / (ef * ef) - 0.25 / (rmaxfrac * rmaxfrac))));
        }
        if (mulcl > 1.e-8 && mulcl > mulclstar) {
            mumin2 = orc_compute_mumin2(mulcl, rmaxfrac, mu);
            if (mu > mumin2) { mumin2_fail = 1; }
            mu = dmax(mu, mumin2);
            if (mu == mumin2) limit_ufrc = 1;
            PROBE(PRB(ORC_P_MUMIN2_APPLIED) = (mu == mumin2) ? 1 : 0);
        }
        if (mu == mumin1) limit_ufrc = 1;
        cbmf = (rho0inv * sigmaw / 2.5066) * XEXP(-(mu * mu));
        winv = sigmaw * (2.0 / 2.5066) * XEXP(-(mu * mu)) / XERFC(mu);
        wtw = winv * winv - 2.0 * cinlcl * rbuoy;
        PROBE(PRB(ORC_P_CLOSURE) = 1; PRB(ORC_P_WCRIT) = wcrit; PRB(ORC_P_SIGMAW) = sigmaw; PRB(ORC_P_MU) = mu;
              PRB(ORC_P_RHO0INV) = rho0inv; PRB(ORC_P_CBMF) = cbmf; PRB(ORC_P_WINV) = winv;
              PRB(ORC_P_CINLCL) = cinlcl; PRB(ORC_P_RBUOY) = rbuoy; PRB(ORC_P_RMAXFRAC) = rmaxfrac;
              PRB(ORC_P_CBMFLIMIT) = 0.9 * dp0[kinv - 1] / g / dt; PRB(ORC_P_WLCL) = 0; PRB(ORC_P_UFRCLCL) = 0);
        if (wtw <= 0.0) { exit_code = EX_WTW; id_exit = 1; goto L333; }
        wlcl = sqrt(wtw);
        ufrclcl = cbmf / wlcl / rho0inv;
        PROBE(PRB(ORC_P_WLCL) = wlcl; PRB(ORC_P_UFRCLCL) = ufrclcl);
        wrel = wlcl;
        if (ufrclcl <= 0.0001) { exit_code = EX_UFRC; id_exit = 1; goto L333; }
        ufrc[krel - 1] = ufrclcl;
        for (k = kinv - 1; k <= krel - 1; ++k) { umf[k] = cbmf; wu[k] = winv; }

        /* ---- S10 release-level updraft (uwshcu.F90:2205-2255) ---- */
        emf[krel - 1] = 0.0;
        umf[krel - 1] = cbmf;
        wu[krel - 1] = wrel;
        thlu[krel - 1] = thlsrc;
        qtu[krel - 1] = qtsrc;
        orc_conden(c, prel, thlsrc, qtsrc, &thj, &qvj, &qlj, &qij, &qse, &id_check);
        if (id_check == 1) { exit_code = EX_CONDEN; id_exit = 1; goto L333; }
        thvu[krel - 1] = THV(thj, qvj, qlj, qij);
        uplus = 0.0;
        vplus = 0.0;
        if (krel == kinv) {
            uplus = PGFc * ssu0[kinv] * (prel - ps0[kinv - 1]);
            vplus = PGFc * ssv0[kinv] * (prel - ps0[kinv - 1]);
        } else {
            int kend = (krel - 1 > kinv) ? krel - 1 : kinv;
            for (k = kinv; k <= kend; ++k) {
                uplus = uplus + PGFc * ssu0[k] * (ps0[k] - ps0[k - 1]);
                vplus = vplus + PGFc * ssv0[k] * (ps0[k] - ps0[k - 1]);
            }
            uplus = uplus + PGFc * ssu0[krel] * (prel - ps0[krel - 1]);
            vplus = vplus + PGFc * ssv0[krel] * (prel - ps0[krel - 1]);
        }
        uu[krel - 1] = usrc + uplus;
        vu[krel - 1] = vsrc + vplus;
        for (m = 1; m <= ntr; ++m) W->tru[m][krel - 1] = trsrc[m];

        /* ---- S11 buoyancy sorting (uwshcu.F90:2296-2952) ---- */
        scaleh = tscaleh;
        if (tscaleh < 0.0) scaleh = 1000.0;
        for (iter_scaleh = 1; iter_scaleh <= 3; ++iter_scaleh) {
            PROBE(for (int f_ = 0; f_ < ORC_PL_NFIELD; ++f_) for (int k_ = 0; k_ < ORC_PRB_KMAX; ++k_) PRBL(f_, k_) = 0.0;
                  PRB(ORC_P_PPEN_SET) = 0; PRB(ORC_P_KRELP) = krel; PRB(ORC_P_PGFC) = PGFc; PRB(ORC_P_G) = g;
                  PRB(ORC_P_CRIQC) = criqc);
            kbup = krel;
            kpen = krel;
            wtw = wlcl * wlcl;
            pe = 0.5 * (prel + ps0[krel]);
            dpe = prel - ps0[krel];
            exne = exnf(c, pe);
            thvebot = thv0rel;
            thle = thl0[krel] + ssthl0[krel] * (pe - p0[krel]);
            qte = qt0[krel] + ssqt0[krel] * (pe - p0[krel]);
            ue = u0[krel] + ssu0[krel] * (pe - p0[krel]);
            ve = v0[krel] + ssv0[krel] * (pe - p0[krel]);
            for (m = 1; m <= ntr; ++m) tre[m] = W->tr0[m][krel] + W->sstr0[m][krel] * (pe - p0[krel]);

            for (k = krel; k <= mkx - 1; ++k) {
                km1 = k - 1;
                thlue = thlu[km1];
                qtue = qtu[km1];
                wue = wu[km1];
                wtwb = wtw;
                for (iter_xc = 1; iter_xc <= 2; ++iter_xc) {
                    wtw = wu[km1] * wu[km1];
                    orc_conden(c, pe, thle, qte, &thj, &qvj, &qlj, &qij, &qse, &id_check);
                    if (id_check == 1) { exit_code = EX_CONDEN; id_exit = 1; goto L333; }
                    thv0j = THV(thj, qvj, qlj, qij);
                    rho0j = pe / (r * thv0j * exne);
                    qsat_arg = thle * exne;
                    orc_qsat(c, qsat_arg, pe, &es, &qs, &gam);
                    excess0 = qte - qs;
                    orc_conden(c, pe, thlue, qtue, &thj, &qvj, &qlj, &qij, &qse, &id_check);
                    if (id_check == 1) { exit_code = EX_CONDEN; id_exit = 1; goto L333; }
                    if ((qlj + qij) > criqc) {
                        exql = ((qlj + qij) - criqc) * qlj / (qlj + qij);
                        exqi = ((qlj + qij) - criqc) * qij / (qlj + qij);
                        qtue = qtue - exql - exqi;
                        thlue = thlue + (xlv / cp / exne) * exql + (xls / cp / exne) * exqi;
                    }
                    orc_conden(c, pe, thlue, qtue, &thj, &qvj, &qlj, &qij, &qse, &id_check);
                    if (id_check == 1) { exit_code = EX_CONDEN; id_exit = 1; goto L333; }
                    thvj = THV(thj, qvj, qlj, qij);
                    tj = thj * exne;
                    qsat_arg = thlue * exne;
                    orc_qsat(c, qsat_arg, pe, &es, &qs, &gam);
                    excessu = qtue - qs;
                    cridis = rle * scaleh;
                    double xsat_p = 0.0, thvxsat_p = 0.0;   /* probe copies */
                    if ((excessu <= 0.0 && excess0 <= 0.0) || (excessu >= 0.0 && excess0 >= 0.0)) {
                        xc = dmin(1.0, dmax(0.0, 1.0 - 2.0 * rbuoy * g * cridis / (wue * wue) * (1.0 - thvj / thv0j)));
                    } else {
                        xsat = excessu / (excessu - excess0);
                        thlxsat = thlue + xsat * (thle - thlue);
                        qtxsat = qtue + xsat * (qte - qtue);
                        orc_conden(c, pe, thlxsat, qtxsat, &thj, &qvj, &qlj, &qij, &qse, &id_check);
                        if (id_check == 1) { exit_code = EX_CONDEN; id_exit = 1; goto L333; }
                        thvxsat = THV(thj, qvj, qlj, qij);
                        xsat_p = xsat;
                        thvxsat_p = thvxsat;
                        for (kk = 1; kk <= 2; ++kk) {
                            if (kk == 1) {
                                thv_x0 = thvj;
                                thv_x1 = (1.0 - 1.0 / xsat) * thvj + (1.0 / xsat) * thvxsat;
                            } else {
                                thv_x1 = thv0j;
                                thv_x0 = (xsat / (xsat - 1.0)) * thv0j + (1.0 / (1.0 - xsat)) * thvxsat;
                            }
                            aquad = wue * wue;
                            bquad = 2.0 * rbuoy * g * cridis * (thv_x1 - thv_x0) / thv0j - 2.0 * (wue * wue);
                            cquad = 2.0 * rbuoy * g * cridis * (thv_x0 - thv0j) / thv0j + (wue * wue);
                            if (kk == 1) {
                                if ((bquad * bquad - 4.0 * aquad * cquad) >= 0.0) {
                                    orc_roots(aquad, bquad, cquad, &xs1, &xs2, &status);
                                    x_cu = dmin(1.0, dmax(0.0, dmin(xsat, dmin(xs1, xs2))));
                                } else {
                                    x_cu = xsat;
                                }
                            } else {
                                if ((bquad * bquad - 4.0 * aquad * cquad) >= 0.0) {
                                    orc_roots(aquad, bquad, cquad, &xs1, &xs2, &status);
                                    x_en = dmin(1.0, dmax(0.0, dmax(xsat, dmin(xs1, xs2))));
                                } else {
                                    x_en = 1.0;
                                }
                            }
                        }
                        if (x_cu == xsat) xc = dmax(x_cu, x_en);
                        else xc = x_cu;
                    }
                    PROBE(int sat_ = !((excessu <= 0.0 && excess0 <= 0.0) || (excessu >= 0.0 && excess0 >= 0.0));
                          PRBL(ORC_PL_CASE, k) = sat_ ? 2 : 1; PRBL(ORC_PL_XC, k) = xc; PRBL(ORC_PL_WUE, k) = wue;
                          PRBL(ORC_PL_CRIDIS, k) = cridis; PRBL(ORC_PL_THVJ, k) = thvj; PRBL(ORC_PL_THV0J, k) = thv0j;
                          PRBL(ORC_PL_XSAT, k) = xsat_p; PRBL(ORC_PL_THVXSAT, k) = thvxsat_p;
                          PRBL(ORC_PL_THLUE, k) = thlue; PRBL(ORC_PL_QTUE, k) = qtue; PRBL(ORC_PL_EXCESSU, k) = excessu;
                          PRBL(ORC_PL_EXCESS0, k) = excess0);
                    ee2 = xc * xc;
                    ud2 = 1.0 - 2.0 * xc + xc * xc;
                    rei[k] = (0.5 * rkm / z0[k] / g / rho0j);
                    if (xc > 0.5)
                        rei[k] = dmin(rei[k], 0.9 * XLOG(dp0[k] / g / dt / umf[km1] + 1.0) / dpe / (2.0 * xc - 1.0));
                    fer[k] = rei[k] * ee2;
                    fdr[k] = rei[k] * ud2;
                    if (zw && zw->ent) {
                        /* zwshcu.F90:2818-2844: entrainment from the updraft buoyancy */
                        double fer_min;
                        bogbot = rbuoy * (thvu[km1] / thvebot - 1.0) * g;
                        fer[k] = zw->ent0 * bogbot / dmax(wue * wue, 1.0e-4);
                        fer_min = 0.0;
                        if (bogbot > 1.e-4) fer_min = 1.e-4;
                        fer[k] = dmin(dmax(fer[k], fer_min), 3.0e-2);
                    }
                    umf[k] = umf[km1] * XEXP(dpe * (fer[k] - fdr[k]));
                    emf[k] = 0.0;
                    if (fer[k] * dpe < 1.e-4) {
                        thlu[k] = thlu[km1] + (thle + ssthl0[k] * dpe / 2.0 - thlu[km1]) * fer[k] * dpe;
                        qtu[k] = qtu[km1] + (qte + ssqt0[k] * dpe / 2.0 - qtu[km1]) * fer[k] * dpe;
                        uu[k] = uu[km1] + (ue + ssu0[k] * dpe / 2.0 - uu[km1]) * fer[k] * dpe - PGFc * ssu0[k] * dpe;
                        vu[k] = vu[km1] + (ve + ssv0[k] * dpe / 2.0 - vu[km1]) * fer[k] * dpe - PGFc * ssv0[k] * dpe;
                        for (m = 1; m <= ntr; ++m)
                            W->tru[m][k] = W->tru[m][km1] +
                                           (tre[m] + W->sstr0[m][k] * dpe / 2.0 - W->tru[m][km1]) * fer[k] * dpe;
                    } else {
                        double ef = XEXP(-fer[k] * dpe);
                        thlu[k] = (thle + ssthl0[k] / fer[k] - ssthl0[k] * dpe / 2.0) -
                                  (thle + ssthl0[k] * dpe / 2.0 - thlu[km1] + ssthl0[k] / fer[k]) * ef;
                        qtu[k] = (qte + ssqt0[k] / fer[k] - ssqt0[k] * dpe / 2.0) -
                                 (qte + ssqt0[k] * dpe / 2.0 - qtu[km1] + ssqt0[k] / fer[k]) * ef;
                        uu[k] = (ue + (1.0 - PGFc) * ssu0[k] / fer[k] - ssu0[k] * dpe / 2.0) -
                                (ue + ssu0[k] * dpe / 2.0 - uu[km1] + (1.0 - PGFc) * ssu0[k] / fer[k]) * ef;
                        vu[k] = (ve + (1.0 - PGFc) * ssv0[k] / fer[k] - ssv0[k] * dpe / 2.0) -
                                (ve + ssv0[k] * dpe / 2.0 - vu[km1] + (1.0 - PGFc) * ssv0[k] / fer[k]) * ef;
                        for (m = 1; m <= ntr; ++m)
                            W->tru[m][k] = (tre[m] + W->sstr0[m][k] / fer[k] - W->sstr0[m][k] * dpe / 2.0) -
                                           (tre[m] + W->sstr0[m][k] * dpe / 2.0 - W->tru[m][km1] +
                                            W->sstr0[m][k] / fer[k]) * ef;
                    }
                    PROBE(PRBL(ORC_PL_SET, k) = 1; PRBL(ORC_PL_THLE, k) = thle; PRBL(ORC_PL_QTE, k) = qte;
                          PRBL(ORC_PL_UE, k) = ue; PRBL(ORC_PL_VE, k) = ve; PRBL(ORC_PL_PE, k) = pe;
                          PRBL(ORC_PL_DPE, k) = dpe; PRBL(ORC_PL_FER, k) = fer[k]; PRBL(ORC_PL_FDR, k) = fdr[k];
                          PRBL(ORC_PL_THLU0, k) = thlu[km1]; PRBL(ORC_PL_QTU0, k) = qtu[km1];
                          PRBL(ORC_PL_UU0, k) = uu[km1]; PRBL(ORC_PL_VU0, k) = vu[km1];
                          PRBL(ORC_PL_THLU1, k) = thlu[k]; PRBL(ORC_PL_QTU1, k) = qtu[k];
                          PRBL(ORC_PL_UU1, k) = uu[k]; PRBL(ORC_PL_VU1, k) = vu[k];
                          PRBL(ORC_PL_SSTHL, k) = ssthl0[k]; PRBL(ORC_PL_SSQT, k) = ssqt0[k];
                          PRBL(ORC_PL_SSU, k) = ssu0[k]; PRBL(ORC_PL_SSV, k) = ssv0[k];
                          PRBL(ORC_PL_UMF0, k) = umf[km1]; PRBL(ORC_PL_UMF1, k) = umf[k]; PRBL(ORC_PL_REI, k) = rei[k]);
                    orc_conden(c, ps0[k], thlu[k], qtu[k], &thj, &qvj, &qlj, &qij, &qse, &id_check);
                    if (id_check == 1) { exit_code = EX_CONDEN; id_exit = 1; goto L333; }
                    if ((qlj + qij) > criqc) {
                        exql = ((qlj + qij) - criqc) * qlj / (qlj + qij);
                        exqi = ((qlj + qij) - criqc) * qij / (qlj + qij);
                        qtu[k] = qtu[k] - exql - exqi;
                        thlu[k] = thlu[k] + (xlv / cp / exns0[k]) * exql + (xls / cp / exns0[k]) * exqi;
                        dwten[k] = exql;
                        diten[k] = exqi;
                    } else {
                        dwten[k] = 0.0;
                        diten[k] = 0.0;
                    }
                    orc_conden(c, ps0[k], thlu[k], qtu[k], &thj, &qvj, &qlj, &qij, &qse, &id_check);
                    if (id_check == 1) { exit_code = EX_CONDEN; id_exit = 1; goto L333; }
                    thvu[k] = THV(thj, qvj, qlj, qij);
                    bogbot = rbuoy * (thvu[km1] / thvebot - 1.0);
                    bogtop = rbuoy * (thvu[k] / thv0top[k] - 1.0);
                    delbog = bogtop - bogbot;
                    drage = fer[k] * (1.0 + rdrag);
                    expfac = XEXP(-2.0 * drage * dpe);
                    wtwb = wtw;
                    if (drage * dpe > 1.e-3)
                        wtw = wtw * expfac + (delbog + (1.0 - expfac) * (bogbot + delbog / (-2.0 * drage * dpe))) /
                                                 (rho0j * drage);
                    else
                        wtw = wtw + dpe * (bogbot + bogtop) / rho0j;
                    PROBE(PRBL(ORC_PL_WTWB, k) = wtwb; PRBL(ORC_PL_WTW, k) = wtw; PRBL(ORC_PL_DRAGE, k) = drage;
                          PRBL(ORC_PL_BOGBOT, k) = bogbot; PRBL(ORC_PL_BOGTOP, k) = bogtop;
                          PRBL(ORC_PL_RHO0J, k) = rho0j);
                    if (wtw > 0.0) {
                        thlue = 0.5 * (thlu[km1] + thlu[k]);
                        qtue = 0.5 * (qtu[km1] + qtu[k]);
                        wue = 0.5 * sqrt(dmax(wtwb + wtw, 0.0));
                    } else {
                        break;  /* go to 111 */
                    }
                }
                /* 111 continue */
                if (umf[k] == 0.0) wtw = -1.0;
                if (bogtop > 0.0 && wtw > 0.0) kbup = k;
                if (wtw <= 0.0) { kpen = k; break; /* go to 45 */ }
                wu[k] = sqrt(wtw);
                if (wu[k] > 100.0) { exit_code = EX_WU; id_exit = 1; goto L333; }
                rhos0j = ps0[k] / (r * 0.5 * (thv0bot[k + 1] + thv0top[k]) * exns0[k]);
                ufrc[k] = umf[k] / (rhos0j * wu[k]);
                if (ufrc[k] > rmaxfrac) {
                    limit_ufrc = 1;
                    ufrc[k] = rmaxfrac;
                    umf[k] = rmaxfrac * rhos0j * wu[k];
                    fdr[k] = fer[k] - XLOG(umf[k] / umf[km1]) / dpe;
                }
                pe = p0[k + 1];
                dpe = dp0[k + 1];
                exne = exn0[k + 1];
                thvebot = thv0bot[k + 1];
                thle = thl0[k + 1];
                qte = qt0[k + 1];
                ue = u0[k + 1];
                ve = v0[k + 1];
                for (m = 1; m <= ntr; ++m) tre[m] = W->tr0[m][k + 1];
            }
            /* 45 continue */
            if (drage == 0.0) {
                aquad = (bogtop - bogbot) / (ps0[kpen] - ps0[kpen - 1]);
                bquad = 2.0 * bogbot;
                cquad = -(wu[kpen - 1] * wu[kpen - 1]) * rho0j;
                orc_roots(aquad, bquad, cquad, &xc1, &xc2, &status);
                if (status == 0) {
                    if (xc1 <= 0.0 && xc2 <= 0.0) {
                        ppen = dmax(xc1, xc2);
                        ppen = dmin(0.0, dmax(-dp0[kpen], ppen));
                    } else if (xc1 > 0.0 && xc2 > 0.0) {
                        ppen = -dp0[kpen];
                    } else {
                        ppen = dmin(xc1, xc2);
                        ppen = dmin(0.0, dmax(-dp0[kpen], ppen));
                    }
                } else {
                    ppen = -dp0[kpen];
                }
            } else {
                ppen = orc_compute_ppen(wtwb, drage, bogbot, bogtop, rho0j, dp0[kpen]);
            }
            PROBE(PRB(ORC_P_PPEN_SET) = 1; PRB(ORC_P_WTWB) = wtwb; PRB(ORC_P_DRAGE) = drage; PRB(ORC_P_BOGBOT) = bogbot;
                  PRB(ORC_P_BOGTOP) = bogtop; PRB(ORC_P_RHO0J) = rho0j; PRB(ORC_P_DPEN) = dp0[kpen];
                  PRB(ORC_P_PPEN) = ppen; PRB(ORC_P_KPEN) = kpen);
            if (ppen == -dp0[kpen] || ppen == 0.0) limit_ppen = 1;
            if (fer[kpen] * (-ppen) < 1.e-4) {
                thlu_top = thlu[kpen - 1] +
                           (thl0[kpen] + ssthl0[kpen] * (-ppen) / 2.0 - thlu[kpen - 1]) * fer[kpen] * (-ppen);
                qtu_top = qtu[kpen - 1] + (qt0[kpen] + ssqt0[kpen] * (-ppen) / 2.0 - qtu[kpen - 1]) * fer[kpen] * (-ppen);
            } else {
                double ef = XEXP(-fer[kpen] * (-ppen));
                thlu_top = (thl0[kpen] + ssthl0[kpen] / fer[kpen] - ssthl0[kpen] * (-ppen) / 2.0) -
                           (thl0[kpen] + ssthl0[kpen] * (-ppen) / 2.0 - thlu[kpen - 1] + ssthl0[kpen] / fer[kpen]) * ef;
                qtu_top = (qt0[kpen] + ssqt0[kpen] / fer[kpen] - ssqt0[kpen] * (-ppen) / 2.0) -
                          (qt0[kpen] + ssqt0[kpen] * (-ppen) / 2.0 - qtu[kpen - 1] + ssqt0[kpen] / fer[kpen]) * ef;
            }
            orc_conden(c, ps0[kpen - 1] + ppen, thlu_top, qtu_top, &thj, &qvj, &qlj, &qij, &qse, &id_check);
            if (id_check == 1) { exit_code = EX_CONDEN; id_exit = 1; goto L333; }
            exntop = XPOW((ps0[kpen - 1] + ppen) / c->p00, c->rovcp);
            if ((qlj + qij) > criqc) {
                dwten[kpen] = ((qlj + qij) - criqc) * qlj / (qlj + qij);
                diten[kpen] = ((qlj + qij) - criqc) * qij / (qlj + qij);
                qtu_top = qtu_top - dwten[kpen] - diten[kpen];
                thlu_top = thlu_top + (xlv / cp / exntop) * dwten[kpen] + (xls / cp / exntop) * diten[kpen];
            } else {
                dwten[kpen] = 0.0;
                diten[kpen] = 0.0;
            }
            rhos0j = ps0[kpen - 1] / (r * 0.5 * (thv0bot[kpen] + thv0top[kpen - 1]) * exns0[kpen - 1]);
            cush = zs0[kpen - 1] - ppen / rhos0j / g;
            scaleh = cush;
        }

        /* ---- S12 forcedCu filter (uwshcu.F90:2964-2992) ---- */
        forcedCu = (kbup == krel);
        if (forcedCu) { exit_code = EX_CUFILTER; id_exit = 1; goto L333; }

        for (k = kpen; k <= mkx; ++k) { umf[k] = 0.0; emf[k] = 0.0; ufrc[k] = 0.0; }
        for (k = kpen + 1; k <= mkx; ++k) { dwten[k] = 0.0; diten[k] = 0.0; fer[k] = 0.0; fdr[k] = 0.0; }

        /* ---- S13 penetrative entrainment (uwshcu.F90:3055-3183) ---- */
        for (k = 0; k <= mkx; ++k) {
            thlu_emf[k] = thlu[k];
            qtu_emf[k] = qtu[k];
            uu_emf[k] = uu[k];
            vu_emf[k] = vu[k];
            for (m = 1; m <= ntr; ++m) W->tru_emf[m][k] = W->tru[m][k];
        }
        for (k = kpen - 1; k >= kbup; --k) {
            rhos0j = ps0[k] / (r * 0.5 * (thv0bot[k + 1] + thv0top[k]) * exns0[k]);
            if (k == kpen - 1) {
                if ((umf[k] * ppen * rei[kpen] * rpen) < -0.1 * rhos0j) limit_emf = 1;
                if ((umf[k] * ppen * rei[kpen] * rpen) < -0.9 * dp0[kpen] / g / dt) limit_emf = 1;
                emf[k] = dmax(dmax(umf[k] * ppen * rei[kpen] * rpen, -0.1 * rhos0j), -0.9 * dp0[kpen] / g / dt);
                thlu_emf[k] = thl0[kpen] + ssthl0[kpen] * (ps0[k] - p0[kpen]);
                qtu_emf[k] = qt0[kpen] + ssqt0[kpen] * (ps0[k] - p0[kpen]);
                uu_emf[k] = u0[kpen] + ssu0[kpen] * (ps0[k] - p0[kpen]);
                vu_emf[k] = v0[kpen] + ssv0[kpen] * (ps0[k] - p0[kpen]);
                for (m = 1; m <= ntr; ++m)
                    W->tru_emf[m][k] = W->tr0[m][kpen] + W->sstr0[m][kpen] * (ps0[k] - p0[kpen]);
            } else {
                if ((emf[k + 1] - umf[k] * dp0[k + 1] * rei[k + 1] * rpen) < -0.1 * rhos0j) limit_emf = 1;
                if ((emf[k + 1] - umf[k] * dp0[k + 1] * rei[k + 1] * rpen) < -0.9 * dp0[k + 1] / g / dt) limit_emf = 1;
                emf[k] = dmax(dmax(emf[k + 1] - umf[k] * dp0[k + 1] * rei[k + 1] * rpen, -0.1 * rhos0j),
                              -0.9 * dp0[k + 1] / g / dt);
                if (fabs(emf[k]) > fabs(emf[k + 1])) {
                    thlu_emf[k] = (thlu_emf[k + 1] * emf[k + 1] + thl0[k + 1] * (emf[k] - emf[k + 1])) / emf[k];
                    qtu_emf[k] = (qtu_emf[k + 1] * emf[k + 1] + qt0[k + 1] * (emf[k] - emf[k + 1])) / emf[k];
                    uu_emf[k] = (uu_emf[k + 1] * emf[k + 1] + u0[k + 1] * (emf[k] - emf[k + 1])) / emf[k];
                    vu_emf[k] = (vu_emf[k + 1] * emf[k + 1] + v0[k + 1] * (emf[k] - emf[k + 1])) / emf[k];
                    for (m = 1; m <= ntr; ++m)
                        W->tru_emf[m][k] =
                            (W->tru_emf[m][k + 1] * emf[k + 1] + W->tr0[m][k + 1] * (emf[k] - emf[k + 1])) / emf[k];
                } else {
                    thlu_emf[k] = thl0[k + 1];
                    qtu_emf[k] = qt0[k + 1];
                    uu_emf[k] = u0[k + 1];
                    vu_emf[k] = v0[k + 1];
                    for (m = 1; m <= ntr; ++m) W->tru_emf[m][k] = W->tr0[m][k + 1];
                }
            }
        }

        PROBE(PRB(ORC_P_KBUP) = kbup; PRB(ORC_P_KPEN) = kpen; PRB(ORC_P_RPEN) = rpen; PRB(ORC_P_EMF_SET) = 1;
              PRB(ORC_P_LIMIT_EMF) = limit_emf;
              for (int k_ = 0; k_ <= mkx && k_ < ORC_PRB_KMAX; ++k_) {
                  PRBL(ORC_PL_EMF, k_) = emf[k_]; PRBL(ORC_PL_THLU_EMF, k_) = thlu_emf[k_];
                  PRBL(ORC_PL_QTU_EMF, k_) = qtu_emf[k_]; PRBL(ORC_PL_THL0, k_) = k_ >= 1 ? thl0[k_] : 0.0;
                  PRBL(ORC_PL_QT0, k_) = k_ >= 1 ? qt0[k_] : 0.0; PRBL(ORC_PL_PS0, k_) = ps0[k_];
                  PRBL(ORC_PL_P0, k_) = k_ >= 1 ? p0[k_] : 0.0; PRBL(ORC_PL_SSTHL0, k_) = k_ >= 1 ? ssthl0[k_] : 0.0;
                  PRBL(ORC_PL_SSQT0, k_) = k_ >= 1 ? ssqt0[k_] : 0.0; PRBL(ORC_PL_UMF, k_) = umf[k_];
                  PRBL(ORC_PL_REIF, k_) = rei[k_]; PRBL(ORC_PL_DP0, k_) = k_ >= 1 ? dp0[k_] : 0.0;
              });

        /* ---- S14 fluxes (uwshcu.F90:3255-3401) ---- */
        xsrc = qtsrc;
        xmean = qt0[kinv];
        xtop = qt0[kinv + 1] + ssqt0[kinv + 1] * (ps0[kinv] - p0[kinv + 1]);
        xbot = qt0[kinv - 1] + ssqt0[kinv - 1] * (ps0[kinv - 1] - p0[kinv - 1]);
        orc_fluxbelowinv(c, cbmf, ps0, mkx, kinv, dt, xsrc, xmean, xtop, xbot, xflx);
        for (k = 0; k <= kinv - 1; ++k) qtflx[k] = xflx[k];

        xsrc = thlsrc;
        xmean = thl0[kinv];
        xtop = thl0[kinv + 1] + ssthl0[kinv + 1] * (ps0[kinv] - p0[kinv + 1]);
        xbot = thl0[kinv - 1] + ssthl0[kinv - 1] * (ps0[kinv - 1] - p0[kinv - 1]);
        orc_fluxbelowinv(c, cbmf, ps0, mkx, kinv, dt, xsrc, xmean, xtop, xbot, xflx);
        for (k = 0; k <= kinv - 1; ++k) slflx[k] = cp * exns0[k] * xflx[k];

        xsrc = usrc;
        xmean = u0[kinv];
        xtop = u0[kinv + 1] + ssu0[kinv + 1] * (ps0[kinv] - p0[kinv + 1]);
        xbot = u0[kinv - 1] + ssu0[kinv - 1] * (ps0[kinv - 1] - p0[kinv - 1]);
        orc_fluxbelowinv(c, cbmf, ps0, mkx, kinv, dt, xsrc, xmean, xtop, xbot, xflx);
        for (k = 0; k <= kinv - 1; ++k) uflx[k] = xflx[k];

        xsrc = vsrc;
        xmean = v0[kinv];
        xtop = v0[kinv + 1] + ssv0[kinv + 1] * (ps0[kinv] - p0[kinv + 1]);
        xbot = v0[kinv - 1] + ssv0[kinv - 1] * (ps0[kinv - 1] - p0[kinv - 1]);
        orc_fluxbelowinv(c, cbmf, ps0, mkx, kinv, dt, xsrc, xmean, xtop, xbot, xflx);
        for (k = 0; k <= kinv - 1; ++k) vflx[k] = xflx[k];

        for (m = 1; m <= ntr; ++m) {
            xsrc = trsrc[m];
            xmean = W->tr0[m][kinv];
            xtop = W->tr0[m][kinv + 1] + W->sstr0[m][kinv + 1] * (ps0[kinv] - p0[kinv + 1]);
            xbot = W->tr0[m][kinv - 1] + W->sstr0[m][kinv - 1] * (ps0[kinv - 1] - p0[kinv - 1]);
            orc_fluxbelowinv(c, cbmf, ps0, mkx, kinv, dt, xsrc, xmean, xtop, xbot, xflx);
            for (k = 0; k <= kinv - 1; ++k) W->trflx[m][k] = xflx[k];
        }

        uplus = 0.0;
        vplus = 0.0;
        for (k = kinv; k <= krel - 1; ++k) {
            kp1 = k + 1;
            qtflx[k] = cbmf * (qtsrc - (qt0[kp1] + ssqt0[kp1] * (ps0[k] - p0[kp1])));
            slflx[k] = cbmf * (thlsrc - (thl0[kp1] + ssthl0[kp1] * (ps0[k] - p0[kp1]))) * cp * exns0[k];
            uplus = uplus + PGFc * ssu0[k] * (ps0[k] - ps0[k - 1]);
            vplus = vplus + PGFc * ssv0[k] * (ps0[k] - ps0[k - 1]);
            uflx[k] = cbmf * (usrc + uplus - (u0[kp1] + ssu0[kp1] * (ps0[k] - p0[kp1])));
            vflx[k] = cbmf * (vsrc + vplus - (v0[kp1] + ssv0[kp1] * (ps0[k] - p0[kp1])));
            for (m = 1; m <= ntr; ++m)
                W->trflx[m][k] = cbmf * (trsrc[m] - (W->tr0[m][kp1] + W->sstr0[m][kp1] * (ps0[k] - p0[kp1])));
        }
        for (k = krel; k <= kbup - 1; ++k) {
            kp1 = k + 1;
            slflx[k] = cp * exns0[k] * umf[k] * (thlu[k] - (thl0[kp1] + ssthl0[kp1] * (ps0[k] - p0[kp1])));
            qtflx[k] = umf[k] * (qtu[k] - (qt0[kp1] + ssqt0[kp1] * (ps0[k] - p0[kp1])));
            uflx[k] = umf[k] * (uu[k] - (u0[kp1] + ssu0[kp1] * (ps0[k] - p0[kp1])));
            vflx[k] = umf[k] * (vu[k] - (v0[kp1] + ssv0[kp1] * (ps0[k] - p0[kp1])));
            for (m = 1; m <= ntr; ++m)
                W->trflx[m][k] = umf[k] * (W->tru[m][k] - (W->tr0[m][kp1] + W->sstr0[m][kp1] * (ps0[k] - p0[kp1])));
        }
        for (k = kbup; k <= kpen - 1; ++k) {
            slflx[k] = cp * exns0[k] * emf[k] * (thlu_emf[k] - (thl0[k] + ssthl0[k] * (ps0[k] - p0[k])));
            qtflx[k] = emf[k] * (qtu_emf[k] - (qt0[k] + ssqt0[k] * (ps0[k] - p0[k])));
            uflx[k] = emf[k] * (uu_emf[k] - (u0[k] + ssu0[k] * (ps0[k] - p0[k])));
            vflx[k] = emf[k] * (vu_emf[k] - (v0[k] + ssv0[k] * (ps0[k] - p0[k])));
            for (m = 1; m <= ntr; ++m)
                W->trflx[m][k] = emf[k] * (W->tru_emf[m][k] - (W->tr0[m][k] + W->sstr0[m][k] * (ps0[k] - p0[k])));
        }

        /* ---- S15 subsidence condensate sinks (uwshcu.F90:3416-3476) ---- */
        for (k = 0; k <= mkx; ++k) uemf[k] = 0.0;
        for (k = 0; k <= kinv - 2; ++k) uemf[k] = cbmf * (ps0[0] - ps0[k]) / (ps0[0] - ps0[kinv - 1]);
        for (k = kinv - 1; k <= krel - 1; ++k) uemf[k] = cbmf;
        for (k = krel; k <= kbup - 1; ++k) uemf[k] = umf[k];
        for (k = kbup; k <= kpen - 1; ++k) uemf[k] = emf[k];
        for (k = 1; k <= mkx; ++k) comsub[k] = 0.0;
        for (k = 1; k <= kpen; ++k) comsub[k] = 0.5 * (uemf[k] + uemf[k - 1]);
        for (k = 1; k <= kpen; ++k) {
            if (comsub[k] >= 0.0) {
                if (k == mkx) {
                    thlten_sub = qtten_sub = qlten_sub = qiten_sub = nlten_sub = niten_sub = 0.0;
                } else {
                    thlten_sub = g * comsub[k] * (thl0[k + 1] - thl0[k]) / (p0[k] - p0[k + 1]);
                    qtten_sub = g * comsub[k] * (qt0[k + 1] - qt0[k]) / (p0[k] - p0[k + 1]);
                    qlten_sub = g * comsub[k] * (ql0[k + 1] - ql0[k]) / (p0[k] - p0[k + 1]);
                    qiten_sub = g * comsub[k] * (qi0[k + 1] - qi0[k]) / (p0[k] - p0[k + 1]);
                    nlten_sub = g * comsub[k] * (W->tr0[ixnumliq][k + 1] - W->tr0[ixnumliq][k]) / (p0[k] - p0[k + 1]);
                    niten_sub = g * comsub[k] * (W->tr0[ixnumice][k + 1] - W->tr0[ixnumice][k]) / (p0[k] - p0[k + 1]);
                }
            } else {
                if (k == 1) {
                    thlten_sub = qtten_sub = qlten_sub = qiten_sub = nlten_sub = niten_sub = 0.0;
                } else {
                    thlten_sub = g * comsub[k] * (thl0[k] - thl0[k - 1]) / (p0[k - 1] - p0[k]);
                    qtten_sub = g * comsub[k] * (qt0[k] - qt0[k - 1]) / (p0[k - 1] - p0[k]);
                    qlten_sub = g * comsub[k] * (ql0[k] - ql0[k - 1]) / (p0[k - 1] - p0[k]);
                    qiten_sub = g * comsub[k] * (qi0[k] - qi0[k - 1]) / (p0[k - 1] - p0[k]);
                    nlten_sub = g * comsub[k] * (W->tr0[ixnumliq][k] - W->tr0[ixnumliq][k - 1]) / (p0[k - 1] - p0[k]);
                    niten_sub = g * comsub[k] * (W->tr0[ixnumice][k] - W->tr0[ixnumice][k - 1]) / (p0[k - 1] - p0[k]);
                }
            }
            thl_prog = thl0[k] + thlten_sub * dt;
            qt_prog = dmax(qt0[k] + qtten_sub * dt, 1.e-12);
            orc_conden(c, p0[k], thl_prog, qt_prog, &thj, &qvj, &qlj, &qij, &qse, &id_check);
            if (id_check == 1) { exit_code = EX_SUBCONDEN; id_exit = 1; goto L333; }
            qlten_sink[k] = dmax(qlten_sub, -ql0[k] / dt);
            qiten_sink[k] = dmax(qiten_sub, -qi0[k] / dt);
            nlten_sink[k] = dmax(nlten_sub, -W->tr0[ixnumliq][k] / dt);
            niten_sink[k] = dmax(niten_sub, -W->tr0[ixnumice][k] / dt);
        }

        /* ---- S16 tendencies (uwshcu.F90:3488-3794) ---- */
        for (k = 1; k <= kpen; ++k) {
            km1 = k - 1;
            uten[k] = (uflx[km1] - uflx[k]) * g / dp0[k];
            vten[k] = (vflx[km1] - vflx[k]) * g / dp0[k];
            uf[k] = u0[k] + uten[k] * dt;
            vf[k] = v0[k] + vten[k] * dt;
        }
        rliq = 0.0;
        rainflx = 0.0;
        snowflx = 0.0;
        for (k = 1; k <= kpen; ++k) {
            km1 = k - 1;
            dwten[k] = dwten[k] * 0.5 * (umf[k - 1] + umf[k]) * g / dp0[k];
            diten[k] = diten[k] * 0.5 * (umf[k - 1] + umf[k]) * g / dp0[k];
            qrten[k] = frc_rasn * dwten[k];
            qsten[k] = frc_rasn * diten[k];
            rainflx = rainflx + qrten[k] * dp0[k] / g;
            snowflx = snowflx + qsten[k] * dp0[k] / g;
            slten[k] = (slflx[km1] - slflx[k]) * g / dp0[k];
            if (k == 1) {
                slten[k] = slten[k] - g / 4.0 / dp0[k] *
                                          (uflx[k] * (uf[k + 1] - uf[k] + u0[k + 1] - u0[k]) +
                                           vflx[k] * (vf[k + 1] - vf[k] + v0[k + 1] - v0[k]));
            } else if (k >= 2 && k <= kpen - 1) {
                slten[k] = slten[k] - g / 4.0 / dp0[k] *
                                          (uflx[k] * (uf[k + 1] - uf[k] + u0[k + 1] - u0[k]) +
                                           uflx[k - 1] * (uf[k] - uf[k - 1] + u0[k] - u0[k - 1]) +
                                           vflx[k] * (vf[k + 1] - vf[k] + v0[k + 1] - v0[k]) +
                                           vflx[k - 1] * (vf[k] - vf[k - 1] + v0[k] - v0[k - 1]));
            } else if (k == kpen) {
                slten[k] = slten[k] - g / 4.0 / dp0[k] *
                                          (uflx[k - 1] * (uf[k] - uf[k - 1] + u0[k] - u0[k - 1]) +
                                           vflx[k - 1] * (vf[k] - vf[k - 1] + v0[k] - v0[k - 1]));
            }
            qtten[k] = (qtflx[km1] - qtflx[k]) * g / dp0[k];

            if (k < krel || k > kpen) {
                qlu_mid = 0.0; qiu_mid = 0.0; qlj = 0.0; qij = 0.0;
            } else if (k == krel) {
                orc_conden(c, prel, thlu[krel - 1], qtu[krel - 1], &thj, &qvj, &qlj, &qij, &qse, &id_check);
                if (id_check == 1) { exit_code = EX_CONDEN; id_exit = 1; goto L333; }
                qlubelow = qlj;
                qiubelow = qij;
                orc_conden(c, ps0[k], thlu[k], qtu[k], &thj, &qvj, &qlj, &qij, &qse, &id_check);
                if (id_check == 1) { exit_code = EX_CONDEN; id_exit = 1; goto L333; }
                qlu_mid = 0.5 * (qlubelow + qlj) * (prel - ps0[k]) / (ps0[k - 1] - ps0[k]);
                qiu_mid = 0.5 * (qiubelow + qij) * (prel - ps0[k]) / (ps0[k - 1] - ps0[k]);
            } else if (k == kpen) {
                orc_conden(c, ps0[k - 1] + ppen, thlu_top, qtu_top, &thj, &qvj, &qlj, &qij, &qse, &id_check);
                if (id_check == 1) { exit_code = EX_CONDEN; id_exit = 1; goto L333; }
                qlu_mid = 0.5 * (qlubelow + qlj) * (-ppen) / (ps0[k - 1] - ps0[k]);
                qiu_mid = 0.5 * (qiubelow + qij) * (-ppen) / (ps0[k - 1] - ps0[k]);
                qlu_top = qlj;
                qiu_top = qij;
            } else {
                orc_conden(c, ps0[k], thlu[k], qtu[k], &thj, &qvj, &qlj, &qij, &qse, &id_check);
                if (id_check == 1) { exit_code = EX_CONDEN; id_exit = 1; goto L333; }
                qlu_mid = 0.5 * (qlubelow + qlj);
                qiu_mid = 0.5 * (qiubelow + qij);
            }
            qlubelow = qlj;
            qiubelow = qij;

            qc_l[k] = (1.0 - frc_rasn) * dwten[k];
            qc_i[k] = (1.0 - frc_rasn) * diten[k];
            if (k <= kbup) {
                qc_l[k] = qc_l[k] + g * 0.5 * (umf[k - 1] + umf[k]) * fdr[k] * qlu_mid;
                qc_i[k] = qc_i[k] + g * 0.5 * (umf[k - 1] + umf[k]) * fdr[k] * qiu_mid;
                qc_lm = -g * 0.5 * (umf[k - 1] + umf[k]) * fdr[k] * ql0[k];
                qc_im = -g * 0.5 * (umf[k - 1] + umf[k]) * fdr[k] * qi0[k];
                nc_lm = -g * 0.5 * (umf[k - 1] + umf[k]) * fdr[k] * W->tr0[ixnumliq][k];
                nc_im = -g * 0.5 * (umf[k - 1] + umf[k]) * fdr[k] * W->tr0[ixnumice][k];
            } else {
                qc_lm = 0.0; qc_im = 0.0; nc_lm = 0.0; nc_im = 0.0;
            }
            if (k == kbup) {
                qc_l[k] = qc_l[k] + g * umf[k] * qlj / (ps0[k - 1] - ps0[k]);
                qc_i[k] = qc_i[k] + g * umf[k] * qij / (ps0[k - 1] - ps0[k]);
                qc_lm = qc_lm - g * umf[k] * ql0[k] / (ps0[k - 1] - ps0[k]);
                qc_im = qc_im - g * umf[k] * qi0[k] / (ps0[k - 1] - ps0[k]);
                nc_lm = nc_lm - g * umf[k] * W->tr0[ixnumliq][k] / (ps0[k - 1] - ps0[k]);
                nc_im = nc_im - g * umf[k] * W->tr0[ixnumice][k] / (ps0[k - 1] - ps0[k]);
            }
            if (k == kbup) {
                orc_conden(c, p0[k], thlu_emf[k], qtu_emf[k], &thj, &qvj, &ql_emf_kbup, &qi_emf_kbup, &qse, &id_check);
                if (id_check == 1) { exit_code = EX_KBUPCONDEN; id_exit = 1; goto L333; }
                nl_emf_kbup = (ql_emf_kbup > 0.0) ? W->tru_emf[ixnumliq][k] : 0.0;
                ni_emf_kbup = (qi_emf_kbup > 0.0) ? W->tru_emf[ixnumice][k] : 0.0;
                qc_lm = qc_lm - g * emf[k] * (ql_emf_kbup - ql0[k]) / (ps0[k - 1] - ps0[k]);
                qc_im = qc_im - g * emf[k] * (qi_emf_kbup - qi0[k]) / (ps0[k - 1] - ps0[k]);
                nc_lm = nc_lm - g * emf[k] * (nl_emf_kbup - W->tr0[ixnumliq][k]) / (ps0[k - 1] - ps0[k]);
                nc_im = nc_im - g * emf[k] * (ni_emf_kbup - W->tr0[ixnumice][k]) / (ps0[k - 1] - ps0[k]);
            }
            qlten[k] = qc_l[k] + frc_rasn * dwten[k] + (dmax(0.0, ql0[k] + (qc_lm + qlten_sink[k]) * dt) - ql0[k]) / dt;
            qiten[k] = qc_i[k] + frc_rasn * diten[k] + (dmax(0.0, qi0[k] + (qc_im + qiten_sink[k]) * dt) - qi0[k]) / dt;
            W->trten[ixnumliq][k] = dmax(nc_lm + nlten_sink[k], -W->tr0[ixnumliq][k] / dt);
            W->trten[ixnumice][k] = dmax(nc_im + niten_sink[k], -W->tr0[ixnumice][k] / dt);
            qvten[k] = qtten[k] - qlten[k] - qiten[k];
            sten[k] = slten[k] + xlv * qlten[k] + xls * qiten[k];
            qc[k] = qc_l[k] + qc_i[k];
            rliq = rliq + qc[k] * dp0[k] / g / 1000.0;
        }
        precip = rainflx + snowflx;
        snow = snowflx;

        /* ---- S17 precipitation evaporation (uwshcu.F90:3808-3934) ---- */
        evpint_rain = 0.0;
        evpint_snow = 0.0;
        for (k = 0; k <= mkx; ++k) { flxrain[k] = 0.0; flxsnow[k] = 0.0; }
        for (k = 1; k <= mkx; ++k) { ntraprd[k] = 0.0; ntsnprd[k] = 0.0; }
        for (k = mkx; k >= 1; --k) {
            if (t0[k] > 273.16) snowmlt = dmax(0.0, flxsnow[k] * g / dp0[k]);
            else snowmlt = 0.0;
            orc_qsat(c, t0[k], p0[k], &es, &qs, &gam);
            subsat = dmax((1.0 - qv0[k] / qs), 0.0);
            evprain = kevp * subsat * sqrt(flxrain[k] + snowmlt * dp0[k] / g);
            evpsnow = kevp * subsat * sqrt(dmax(flxsnow[k] - snowmlt * dp0[k] / g, 0.0));
            evplimit = dmax(0.0, (W->qw0[k] - qv0[k]) / dt);
            evplimit_rain = dmin(evplimit, (flxrain[k] + snowmlt * dp0[k] / g) * g / dp0[k]);
            evplimit_rain = dmin(evplimit_rain, (rainflx - evpint_rain) * g / dp0[k]);
            evprain = dmax(0.0, dmin(evplimit_rain, evprain));
            evplimit_snow = dmin(evplimit, dmax(flxsnow[k] - snowmlt * dp0[k] / g, 0.0) * g / dp0[k]);
            evplimit_snow = dmin(evplimit_snow, (snowflx - evpint_snow) * g / dp0[k]);
            evpsnow = dmax(0.0, dmin(evplimit_snow, evpsnow));
            if ((evprain + evpsnow) > evplimit) {
                tmp1 = evprain * evplimit / (evprain + evpsnow);
                tmp2 = evpsnow * evplimit / (evprain + evpsnow);
                evprain = tmp1;
                evpsnow = tmp2;
            }
            evapc[k] = evprain + evpsnow;
            evpint_rain = evpint_rain + evprain * dp0[k] / g;
            evpint_snow = evpint_snow + evpsnow * dp0[k] / g;
            ntraprd[k] = qrten[k] - evprain + snowmlt;
            ntsnprd[k] = qsten[k] - evpsnow - snowmlt;
            flxrain[k - 1] = flxrain[k] + ntraprd[k] * dp0[k] / g;
            flxsnow[k - 1] = flxsnow[k] + ntsnprd[k] * dp0[k] / g;
            flxrain[k - 1] = dmax(flxrain[k - 1], 0.0);
            if (flxrain[k - 1] == 0.0) ntraprd[k] = -flxrain[k] * g / dp0[k];
            flxsnow[k - 1] = dmax(flxsnow[k - 1], 0.0);
            if (flxsnow[k - 1] == 0.0) ntsnprd[k] = -flxsnow[k] * g / dp0[k];
            qlten[k] = qlten[k] - qrten[k];
            qiten[k] = qiten[k] - qsten[k];
            qvten[k] = qvten[k] + evprain + evpsnow;
            qtten[k] = qlten[k] + qiten[k] + qvten[k];
            sten[k] = sten[k] - xlv * evprain - xls * evpsnow - (xls - xlv) * snowmlt;
            slten[k] = sten[k] - xlv * qlten[k] - xls * qiten[k];
        }
        precip = (flxrain[0] + flxsnow[0]) / 1000.0;
        snow = flxsnow[0] / 1000.0;

        /* ---- S18 reserved condensate (uwshcu.F90:3966-3981) ---- */
        for (k = 1; k <= kpen; ++k) {
            qtten[k] = qtten[k] - qc[k];
            qlten[k] = qlten[k] - qc_l[k];
            qiten[k] = qiten[k] - qc_i[k];
            slten[k] = slten[k] + (xlv * qc_l[k] + xls * qc_i[k]);
            sten[k] = sten[k] - (xls - xlv) * qc_i[k];
        }

        /* ---- S19 positive moisture (uwshcu.F90:3995-4001) ---- */
        for (k = 1; k <= mkx; ++k) {
            W->qv0_star[k] = qv0[k] + qvten[k] * dt;
            W->ql0_star[k] = ql0[k] + qlten[k] * dt;
            W->qi0_star[k] = qi0[k] + qiten[k] * dt;
            W->s0_star[k] = s0[k] + sten[k] * dt;
        }
        orc_positive_moisture_single(xlv, xls, mkx, dt, c->qmin[1], c->qmin[2], c->qmin[3], dp0, W->qv0_star,
                                     W->ql0_star, W->qi0_star, W->s0_star, qvten, qlten, qiten, sten);
        for (k = 1; k <= mkx; ++k) {
            qtten[k] = qvten[k] + qlten[k] + qiten[k];
            slten[k] = sten[k] - xlv * qlten[k] - xls * qiten[k];
        }

        /* ---- S20 tracer limiter (uwshcu.F90:4007-4076) ---- */
        for (m = 4; m <= ntr; ++m) {
            if (m != ixnumliq && m != ixnumice) {
                trmin = c->qmin[m];
                for (k = 0; k <= mkx; ++k) { trflx_d[k] = 0.0; trflx_u[k] = 0.0; }
                for (k = 1; k <= mkx - 1; ++k) {
                    pdelx = c->wet[m] ? dp0[k] : W->dpdry0[k];
                    km1 = k - 1;
                    dum = (W->tr0[m][k] - trmin) * pdelx / g / dt + W->trflx[m][km1] - W->trflx[m][k] + trflx_d[km1];
                    trflx_d[k] = dmin(0.0, dum);
                }
                for (k = mkx; k >= 2; --k) {
                    pdelx = c->wet[m] ? dp0[k] : W->dpdry0[k];
                    km1 = k - 1;
                    dum = (W->tr0[m][k] - trmin) * pdelx / g / dt + W->trflx[m][km1] - W->trflx[m][k] + trflx_d[km1] -
                          trflx_d[k] - trflx_u[k];
                    trflx_u[km1] = dmax(0.0, -dum);
                }
                for (k = 1; k <= mkx; ++k) {
                    pdelx = c->wet[m] ? dp0[k] : W->dpdry0[k];
                    km1 = k - 1;
                    W->trten[m][k] = (W->trflx[m][km1] - W->trflx[m][k] + trflx_d[km1] - trflx_d[k] + trflx_u[km1] -
                                      trflx_u[k]) * g / pdelx;
                }
            }
        }

        /* ---- S21 in-cloud diagnostics (uwshcu.F90:4092-4158) ---- */
        orc_conden(c, prel, thlu[krel - 1], qtu[krel - 1], &thj, &qvj, &qlj, &qij, &qse, &id_check);
        if (id_check == 1) { exit_code = EX_CONDEN; id_exit = 1; goto L333; }
        qcubelow = qlj + qij;
        qlubelow = qlj;
        qiubelow = qij;
        for (k = krel; k <= kpen; ++k) {
            if (k == kpen)
                orc_conden(c, ps0[k - 1] + ppen, thlu_top, qtu_top, &thj, &qvj, &qlj, &qij, &qse, &id_check);
            else
                orc_conden(c, ps0[k], thlu[k], qtu[k], &thj, &qvj, &qlj, &qij, &qse, &id_check);
            if (id_check == 1) { exit_code = EX_CONDEN; id_exit = 1; goto L333; }
            qcu[k] = 0.5 * (qcubelow + qlj + qij);
            qlu[k] = 0.5 * (qlubelow + qlj);
            qiu[k] = 0.5 * (qiubelow + qij);
            /* zwshcu.F90:4412-4429: widened cumulus fraction (zmh_cloud_sh) */
            tmp2 = 0.0;
            tmp3 = 1.0;
            if (zw && zw->cloud_sh) {
                tmp3 = 0.5 * zw->ufrc0;
                tmp1 = (z0[k] - z0[krel]) / (z0[kpen] - z0[krel] + 10.0);
                tmp1 = 1.0 + zw->ufrcz * tmp1;
                tmp3 = tmp3 * tmp1;
                tmp2 = fdr[k] * umf[k] * (zs0[k] - zs0[k - 1]) * dt / dp0[k] * 9.8 * zw->ufrcd;
                tmp2 = dmin(1.0 - rmaxfrac, tmp2);
            }
            cufrc[k] = (ufrc[k - 1] + ufrc[k]);
            if (k == krel) {
                cufrc[k] = (ufrclcl + ufrc[k]) * (prel - ps0[k]) / (ps0[k - 1] - ps0[k]);
            } else if (k == kpen) {
                cufrc[k] = (ufrc[k - 1] + 0.0) * (-ppen) / (ps0[k - 1] - ps0[k]);
                if ((qlj + qij) > criqc) {
                    qcu[k] = 0.5 * (qcubelow + criqc);
                    qlu[k] = 0.5 * (qlubelow + criqc * qlj / (qlj + qij));
                    qiu[k] = 0.5 * (qiubelow + criqc * qij / (qlj + qij));
                }
            }
            if (zw) cufrc[k] = tmp3 * cufrc[k] + tmp2;   /* zwshcu.F90:4449 */
            qcubelow = qlj + qij;
            qlubelow = qlj;
            qiubelow = qij;
        }
        cnt = (double)kpen;
        cnb = (double)(krel - 1);

        /* ---- S22 save and re-derive the environment (uwshcu.F90:4169-4329) ---- */
        if (iter != 2) {
            for (k = 1; k <= mkx; ++k) {
                W->qv0_s[k] = qv0[k] + qvten[k] * dt;
                W->ql0_s[k] = ql0[k] + qlten[k] * dt;
                W->qi0_s[k] = qi0[k] + qiten[k] * dt;
                W->s0_s[k] = s0[k] + sten[k] * dt;
                W->u0_s[k] = u0[k] + uten[k] * dt;
                W->v0_s[k] = v0[k] + vten[k] * dt;
                W->qt0_s[k] = W->qv0_s[k] + W->ql0_s[k] + W->qi0_s[k];
                W->t0_s[k] = t0[k] + sten[k] * dt / cp;
                for (m = 1; m <= ntr; ++m) W->tr0_s[m][k] = W->tr0[m][k] + W->trten[m][k] * dt;
                W->qvten_s[k] = qvten[k]; W->qlten_s[k] = qlten[k]; W->qiten_s[k] = qiten[k];
                W->sten_s[k] = sten[k]; W->uten_s[k] = uten[k]; W->vten_s[k] = vten[k];
                W->qrten_s[k] = qrten[k]; W->qsten_s[k] = qsten[k]; W->evapc_s[k] = evapc[k];
                W->cufrc_s[k] = cufrc[k]; W->qcu_s[k] = qcu[k]; W->qlu_s[k] = qlu[k]; W->qiu_s[k] = qiu[k];
                W->qc_s[k] = qc[k];
                for (m = 1; m <= ntr; ++m) W->trten_s[m][k] = W->trten[m][k];
            }
            for (k = 0; k <= mkx; ++k) { W->umf_s[k] = umf[k]; W->slflx_s[k] = slflx[k]; W->qtflx_s[k] = qtflx[k]; }
            W->precip_s = precip; W->snow_s = snow; W->cush_s = cush; W->cbmf_s = cbmf; W->rliq_s = rliq;
            W->cnt_s = cnt; W->cnb_s = cnb;

            for (k = 1; k <= mkx; ++k) {
                qv0[k] = W->qv0_s[k]; ql0[k] = W->ql0_s[k]; qi0[k] = W->qi0_s[k]; s0[k] = W->s0_s[k]; t0[k] = W->t0_s[k];
                qt0[k] = (qv0[k] + ql0[k] + qi0[k]);
                thl0[k] = (t0[k] - xlv * ql0[k] / cp - xls * qi0[k] / cp) / exn0[k];
                thvl0[k] = (1.0 + zvir * qt0[k]) * thl0[k];
            }
            orc_slope(mkx, thl0, p0, ssthl0);
            orc_slope(mkx, qt0, p0, ssqt0);
            orc_slope(mkx, u0, p0, ssu0);
            orc_slope(mkx, v0, p0, ssv0);
            for (m = 1; m <= ntr; ++m) orc_slope(mkx, W->tr0[m], p0, W->sstr0[m]);
            for (k = 1; k <= mkx; ++k) {
                thl0bot = thl0[k] + ssthl0[k] * (ps0[k - 1] - p0[k]);
                qt0bot = qt0[k] + ssqt0[k] * (ps0[k - 1] - p0[k]);
                orc_conden(c, ps0[k - 1], thl0bot, qt0bot, &thj, &qvj, &qlj, &qij, &qse, &id_check);
                if (id_check == 1) { exit_code = EX_CONDEN; id_exit = 1; goto L333; }
                thv0bot[k] = THV(thj, qvj, qlj, qij);
                thvl0bot[k] = thl0bot * (1.0 + zvir * qt0bot);
                thl0top = thl0[k] + ssthl0[k] * (ps0[k] - p0[k]);
                qt0top = qt0[k] + ssqt0[k] * (ps0[k] - p0[k]);
                orc_conden(c, ps0[k], thl0top, qt0top, &thj, &qvj, &qlj, &qij, &qse, &id_check);
                if (id_check == 1) { exit_code = EX_CONDEN; id_exit = 1; goto L333; }
                thv0top[k] = THV(thj, qvj, qlj, qij);
                thvl0top[k] = thl0top * (1.0 + zvir * qt0top);
            }
        }
    }

    /* ---- update outputs (uwshcu.F90:4337-4370) ---- */
    for (k = 0; k <= mkx; ++k) {
        W->o_umf[k] = umf[k];
        W->o_wu[k] = wu[k];
        W->o_slflx[k] = slflx[k];
        W->o_qtflx[k] = qtflx[k];
        W->o_flxprc1[k] = flxrain[k] + flxsnow[k];
        W->o_flxsnow1[k] = flxsnow[k];
    }
    for (k = 1; k <= mkx; ++k) {
        W->o_qvten[k] = qvten[k]; W->o_qlten[k] = qlten[k]; W->o_qiten[k] = qiten[k]; W->o_sten[k] = sten[k];
        W->o_uten[k] = uten[k]; W->o_vten[k] = vten[k]; W->o_qrten[k] = qrten[k]; W->o_qsten[k] = qsten[k];
        W->o_evapc[k] = evapc[k]; W->o_cufrc[k] = cufrc[k]; W->o_qcu[k] = qcu[k]; W->o_qlu[k] = qlu[k];
        W->o_qiu[k] = qiu[k]; W->o_qc[k] = qc[k];
        for (m = 1; m <= ntr; ++m) W->o_trten[m][k] = W->trten[m][k];
    }
    W->o_precip = precip; W->o_snow = snow; W->cush = cush; W->o_cbmf = cbmf; W->o_rliq = rliq;
    W->o_cnt = cnt; W->o_cnb = cnb;
    path = 2;
    goto Ldone;

L333:
    /* exit without convection (uwshcu.F90:4448-4482); flxprc1/flxsnow1 keep their
     * initial zero (uwshcu.F90:1090-1091) */
    for (k = 0; k <= mkx; ++k) {
        W->o_umf[k] = 0.0; W->o_wu[k] = 0.0; W->o_slflx[k] = 0.0; W->o_qtflx[k] = 0.0;
        W->o_flxprc1[k] = 0.0; W->o_flxsnow1[k] = 0.0;
    }
    for (k = 1; k <= mkx; ++k) {
        W->o_qvten[k] = 0.0; W->o_qlten[k] = 0.0; W->o_qiten[k] = 0.0; W->o_sten[k] = 0.0;
        W->o_uten[k] = 0.0; W->o_vten[k] = 0.0; W->o_qrten[k] = 0.0; W->o_qsten[k] = 0.0;
        W->o_evapc[k] = 0.0; W->o_cufrc[k] = 0.0; W->o_qcu[k] = 0.0; W->o_qlu[k] = 0.0;
        W->o_qiu[k] = 0.0; W->o_qc[k] = 0.0;
        for (m = 1; m <= ntr; ++m) W->o_trten[m][k] = 0.0;
    }
    W->o_precip = 0.0; W->o_snow = 0.0; W->cush = -1.0; W->o_cbmf = 0.0; W->o_rliq = 0.0;
    W->o_cnt = 1.0; W->o_cnb = (double)mkx;
    path = 0;
    (void)id_exit;

Ldone:
    W->diag[0] = (path == 0) ? exit_code : 0;
    W->diag[1] = kinv; W->diag[2] = klcl; W->diag[3] = klfc; W->diag[4] = krel;
    W->diag[5] = kbup; W->diag[6] = kpen; W->diag[7] = path; W->diag[8] = limit_ufrc;
    W->diag[9] = limit_ppen; W->diag[10] = limit_emf; W->diag[11] = exit_iter;
    return mumin2_fail ? -3 : 0;
}

/* ------------------------------------------------------------------------- */
/* compute_uwshcu_inv (uwshcu.F90:251-476)                                    */
/* ------------------------------------------------------------------------- */

/* compute_uwshcu_inv, or compute_zwshcu_inv (zwshcu.F90:336-628) when zw is not
 * NULL.  The zwshcu wrapper caps tpert and qpert (zwshcu.F90:547-551) and, with
 * pertt, adds them to t0, qv0 and s0 (cp * tpert) in every UW layer whose bottom
 * interface height zs0(k-1) is not above pblh (zwshcu.F90:556-566), before the
 * column physics and its wet-bulb computation see the inputs. */
static int inv_common(const orc_consts *c, const orc_zw *zw, int mix, int mkx, int iend, int ncnst, double dt,
    const double *ps0_inv, const double *zs0_inv, const double *p0_inv, const double *z0_inv,
    const double *dp0_inv, const double *u0_inv, const double *v0_inv, const double *qv0_inv,
    const double *ql0_inv, const double *qi0_inv, const double *t0_inv, const double *s0_inv,
    const double *tr0_inv, const double *tke_inv, const double *pblh, double *cush,
    const double *omega_inv, const double *tpert, const double *qpert,
    double *umf_inv, double *slflx_inv, double *qtflx_inv, double *flxprc1_inv, double *flxsnow1_inv,
    double *qvten_inv, double *qlten_inv, double *qiten_inv, double *sten_inv, double *uten_inv,
    double *vten_inv, double *trten_inv, double *qrten_inv, double *qsten_inv, double *precip,
    double *snow, double *evapc_inv, double *cufrc_inv, double *qcu_inv, double *qlu_inv,
    double *qiu_inv, double *cbmf, double *qc_inv, double *rliq, double *cnt_inv, double *cnb_inv,
    const double *dpdry0_inv, double *wu_inv, int *diag, int *bad_col)
{
    int ntr = ncnst;
    if (c->ixnumliq > ntr) ntr = c->ixnumliq;
    if (c->ixnumice > ntr) ntr = c->ixnumice;
    if (mkx < 3 || mkx >= KMAX - 1 || ncnst < 0 || ntr > ORC_NCNST_MAX || iend > mix) {
        if (bad_col) *bad_col = 0;
        return -4;
    }
    if (zw && (!omega_inv || !tpert || !qpert)) {
        if (bad_col) *bad_col = 0;
        return -4;
    }
    colws *W = (colws *)calloc(1, sizeof(colws));
    int rc = 0;
#define MID(a, i, kk) (a)[(size_t)((mkx + 1 - (kk)) - 1) * mix + (i)]
#define INT(a, i, kk) (a)[(size_t)((mkx + 1 - (kk)) - 1) * mix + (i)]
    /* the zwshcu input perturbation of column i, UW layer k: 1 where it applies */
#define PERT(i, kk) (zw && zw->pertt && !(zs0_inv[(size_t)(mkx - ((kk) - 1)) * mix + (i)] > pblh[i]))
#define TPERT(i) dmin(tpert[i], zw->tpertmax)
#define QPERT(i) dmin(qpert[i], zw->qpertmax)
    /* findsp over the whole chunk first (uwshcu.F90:1216), in UW level order */
    for (int i = 0; i < iend && rc == 0; ++i) {
        for (int k = 1; k <= mkx; ++k) {
            double tsp, qsp, enerr;
            double qvk = MID(qv0_inv, i, k), tk = MID(t0_inv, i, k);
            if (PERT(i, k)) { qvk = qvk + QPERT(i); tk = tk + TPERT(i); }
            int doit = orc_findsp_point(c, qvk, tk, MID(p0_inv, i, k), &tsp, &qsp, &enerr);
            if (doit == 0) { rc = -1; if (bad_col) *bad_col = i; break; }
            if (doit == 2 && enerr > 1.e-4) { rc = -2; if (bad_col) *bad_col = i; break; }
        }
    }
    if (rc != 0) { free(W); return rc; }

    for (int i = 0; i < iend; ++i) {
        for (int k = 1; k <= mkx; ++k) {
            /* level flip, uwshcu.F90:380-398 (k_inv = mkx+1-k) */
            W->p0[k] = MID(p0_inv, i, k); W->u0[k] = MID(u0_inv, i, k); W->v0[k] = MID(v0_inv, i, k);
            W->z0[k] = MID(z0_inv, i, k); W->dp0[k] = MID(dp0_inv, i, k); W->dpdry0[k] = MID(dpdry0_inv, i, k);
            W->qv0[k] = MID(qv0_inv, i, k); W->ql0[k] = MID(ql0_inv, i, k); W->qi0[k] = MID(qi0_inv, i, k);
            W->t0[k] = MID(t0_inv, i, k); W->s0[k] = MID(s0_inv, i, k);
            if (zw) W->omega[k] = MID(omega_inv, i, k);
            if (PERT(i, k)) {
                W->qv0[k] = W->qv0[k] + QPERT(i);
                W->t0[k] = W->t0[k] + TPERT(i);
                W->s0[k] = W->s0[k] + c->cp * TPERT(i);
            }
            for (int m = 1; m <= ncnst; ++m)
                W->tr0[m][k] = tr0_inv[((size_t)(m - 1) * mkx + (mkx - k)) * mix + i];
            double tsp, qsp, enerr;
            orc_findsp_point(c, W->qv0[k], W->t0[k], W->p0[k], &tsp, &qsp, &enerr);
            W->qw0[k] = qsp;
        }
        for (int k = 0; k <= mkx; ++k) {
            /* uwshcu.F90:409-414 (k_inv = mkx+1-k over 1..mkx+1) */
            W->ps0[k] = ps0_inv[(size_t)(mkx - k) * mix + i];
            W->zs0[k] = zs0_inv[(size_t)(mkx - k) * mix + i];
            W->tke[k] = tke_inv[(size_t)(mkx - k) * mix + i];
        }
        W->pblh = pblh[i];
        W->cush = cush[i];
        W->tpert = zw ? TPERT(i) : 0.0;
        int crc = column(c, W, mkx, ntr, dt, zw);
        if (crc != 0 && rc == 0) { rc = crc; if (bad_col) *bad_col = i; }
        /* flip back, uwshcu.F90:434-469 */
        cush[i] = W->cush;
        cnt_inv[i] = (double)(mkx + 1) - W->o_cnt;
        cnb_inv[i] = (double)(mkx + 1) - W->o_cnb;
        precip[i] = W->o_precip; snow[i] = W->o_snow; cbmf[i] = W->o_cbmf; rliq[i] = W->o_rliq;
        for (int k = 0; k <= mkx; ++k) {
            size_t o = (size_t)(mkx - k) * mix + i;
            umf_inv[o] = W->o_umf[k]; wu_inv[o] = W->o_wu[k]; slflx_inv[o] = W->o_slflx[k];
            qtflx_inv[o] = W->o_qtflx[k]; flxprc1_inv[o] = W->o_flxprc1[k]; flxsnow1_inv[o] = W->o_flxsnow1[k];
        }
        for (int k = 1; k <= mkx; ++k) {
            size_t o = (size_t)(mkx - k) * mix + i;
            qvten_inv[o] = W->o_qvten[k]; qlten_inv[o] = W->o_qlten[k]; qiten_inv[o] = W->o_qiten[k];
            sten_inv[o] = W->o_sten[k]; uten_inv[o] = W->o_uten[k]; vten_inv[o] = W->o_vten[k];
            qrten_inv[o] = W->o_qrten[k]; qsten_inv[o] = W->o_qsten[k]; evapc_inv[o] = W->o_evapc[k];
            cufrc_inv[o] = W->o_cufrc[k]; qcu_inv[o] = W->o_qcu[k]; qlu_inv[o] = W->o_qlu[k];
            qiu_inv[o] = W->o_qiu[k]; qc_inv[o] = W->o_qc[k];
            for (int m = 1; m <= ncnst; ++m)
                trten_inv[((size_t)(m - 1) * mkx + (mkx - k)) * mix + i] = W->o_trten[m][k];
        }
        if (diag)
            for (int d = 0; d < ORC_NDIAG; ++d) diag[(size_t)i * ORC_NDIAG + d] = W->diag[d];
    }
#undef MID
#undef INT
#undef PERT
#undef TPERT
#undef QPERT
    free(W);
    return rc;
}

int orc_compute_uwshcu_inv(const orc_consts *c, int mix, int mkx, int iend, int ncnst, double dt,
    const double *ps0_inv, const double *zs0_inv, const double *p0_inv, const double *z0_inv,
    const double *dp0_inv, const double *u0_inv, const double *v0_inv, const double *qv0_inv,
    const double *ql0_inv, const double *qi0_inv, const double *t0_inv, const double *s0_inv,
    const double *tr0_inv, const double *tke_inv, const double *cldfrct_inv,
    const double *concldfrct_inv, const double *pblh, double *cush,
    double *umf_inv, double *slflx_inv, double *qtflx_inv, double *flxprc1_inv, double *flxsnow1_inv,
    double *qvten_inv, double *qlten_inv, double *qiten_inv, double *sten_inv, double *uten_inv,
    double *vten_inv, double *trten_inv, double *qrten_inv, double *qsten_inv, double *precip,
    double *snow, double *evapc_inv, double *cufrc_inv, double *qcu_inv, double *qlu_inv,
    double *qiu_inv, double *cbmf, double *qc_inv, double *rliq, double *cnt_inv, double *cnb_inv,
    const double *dpdry0_inv, double *wu_inv, int *diag, int *bad_col)
{
    (void)cldfrct_inv; (void)concldfrct_inv;
    return inv_common(c, NULL, mix, mkx, iend, ncnst, dt, ps0_inv, zs0_inv, p0_inv, z0_inv, dp0_inv, u0_inv,
        v0_inv, qv0_inv, ql0_inv, qi0_inv, t0_inv, s0_inv, tr0_inv, tke_inv, pblh, cush, NULL, NULL, NULL,
        umf_inv, slflx_inv, qtflx_inv, flxprc1_inv, flxsnow1_inv, qvten_inv, qlten_inv, qiten_inv, sten_inv,
        uten_inv, vten_inv, trten_inv, qrten_inv, qsten_inv, precip, snow, evapc_inv, cufrc_inv, qcu_inv,
        qlu_inv, qiu_inv, cbmf, qc_inv, rliq, cnt_inv, cnb_inv, dpdry0_inv, wu_inv, diag, bad_col);
}

int orc_compute_zwshcu_inv(const orc_consts *c, const orc_zw *zw, int mix, int mkx, int iend, int ncnst,
    double dt, const double *ps0_inv, const double *zs0_inv, const double *p0_inv, const double *z0_inv,
    const double *dp0_inv, const double *u0_inv, const double *v0_inv, const double *qv0_inv,
    const double *ql0_inv, const double *qi0_inv, const double *t0_inv, const double *s0_inv,
    const double *tr0_inv, const double *tke_inv, const double *pblh, double *cush,
    const double *omega_inv, const double *tpert, const double *qpert,
    double *umf_inv, double *slflx_inv, double *qtflx_inv, double *flxprc1_inv, double *flxsnow1_inv,
    double *qvten_inv, double *qlten_inv, double *qiten_inv, double *sten_inv, double *uten_inv,
    double *vten_inv, double *trten_inv, double *qrten_inv, double *qsten_inv, double *precip,
    double *snow, double *evapc_inv, double *cufrc_inv, double *qcu_inv, double *qlu_inv,
    double *qiu_inv, double *cbmf, double *qc_inv, double *rliq, double *cnt_inv, double *cnb_inv,
    const double *dpdry0_inv, double *wu_inv, int *diag, int *bad_col)
{
    if (!zw) {
        if (bad_col) *bad_col = 0;
        return -4;
    }
    return inv_common(c, zw, mix, mkx, iend, ncnst, dt, ps0_inv, zs0_inv, p0_inv, z0_inv, dp0_inv, u0_inv,
        v0_inv, qv0_inv, ql0_inv, qi0_inv, t0_inv, s0_inv, tr0_inv, tke_inv, pblh, cush, omega_inv, tpert, qpert,
        umf_inv, slflx_inv, qtflx_inv, flxprc1_inv, flxsnow1_inv, qvten_inv, qlten_inv, qiten_inv, sten_inv,
        uten_inv, vten_inv, trten_inv, qrten_inv, qsten_inv, precip, snow, evapc_inv, cufrc_inv, qcu_inv,
        qlu_inv, qiu_inv, cbmf, qc_inv, rliq, cnt_inv, cnb_inv, dpdry0_inv, wu_inv, diag, bad_col);
}
