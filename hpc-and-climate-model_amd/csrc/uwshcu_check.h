// uwshcu_check.h -- conservation check of one column's outputs, in FP64.
//
// The scheme conserves column water and, with the kinetic energy of the momentum
// tendencies, column energy exactly up to round-off:
//   water   sum_k (qvten+qlten+qiten) dp/g + (precip + rliq)*1000 = 0
//   energy  sum_k (sten + dKE/dt + (xlv+xlf) qvten + xlf qlten) dp/g
//             = -(precip + rliq - snow)*1000*xlf
// (the budget CAM checks after this scheme: check_energy.F90:336-367 with
// flx_cnd = precip + rliq and flx_ice = snow, tphysbc.F90:444-449).  In FP64 both
// close to ~1e-14 of the largest flux; with FP32 plume arithmetic they close to
// ~1e-6.  A column whose residual exceeds tol times its largest flux, or whose
// outputs are not finite, fails the check.
#ifndef UWSHCU_CHECK_H
#define UWSHCU_CHECK_H

#include "uwshcu_column.h"

namespace uwshcu {

UWSHCU_HD inline bool finite_value(double x) { return x - x == 0.0; }   // false for NaN and +-Inf

// Budget residuals of column i: r[0] water (kg/m2/s), r[1] energy (W/m2), r[2] the
// largest |qtflx|, r[3] the largest |slflx|; returns false if an output is not finite.
UWSHCU_HD inline bool column_budget(const FwdArgs& a, const uwshcu_consts& C, int i, double* r) {
    const int mkx = a.mkx;
    const size_t ldc = (size_t)a.ldc;
    auto at = [&](const double* p, int row) { return p[(size_t)row * ldc + i]; };
    const double dt = a.dt;
    double w = 0.0, te = 0.0, acc = 0.0, fq = 0.0, fs = 0.0;
    for (int r0 = 0; r0 < mkx; ++r0) {
        const double dpg = at(a.dp0, r0) / C.g;
        const double qv = at(a.qvten, r0), ql = at(a.qlten, r0), qi = at(a.qiten, r0);
        const double u0 = at(a.u0, r0), v0 = at(a.v0, r0);
        const double uf = u0 + at(a.uten, r0) * dt, vf = v0 + at(a.vten, r0) * dt;
        // (uf-u0)(uf+u0): exactly zero where the wind tendency is, whether or not the
        // compiler fuses the products
        const double dke = 0.5 * ((uf - u0) * (uf + u0) + (vf - v0) * (vf + v0)) / dt;
        w += (qv + ql + qi) * dpg;
        te += (at(a.sten, r0) + dke + (C.xlv + C.xlf) * qv + C.xlf * ql) * dpg;
        acc += fabs(at(a.qrten, r0)) + fabs(at(a.qsten, r0)) + fabs(at(a.evapc, r0)) + fabs(at(a.cufrc, r0)) +
               fabs(at(a.qcu, r0)) + fabs(at(a.qlu, r0)) + fabs(at(a.qiu, r0)) + fabs(at(a.qc, r0));
        for (int m = 0; m < a.ncnst; ++m) acc += fabs(a.trten[((size_t)m * mkx + r0) * ldc + i]);
    }
    for (int r0 = 0; r0 <= mkx; ++r0) {
        fq = fmax(fq, fabs(at(a.qtflx, r0)));
        fs = fmax(fs, fabs(at(a.slflx, r0)));
        acc += fabs(at(a.umf, r0)) + fabs(at(a.wu, r0)) + fabs(at(a.flxprc1, r0)) + fabs(at(a.flxsnow1, r0));
    }
    const double precip = a.precip[i], rliq = a.rliq[i], snow = a.snow[i];
    w += (precip + rliq) * 1000.0;
    const double eres = te + (precip + rliq - snow) * 1000.0 * C.xlf;
    acc += fabs(a.cbmf[i]) + fabs(a.cush[i]) + fs + fq;
    r[0] = w;
    r[1] = eres;
    r[2] = fq;
    r[3] = fs;
    return finite_value(w) && finite_value(eres) && finite_value(acc);
}

UWSHCU_HD inline bool column_fails_check(const FwdArgs& a, const uwshcu_consts& C, int i, double tol) {
    double r[4];
    if (!column_budget(a, C, i, r)) return true;
    return (r[2] > 0.0 && fabs(r[0]) > tol * r[2]) || (r[3] > 0.0 && fabs(r[1]) > tol * r[3]);
}

}  // namespace uwshcu
#endif
