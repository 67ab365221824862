/*
 * uwshcu_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * Serial FP64 CPU restatement of the UW shallow-cumulus column physics
 * (compute_uwshcu_inv / compute_uwshcu and helpers, uwshcu.F90:251-5454;
 * saturation tables wv_saturation.F90:64-204,482-609,734-751; gffgch.F90).
 *
 * This library is the checker for the HIP product path.  Only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark may load
 * it.  Nothing in the shipped library links or calls it.
 */
#ifndef UWSHCU_ORACLE_H
#define UWSHCU_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NCNST_MAX 32

/* Physical constants and tables.  Field meaning follows the module variables of
 * uwshcu.F90:25-37 (xlv..rovcp, rpen) and wv_saturation.F90:48-60. */
typedef struct orc_consts {
    /* uwshcu module (init_uwshcu, uwshcu.F90:230-239) */
    double xlv, xlf, xls, cp, zvir, r, g, ep2, p00, rovcp;
    /* tunables: rpen is a namelist value (uwshcu.F90:37,95); rle and rkm are
     * compile-time parameters in the reference (uwshcu.F90:1036-1038) */
    double rpen, rle, rkm;
    /* wv_saturation module (gestbl, wv_saturation.F90:125-140) */
    double tmin, tmax, ttrice, epsqs, rgasv, hlatf, hlatv, cpws, tmelt;
    int    icephs;
    double pcf[5];
    double estbl[250];
    /* constituents: minimum values and 'wet' (1) / 'dry' (0) type, 1-based index */
    double qmin[ORC_NCNST_MAX + 1];
    int    wet[ORC_NCNST_MAX + 1];
    int    ixnumliq, ixnumice;   /* 1-based constituent indices of NUMLIQ/NUMICE */
} orc_consts;

/* Fill CESM default constants (shr_const values) and build the Goff-Gratch table. */
void orc_default_consts(orc_consts *c);

/* Primitives, exposed for unit tests.  orc_xexp/xlog/xpow/xerfc are the flavour's
 * transcendental functions (host libm or portable; erfc is Cody's in both). */
double orc_xexp(double x);
double orc_xlog(double x);
double orc_xpow(double x, double y);
double orc_xerfc(double x);
void   orc_gffgch(const orc_consts *c, double t, double *es, int itype);
void   orc_gestbl(orc_consts *c, double tmn, double tmx, double trice, int ip);
double orc_estblf(const orc_consts *c, double td);
void   orc_qsat(const orc_consts *c, double t, double p, double *es, double *qs, double *gam);
void   orc_conden(const orc_consts *c, double p, double thl, double qt,
                  double *th, double *qv, double *ql, double *qi, double *rvls, int *id_check);
void   orc_roots(double a, double b, double c, double *r1, double *r2, int *status);
void   orc_slope(int mkx, const double *field, const double *p0, double *out);
double orc_qsinvert(const orc_consts *c, double qt, double thl, double psfc);
double orc_compute_alpha(double del_CIN, double ke);
double orc_compute_mumin2(double mulcl, double rmaxfrac, double mulow);
double orc_compute_ppen(double wtwb, double D, double bogbot, double bogtop, double rho0j, double dpen);
void   orc_fluxbelowinv(const orc_consts *c, double cbmf, const double *ps0, int mkx, int kinv, double dt,
                        double xsrc, double xmean, double xtopin, double xbotin, double *xflx);
void   orc_positive_moisture_single(double xlv, double xls, int mkx, double dt, double qvmin, double qlmin,
                                    double qimin, const double *dp, double *qv, double *ql, double *qi,
                                    double *s, double *qvten, double *qlten, double *qiten, double *sten);
double orc_single_cin(const orc_consts *c, double pbot, double thv0bot, double ptop, double thv0top,
                      double thvubot, double thvutop);
void   orc_getbuoy(const orc_consts *c, double pbot, double thv0bot, double ptop, double thv0top,
                   double thvubot, double thvutop, double *plfc, double *cin);
/* Wet bulb for one (column,level) point; returns doit code (0 not converged, 1 skipped,
 * 2 converged, 4 too cold); *enerr = relative enthalpy error for converged points. */
int    orc_findsp_point(const orc_consts *c, double q, double t, double p,
                        double *tsp, double *qsp, double *enerr);

/* Stage probe (tests only).  orc_set_probe(buf) makes the column code of the calling
 * thread record internal quantities into buf (NULL stops it); meaningful for calls of
 * one column.  buf holds ORC_P_NSCAL scalars, then ORC_PL_NFIELD per-level fields of
 * ORC_PRB_KMAX entries each (field f, UW level k at ORC_P_NSCAL + f*ORC_PRB_KMAX + k).
 * Values are those of the last pass that wrote them; per-level fields are cleared at
 * the start of each cloud-top (iter_scaleh) pass.  UW levels: interfaces 0..mkx,
 * layers 1..mkx. */
#define ORC_PRB_KMAX 128
enum {
    /* cloud-base closure (uwshcu.F90:2092-2184) */
    ORC_P_CLOSURE, ORC_P_WCRIT, ORC_P_SIGMAW, ORC_P_MU, ORC_P_RHO0INV, ORC_P_CBMF, ORC_P_WINV, ORC_P_WLCL,
    ORC_P_UFRCLCL, ORC_P_CINLCL, ORC_P_RBUOY, ORC_P_RMAXFRAC, ORC_P_CBMFLIMIT, ORC_P_MUMIN2_APPLIED,
    /* cloud-top penetration (uwshcu.F90:2873-2899) */
    ORC_P_PPEN_SET, ORC_P_WTWB, ORC_P_DRAGE, ORC_P_BOGBOT, ORC_P_BOGTOP, ORC_P_RHO0J, ORC_P_DPEN, ORC_P_PPEN,
    /* levels and constants */
    ORC_P_KRELP, ORC_P_KBUP, ORC_P_KPEN, ORC_P_RPEN, ORC_P_PGFC, ORC_P_G, ORC_P_CRIQC,
    /* penetrative entrainment (uwshcu.F90:3055-3183) */
    ORC_P_EMF_SET, ORC_P_LIMIT_EMF,
    ORC_P_NSCAL
};
enum {
    /* buoyancy sorting, per layer k of the final pass (uwshcu.F90:2306-2952) */
    ORC_PL_SET, ORC_PL_CASE, ORC_PL_XC, ORC_PL_WUE, ORC_PL_CRIDIS, ORC_PL_THVJ, ORC_PL_THV0J, ORC_PL_XSAT,
    ORC_PL_THVXSAT, ORC_PL_THLUE, ORC_PL_QTUE, ORC_PL_EXCESSU, ORC_PL_EXCESS0,
    ORC_PL_THLE, ORC_PL_QTE, ORC_PL_UE, ORC_PL_VE, ORC_PL_PE, ORC_PL_DPE, ORC_PL_FER, ORC_PL_FDR,
    ORC_PL_THLU0, ORC_PL_QTU0, ORC_PL_UU0, ORC_PL_VU0, ORC_PL_THLU1, ORC_PL_QTU1, ORC_PL_UU1, ORC_PL_VU1,
    ORC_PL_SSTHL, ORC_PL_SSQT, ORC_PL_SSU, ORC_PL_SSV, ORC_PL_UMF0, ORC_PL_UMF1, ORC_PL_REI,
    ORC_PL_WTWB, ORC_PL_WTW, ORC_PL_DRAGE, ORC_PL_BOGBOT, ORC_PL_BOGTOP, ORC_PL_RHO0J,
    /* penetrative entrainment per interface k, and the environment per layer k */
    ORC_PL_EMF, ORC_PL_THLU_EMF, ORC_PL_QTU_EMF, ORC_PL_THL0, ORC_PL_QT0, ORC_PL_PS0, ORC_PL_P0,
    ORC_PL_SSTHL0, ORC_PL_SSQT0, ORC_PL_UMF, ORC_PL_REIF, ORC_PL_DP0,
    ORC_PL_NFIELD
};
void orc_set_probe(double *buf);
int  orc_probe_layout(int *nscal, int *kmax, int *nfield);

/* Number of integer diagnostics per column written by orc_compute_uwshcu_inv. */
#define ORC_NDIAG 12
/* diag[] layout per column:
 *  0 exit code (0 convective, >0 reason of the final exit; see oracle source)
 *  1 kinv 2 klcl 3 klfc 4 krel 5 kbup 6 kpen  (UW bottom-up indices of the final pass)
 *  7 path: 0 exit, 1 iteration-1 result kept (del_CIN<=0), 2 implicit-CIN result
 *  8 limit_ufrc 9 limit_ppen 10 limit_emf 11 iteration in which the column exited */

/*
 * Same argument list and layout as compute_uwshcu_inv (uwshcu.F90:251-265):
 * Fortran column-major arrays with leading dimension mix (= pcols), CAM level
 * order (index 1 = model top).  Constituents are tr0(mix,mkx,ncnst).
 * Returns 0, or a negative code for conditions the reference treats as fatal:
 *   -1 findsp not converging, -2 findsp enthalpy error, -3 mu > mumin2,
 *   -4 unsupported sizes.  *bad_col receives the 0-based offending column.
 */
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
    const double *dpdry0_inv, double *wu_inv, int *diag, int *bad_col);

/*
 * The modified scheme compute_zwshcu_inv (zwshcu.F90:336-628 -> compute_zwshcu
 * zwshcu.F90:630-4969): the same column physics with the options below
 * (zwshcu.F90:30-57; zmh_cloud_sh is cloud_fraction.F90:52).  Flags are 0 / 1.
 */
typedef struct orc_zw {
    int lf, ent, cloud_sh, pertt, pertw, pad_;
    double fwcrit, ent0, ufrcd, ufrc0, ufrcz, tpertmax, qpertmax;
} orc_zw;

/*
 * compute_zwshcu_inv with the arguments of orc_compute_uwshcu_inv plus
 * omega_inv (mix, mkx) in CAM level order and tpert, qpert (mix).  The other
 * inputs of the reference (gradt, gradq, landfrac, lat, lon, massflx_sh) feed
 * only options this restatement does not carry (zwshcu.F90:518-540, 2319,
 * 2382) and are not taken.  rpen and rkm come from orc_consts (the reference
 * reads uwshcu_rpen and zmh_rkm from its namelist, zwshcu.F90:43-44,1251).
 */
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
    const double *dpdry0_inv, double *wu_inv, int *diag, int *bad_col);

#ifdef __cplusplus
}
#endif
#endif
