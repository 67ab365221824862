// uwshcu_types.h -- plain-C types shared by the HIP kernels and the C-ABI host layer.
//
// No torch or HIP types appear here: these structs cross the extern "C" boundary.
#ifndef UWSHCU_TYPES_H
#define UWSHCU_TYPES_H

#ifdef __cplusplus
extern "C" {
#endif

#define UWSHCU_NCNST_MAX 32
#define UWSHCU_ESTBL_LEN 250
#define UWSHCU_NDIAG 12

// Physical constants of the scheme and of the saturation table.  Field meaning
// follows the module variables of uwshcu.F90:25-37 and wv_saturation.F90:48-60.
typedef struct uwshcu_consts {
    double xlv, xlf, xls, cp, zvir, r, g, ep2, p00, rovcp;   // uwshcu.F90:230-239
    double tmin, tmax, ttrice, epsqs, rgasv, hlatf, hlatv, cpws, tmelt;  // wv_saturation.F90:125-140
    int    icephs;
    int    ixnumliq, ixnumice;       // 1-based constituent indices of NUMLIQ / NUMICE
    int    pad_;
    double pcf[5];                   // wv_saturation.F90:178-182
    double qmin[UWSHCU_NCNST_MAX + 1];   // 1-based; qmin[1..3] = Q, CLDLIQ, CLDICE
    int    wet[UWSHCU_NCNST_MAX + 1];    // 1 = 'wet' constituent, 0 = 'dry' (cnst_get_type_byind)
} uwshcu_consts;

// Tunable parameters: rpen is the uwshcu_nl namelist value (uwshcu.F90:58-98);
// rle and rkm are compile-time parameters in the reference (uwshcu.F90:1036-1038)
// and runtime values here.
typedef struct uwshcu_params {
    double rpen, rle, rkm;
} uwshcu_params;

// Options of the modified scheme compute_zwshcu_inv (zwshcu.F90:30-57, namelist
// zwshcu_nl zwshcu.F90:100-107).  Flags are 0 or 1; the reference defaults are in
// brackets.  cloud_sh is cloud_fraction's zmh_cloud_sh (cloud_fraction.F90:52), which
// the ZMH_WXC and WXC macrophysics switch on (cloud_fraction.F90:204-207).
typedef struct uwshcu_zw_params {
    int lf;          // zmh_shallow_lf [1]: large-scale ascent lowers the critical updraft speed
    int ent;         // zmh_shallow_ent [0]: buoyancy-based lateral entrainment rate
    int cloud_sh;    // zmh_cloud_sh [0]: cumulus fraction widened by detrainment
    int pertt;       // zmh_shallow_pertt [0]: tpert / qpert added to the PBL layers
    int pertw;       // zmh_shallow_pertw [0]: tpert widens the updraft-speed distribution
    int pad_;
    double fwcrit;   // zmh_shallow_fwcrit [1.0]: factor on the critical updraft speed
    double ent0;     // zmh_ent0 [0.015 as a single-precision literal]
    double ufrcd, ufrc0, ufrcz;   // zmh_ufrcd, zmh_ufrc0, zmh_ufrcz [2.0 each]
    double tpertmax; // zmh_shallow_tpertmax [2.0]: cap on tpert
    double qpertmax; // zmh_shallow_qpertmax [1.0e-3 as a single-precision literal]: cap on qpert
} uwshcu_zw_params;

// Precision modes.
enum {
    UWSHCU_FP64  = 0,   // the reference's arithmetic in FP64, bit for bit
    // The fast path: every state in FP64, the saturation adjustment (conden) solved by
    // Newton's method with the table's own derivative to 1e-10 of T instead of by the
    // reference's ten fixed steps; then every column that fails the FP64 water/energy
    // conservation check (relative tolerance 1e-4 by default) or took a branch decision
    // within 1e-8 of its branch point is recomputed by the FP64 path on the GPU.
    // Tendencies stay within about 1e-6 of UWSHCU_FP64 (relative L2 over a chunk).
    UWSHCU_MIXED = 1,
    // Experimental: FP32 plume arithmetic (FP64 saturation adjustment, LCL, CIN,
    // closure and fix-ups), no check and no recompute.  Its tendencies differ from
    // UWSHCU_FP64 by 1-2 % (relative L2): the buoyancy differences of ~300 K virtual
    // temperatures carry 1e-4 relative FP32 error, which the plume amplifies.
    UWSHCU_MIXED_UNCHECKED = 2,
    // The fast pass of UWSHCU_MIXED alone: no check, no recompute.
    UWSHCU_MIXED_FAST_UNCHECKED = 3
};

// Cotangents of the outputs for the adjoint: device pointers in the caller's layout
// (the same layout as the outputs); NULL means zero.
typedef struct uwshcu_cotangents {
    const double *umf, *slflx, *qtflx, *flxprc1, *flxsnow1, *wu;
    const double *qvten, *qlten, *qiten, *sten, *uten, *vten, *qrten, *qsten, *evapc;
    const double *cufrc, *qcu, *qlu, *qiu, *qc;
    const double *trten;
    const double *precip, *snow, *cbmf, *rliq, *cnt, *cnb, *cush;
} uwshcu_cotangents;

// Gradients written by the adjoint (device pointers, caller's layout; NULL = not
// wanted): t0, qv0, u0, v0 (pcols, pver); tke (pcols, pver+1); pblh (pcols);
// par_col (3, pcols): each column's contribution to d/d(rpen, rle, rkm);
// par_sum [3]: those contributions summed over the columns in FP64.
typedef struct uwshcu_gradients {
    double *t0, *qv0, *u0, *v0, *tke, *pblh, *par_col, *par_sum;
} uwshcu_gradients;

// One physics chunk for uwshcu_forward_chunks_device: the device arrays of one
// bxcompute_uwshcu_inv_ call (Fortran (pcols, pver[+1]) layout, CAM level order,
// tr0/trten (pcols, pver, ncnst)).  Only the first ncol columns are read or written.
// tr0/trten may be NULL when ncnst is 0; every other array is required when ncol > 0.
typedef struct uwshcu_chunk {
    int pcols, ncol;
    const double *ps0, *zs0, *p0, *z0, *dp0, *u0, *v0, *qv0, *ql0, *qi0, *t0, *s0, *tr0, *tke, *pblh, *dpdry0;
    double *cush;
    double *umf, *slflx, *qtflx, *flxprc1, *flxsnow1, *wu;
    double *qvten, *qlten, *qiten, *sten, *uten, *vten, *trten, *qrten, *qsten, *evapc, *cufrc, *qcu, *qlu, *qiu, *qc;
    double *precip, *snow, *cbmf, *rliq, *cnt, *cnb;
} uwshcu_chunk;

// Fields of the driver-side merge that follows the scheme in convect_shallow_tend
// (convect_shallow.F90:616-660 and 790-836): device pointers in the driver's
// (pcols, ...) layout, CAM level order, all required.
typedef struct uwshcu_merge_fields {
    // in: this library's outputs for the chunk
    const double *umf;                    // cmfmc2 (pcols, pver+1)
    const double *slflx, *qtflx;          // (pcols, pver+1)
    const double *qrten, *qsten, *qcu;    // rprdsh, cmfdqs, iccmr_UW (pcols, pver)
    const double *qc2;                    // qc (pcols, pver)
    const double *rliq2, *cnt2, *cnb2, *cbmf;   // (pcols)
    // in: physics state and deep convection
    const double *pmid, *rprddp;          // (pcols, pver)
    // in/out: the deep-convection fields, merged in place
    double *cmfmc;                        // (pcols, pver+1): rows 1..pver
    double *cnt, *cnb, *rliq;             // (pcols)
    double *qc;                           // (pcols, pver)
    // out
    double *icwmr, *rprdsh, *rprdtot;     // (pcols, pver)
    double *cmfsl, *cmflq;                // (pcols, pver+1)
    double *freqsh, *pcnt, *pcnb;         // (pcols)
    double *massflx_sh;                   // (pcols): all pcols entries, as the driver copies them
} uwshcu_merge_fields;

// Return codes (negative values are conditions the reference treats as fatal).
enum {
    UWSHCU_OK                 = 0,
    UWSHCU_ERR_FINDSP_CONV    = -1,   // findsp not converging (uwshcu.F90:5419-5433)
    UWSHCU_ERR_FINDSP_ENTH    = -2,   // findsp enthalpy error (uwshcu.F90:5434-5449)
    UWSHCU_ERR_MUMIN2         = -3,   // mu > mumin2 (uwshcu.F90:2119-2122)
    UWSHCU_ERR_ARG            = -10,  // invalid sizes or null pointers
    UWSHCU_ERR_NOMEM          = -11,  // workspace allocation failed
    UWSHCU_ERR_HIP            = -12,  // a HIP runtime call failed
    UWSHCU_ERR_NO_DEVICE      = -13,  // no usable GPU
    UWSHCU_ERR_TAPE           = -14,  // adjoint: a column's tape exceeded its capacity
    UWSHCU_ERR_SCRATCH        = -15,  // allocation-free entry point: the chunk needs more than its static scratch
    UWSHCU_ERR_PARAMS         = -16   // drop-in called before uwshcu_set_dropin_params (uwshcu.F90:241-242)
};

#ifdef __cplusplus
}
#endif
#endif
