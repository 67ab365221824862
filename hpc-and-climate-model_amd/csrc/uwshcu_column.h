// uwshcu_column.h -- one UW shallow-cumulus column per GPU thread.
//
// Follows compute_uwshcu (uwshcu.F90:1218-4549) stage by stage for one column i,
// with the level flip of compute_uwshcu_inv (uwshcu.F90:380-414, 434-469) folded
// into the indexing: UW level k (1 = lowest layer, interfaces 0..mkx) maps to row
// mkx-k of the caller's CAM-ordered (pcols, pver[+1]) arrays.
//
// Memory layout.  Inputs and outputs are the caller's column-fastest arrays with
// leading dimension ldc, so a wavefront touches 64 consecutive doubles per level.
// Outputs double as working arrays where the reference's working array ends up in
// an output.  Other per-level state lives in a workspace of (ws_arrays * (mkx+1))
// rows of wld elements (column index fastest).
//
// Template parameter R is the type of the bulk arithmetic and of the workspace:
// double for the FP64 path, float for the mixed path.  The saturation adjustment
// (conden), the LCL inversion (qsinvert), the CIN integrals and the implicit-CIN
// weight, the findsp wet bulb and the positive-moisture fix-up always run in FP64.
#ifndef UWSHCU_COLUMN_H
#define UWSHCU_COLUMN_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "uwshcu_thermo.h"

// UWSHCU_BOUNDS_CHECK (host builds of the test harness only) asserts every index.
#ifdef UWSHCU_BOUNDS_CHECK
#include <cassert>
#define UW_CHECK(cond) assert(cond)
#else
#define UW_CHECK(cond) ((void)0)
#endif

namespace uwshcu {

// Workspace arrays (each mkx+1 rows).  The environment arrays are FP64 in both
// precisions: A_* hold the environment of the input state, B_* the environment
// re-derived from the iteration-1 provisional state (uwshcu.F90:4285-4327), used
// only for the second CIN evaluation; E_QVSTAR is the positive-moisture fix-up's
// provisional vapour.  They feed the CIN and cloud-base stages, which run in FP64.
// E_THV0MID/E_QS0MID hold thv and qs of the environment at layer midpoints
// (conden(p0, thl0, qt0)), which the buoyancy-sorting loop evaluates for every
// layer above the release level in every iteration; 0 marks "not yet evaluated",
// -1 a failed saturation adjustment.
enum EnvArray : int {
    A_THL0, A_QT0, A_SSTHL0, A_SSQT0, A_SSU0, A_SSV0, A_THV0BOT, A_THV0TOP, A_THVL0BOT, A_THVL0TOP,
    A_EXN0, A_EXNS0,
    B_THL0, B_QT0, B_SSTHL0, B_SSQT0, B_THV0BOT, B_THV0TOP, B_THVL0BOT, B_THVL0TOP,
    E_QVSTAR,
    E_THV0MID, E_QS0MID,   // saturation state of the layer-mean environment, filled when first used
    ENV_ARRAYS
};
// Plume and flux working arrays, in the bulk type R.
enum WsArray : int {
    W_EMF, W_THLU, W_QTU, W_UU, W_VU, W_THVU, W_UFRC, W_REI, W_FER, W_FDR, W_DWTEN, W_DITEN,
    W_THLUE, W_QTUE, W_UUE, W_VUE, W_UFLX, W_VFLX, W_QLSINK, W_QISINK, W_NLSINK, W_NISINK,
    W_QCL, W_QCI, W_QW0, W_TRFD, W_TRFU, W_QLR, W_QIR,
    WS_FIXED
};
// Per-constituent blocks follow the fixed plume arrays: SSTR0, TRU, TRUE, TRFLX (ncnst arrays each).
UWSHCU_HD inline int ws_arrays(int ncnst) { return WS_FIXED + 4 * ncnst; }
// Workspace bytes per column: environment block of the FP64 type (d_size bytes per
// value), then the R-typed plume block.
UWSHCU_HD inline size_t ws_bytes_per_column(int mkx, int ncnst, size_t r_size, size_t d_size = sizeof(double)) {
    return (size_t)(mkx + 1) * ((size_t)ENV_ARRAYS * d_size + (size_t)ws_arrays(ncnst) * r_size);
}

// Reverse-mode state of one launch (adjoint kernel only; null pointers otherwise).
struct AdArgs {
    AD* out;           // tape values of the outputs (layout: AdField below)
    TapeCtx* tapes;    // one per workspace slot
};

struct FwdArgs {
    int ncol, mkx, ncnst, ldc;
    int col0;            // first column of this launch; workspace column = i - col0
    long long wld;
    double dt;
    uwshcu_params par;
    const double *ps0, *zs0, *p0, *z0, *dp0, *dpdry0, *u0, *v0, *qv0, *ql0, *qi0, *t0, *s0, *tr0, *tke, *pblh;
    double* cush;
    double *umf, *slflx, *qtflx, *flxprc1, *flxsnow1, *qvten, *qlten, *qiten, *sten, *uten, *vten, *trten;
    double *qrten, *qsten, *precip, *snow, *evapc, *cufrc, *qcu, *qlu, *qiu, *cbmf, *qc, *rliq, *cnt, *cnb, *wu;
    int* diag;          // optional [UWSHCU_NDIAG][ldc]
    int* err;           // [2]: first fatal code, column
    void* ws;           // workspace (see ws_bytes_per_column)
    // Column selection.  With colmap == nullptr, workspace slot j serves column col0 + j.
    // Otherwise slot j serves column colmap[col0 + j] for col0 + j < *colcount.
    const int* colmap;
    const int* colcount;
    const double* cush_in;   // optional: read the input cush here instead of from cush
    int* amb_flag;           // optional [ldc]: 1 where an FP32 branch decision was within rounding
    int env_ready;           // 1: setup_level1/2 already filled the set-A environment and the wet bulb
    // optional: a column whose implicit-CIN pass needs a second convection pass stops
    // there, saves its scalars (NRESUME doubles, [field][slot], wld apart) in
    // defer_state, leaves its outputs to be rewritten and appends itself to defer_list
    int* defer_list;
    int* defer_count;
    double* defer_state;
    // 1 (with defer_list): a column that reaches the cloud-base closure of its first
    // pass stops there instead, saves its scalars and lists itself, so that only the
    // triggered columns run the convection stages, packed densely in a launch of their own
    int defer_at_closure;
    int* defer_mark;         // optional [wld]: mark the deferred column's slot 1 instead of listing it
    // 1: continue each listed column (colmap) at the closure of its first pass; 2: at its
    // second convection pass.  Each runs in the workspace slot it had in the deferring
    // launch (column - slot0), whose set-A environment is intact.
    int resume;
    int slot0;
    // 1: colmap lists a few columns of a large range (the mixed mode's FP64 recompute):
    // run them with the one-wavefront build, unsplit
    int sparse_map;
    AdArgs ad;               // adjoint kernel only
    // the modified scheme (Column<..., true>): its options, layer omega (caller layout)
    // and the per-column perturbations tpert and qpert
    uwshcu_zw_params zw;
    const double *omega, *tpert, *qpert;
};

// Level loops whose first level differs between columns (the plume from krel, the
// penetrative entrainment down from kpen-1) start every lane of a wavefront at the
// wavefront's extreme first level; a lane runs the loop body only from its own first
// level on, so its results are those of its own loop.  The lanes then address the same
// level of the level-major arrays in each iteration, one 64-column row for the
// wavefront, where lanes at different levels would each read a row of their own.
// Returns the smallest k among the wavefront's active lanes (the lane's own k on the host).
UWSHCU_HD inline int wave_min_level(int k) {
#if defined(__HIP_DEVICE_COMPILE__)
    int kw = 0;
    while (__ballot(k <= kw) == 0ull) ++kw;   // ends at this lane's k at the latest
    return kw;
#else
    return k;
#endif
}
UWSHCU_HD inline int wave_max_level(int k) {
#if defined(__HIP_DEVICE_COMPILE__)
    int kw = k;
    while (__ballot(k > kw) != 0ull) ++kw;    // ends at the largest k of the active lanes
    return kw;
#else
    return k;
#endif
}

// Parts (bit PART_*) whose level loops align their lanes (wave_min_level).  Alignment
// makes a wavefront's accesses coalesce, but the lanes idle until their own first level.
// Measured at 1 036 800 x 35 x 3: aligned, the buoyancy sorting moves 29 GB instead of
// 49 GB but takes 0.7 ms longer, and the one-wavefront-per-SIMD build loses 10 % at
// 32768 columns; the stages after the buoyancy sorting, which wait on memory most of the
// time, gain.  Fastest: those alone (51.4 ms against 52.0 ms unaligned, 52.1 ms aligned
// in every split part).
#ifndef UWSHCU_ALIGN_PARTS
#define UWSHCU_ALIGN_PARTS (1 << PART_REST)
#endif

// zmh_ramp (mzfunctions_mod.F90:222-248): 0 below x0, 1 above x0 + xint, linear in
// between; the floor on xint is the single-precision literal 1.e-20 there.
UWSHCU_HD inline double zmh_ramp(double x, double x0, double xint) {
    double x2;
    if (x < x0) x2 = 0.0;
    else if (x > x0 + xint) x2 = 1.0;
    else x2 = (x - x0) / fmax(xint, (double)1.e-20f);
    return fmin(x2, 1.0);
}

// Reverse-mode state of one launch (adjoint kernel only).  Output fields are held as
// tape values while the column runs: field f, level k of workspace slot j lives at
// out[(f * (mkx+1) + k) * wld + j]; fields are ADF_* below, then trten(m) at
// ADF_TR + m - 1, then the column scalars at ADF_TR + ncnst + ADC_*.
enum AdField : int {
    ADF_UMF, ADF_SLFLX, ADF_QTFLX, ADF_FLXPRC1, ADF_FLXSNOW1, ADF_WU,
    ADF_QVTEN, ADF_QLTEN, ADF_QITEN, ADF_STEN, ADF_UTEN, ADF_VTEN, ADF_QRTEN, ADF_QSTEN, ADF_EVAPC,
    ADF_CUFRC, ADF_QCU, ADF_QLU, ADF_QIU, ADF_QC,
    ADF_TR
};
enum AdColumnField : int { ADC_PRECIP, ADC_SNOW, ADC_CBMF, ADC_RLIQ, ADC_CNT, ADC_CNB, ADC_CUSH, ADC_COUNT };
UWSHCU_HD inline int ad_fields(int ncnst) { return ADF_TR + ncnst + ADC_COUNT; }
// Independent variables: node f*(mkx+1)+k for t0, qv0, u0, v0 (f = 0..3, k = 1..mkx)
// and tke (f = 4, k = 0..mkx), then rpen, rle, rkm.
enum AdInput : int { ADI_T0, ADI_QV0, ADI_U0, ADI_V0, ADI_TKE, ADI_PARAMS };
UWSHCU_HD inline int ad_inputs(int mkx) { return ADI_PARAMS * (mkx + 1) + 3; }

// Exit reasons reported in diag row 0 (same numbering as the oracle).
// Scalars of a column deferred at its second convection pass (FwdArgs::defer_state).
// Field RS_CONT (after the 56 scalars) is set by the buoyancy-sorting and the
// first-pass-rest parts of a split launch: 1 where the column continues in the next
// part, 0 where it ended there.
enum { NRESUME = 57, RS_CONT = 56 };

// Ints of a split launch's list storage for n columns (launch_column_pipeline).
UWSHCU_HD inline size_t defer_list_ints(size_t n) { return 3 * n + (n + 255) / 256 + 8; }

// Parts of the column pass that the split pipeline runs as launches of their own.
// PART_ALL runs whatever FwdArgs::resume asks for; the others fix it at compile time,
// so that each kernel holds only its own part's live state.
enum : int {
    PART_ALL = 0,
    PART_SORT = 1,      // a listed column from its closure through the buoyancy sorting (S11)
    PART_REST = 2,      // the same from the forcedCu filter (S12) through the in-cloud
                        // diagnostics (S21)
    PART_CLOSURE = 3,   // every column up to the cloud-base closure of its first pass
    PART_CIN2 = 4       // the same, after the level-parallel set-B environment (S22), through
                        // the implicit-CIN iteration's CIN (S1-S7), deferring a second pass
};

enum { EX_NONE = 0, EX_CONDEN = 1, EX_KINV1 = 2, EX_KLCLMKX = 3, EX_KLFCMKX = 4, EX_MU3 = 5, EX_WTW = 6,
       EX_UFRC = 7, EX_WU = 8, EX_CUFILTER = 9, EX_SUBCONDEN = 10, EX_KBUPCONDEN = 11 };

// Every helper of the column pass is expanded in place.  A call would pass the column
// object, and with it the kernel arguments, by address: the compiler then keeps a copy
// of the arguments in scratch, reloads each array pointer from it, and, unable to tell
// global from LDS memory behind those pointers, issues flat instructions.  Expanded,
// the pointers stay in scalar registers and every access is a global or LDS one
// (scratch 1.2 KB -> 44 B per lane in the one-wavefront build).
#define UW_INLINE __attribute__((always_inline))

// R: bulk (plume) arithmetic type; D: the FP64 type (double, or AD in the adjoint).
// ZW: the modified scheme compute_zwshcu (zwshcu.F90:630-4969), FP64 only; its
// differences from compute_uwshcu are marked where they apply.
// FAST: the fast path's saturation adjustment (ThermoT<D, true>), FP64 only.
template <class R, class D = double, bool ZW = false, bool FAST = false>
struct Column {
    static_assert(!ZW || (std::is_same<R, double>::value && std::is_same<D, double>::value),
                  "the modified scheme runs in FP64");
    static constexpr bool kAD = is_ad<D>::value;
    const FwdArgs& a;
    const ThermoT<D, FAST>& T;
    const uwshcu_consts& C;
    const int i, iw, mkx, KS, ncnst;
    const size_t ldc, wld;
    D* __restrict__ env;        // ENV_ARRAYS * KS rows of wld D values
    R* __restrict__ ws;         // ws_arrays(ncnst) * KS rows of wld R values
    TapeCtx* tape;              // adjoint kernel: this column's tape

    UWSHCU_HD UW_INLINE Column(const FwdArgs& args, const ThermoT<D, FAST>& th, int col, int slot, TapeCtx* own_tape = nullptr)
        : a(args), T(th), C(*th.c), i(col), iw(slot), mkx(args.mkx), KS(args.mkx + 1), ncnst(args.ncnst),
          ldc((size_t)args.ldc), wld((size_t)args.wld), env(reinterpret_cast<D*>(args.ws)),
          ws(reinterpret_cast<R*>(reinterpret_cast<D*>(args.ws) + (size_t)ENV_ARRAYS * (size_t)(args.mkx + 1) *
                                                                    (size_t)args.wld)),
          tape(own_tape ? own_tape : args.ad.tapes ? args.ad.tapes + slot : nullptr) {
        if constexpr (kAD) ad_bind_tape(tape);
    }

    // saved scalar f of a deferred column (FwdArgs::defer_state)
    UWSHCU_HD inline double& rstate(int f) const {
        UW_CHECK(f >= 0 && f < NRESUME && iw >= 0 && (size_t)iw < wld && a.defer_state);
        return a.defer_state[(size_t)f * wld + iw];
    }
    // caller arrays: UW level k <-> CAM row mkx-k (layers k=1..mkx and interfaces k=0..mkx)
    UWSHCU_HD inline size_t row(int k) const { return (size_t)(mkx - k) * ldc + i; }
#ifdef UWSHCU_BOUNDS_CHECK
    UWSHCU_HD inline bool iface(const double* p) const {
        return p == a.ps0 || p == a.zs0 || p == a.tke || p == a.umf || p == a.slflx || p == a.qtflx ||
               p == a.flxprc1 || p == a.flxsnow1 || p == a.wu;
    }
#endif
    // tape value of output field f at level k (adjoint kernel)
    UWSHCU_HD inline D& adout(int f, int k) const {
        if constexpr (kAD) return a.ad.out[((size_t)f * KS + k) * wld + iw];
        else return *(D*)nullptr;
    }
    UWSHCU_HD inline int ad_field(const double* p) const {
        const double* const f[] = {a.umf, a.slflx, a.qtflx, a.flxprc1, a.flxsnow1, a.wu, a.qvten, a.qlten, a.qiten,
                                   a.sten, a.uten, a.vten, a.qrten, a.qsten, a.evapc, a.cufrc, a.qcu, a.qlu, a.qiu,
                                   a.qc};
        for (int q = 0; q < ADF_TR; ++q)
            if (p == f[q]) return q;
        if (p == a.precip) return ADF_TR + ncnst + ADC_PRECIP;
        if (p == a.snow) return ADF_TR + ncnst + ADC_SNOW;
        if (p == a.cbmf) return ADF_TR + ncnst + ADC_CBMF;
        if (p == a.rliq) return ADF_TR + ncnst + ADC_RLIQ;
        if (p == a.cnt) return ADF_TR + ncnst + ADC_CNT;
        if (p == a.cnb) return ADF_TR + ncnst + ADC_CNB;
        return ADF_TR + ncnst + ADC_CUSH;
    }
    UWSHCU_HD inline D in(const double* p, int k) const {
        UW_CHECK(k <= mkx && k >= (iface(p) ? 0 : 1) && i >= 0 && i < a.ncol);
        if constexpr (kAD) {
            const int base = p == a.t0 ? ADI_T0 : p == a.qv0 ? ADI_QV0 : p == a.u0 ? ADI_U0 : p == a.v0 ? ADI_V0
                           : p == a.tke ? ADI_TKE : -1;
            if (base >= 0) return D(p[row(k)], base * KS + k);
        }
        return D(p[row(k)]);
    }
    // t0 and qv0 of layer k as the column physics sees them.  The modified scheme's
    // wrapper (zwshcu.F90:547-566) adds tpert / qpert, capped, to t0 / qv0 in every
    // layer whose bottom interface is not above pblh (and cp*tpert to s0, which the
    // column physics never reads).
    UWSHCU_HD inline D in_t0(int k) const {
        const D v = in(a.t0, k);
        if constexpr (ZW) {
            if (a.zw.pertt && !(a.zs0[row(k - 1)] > a.pblh[i])) return v + fmin(a.tpert[i], a.zw.tpertmax);
        }
        return v;
    }
    UWSHCU_HD inline D in_qv0(int k) const {
        const D v = in(a.qv0, k);
        if constexpr (ZW) {
            if (a.zw.pertt && !(a.zs0[row(k - 1)] > a.pblh[i])) return v + fmin(a.qpert[i], a.zw.qpertmax);
        }
        return v;
    }
    UWSHCU_HD inline D& out(double* p, int k) const {
        UW_CHECK(k <= mkx && k >= (iface(p) ? 0 : 1) && i >= 0 && i < a.ncol);
        if constexpr (kAD) return adout(ad_field(p), k);
        else return p[row(k)];
    }
    // column scalars (precip, snow, cbmf, rliq, cnt, cnb, cush)
    UWSHCU_HD inline D& cout(double* p) const {
        if constexpr (kAD) return adout(ad_field(p), 0);
        else return p[i];
    }
    UWSHCU_HD inline D trin(int m, int k) const {   // m 1-based
        UW_CHECK(m >= 1 && m <= ncnst && k >= 1 && k <= mkx);
        return D(a.tr0[((size_t)(m - 1) * mkx + (mkx - k)) * ldc + i]);
    }
    UWSHCU_HD inline D& trout(int m, int k) const {
        UW_CHECK(m >= 1 && m <= ncnst && k >= 1 && k <= mkx);
        if constexpr (kAD) return adout(ADF_TR + m - 1, k);
        else return a.trten[((size_t)(m - 1) * mkx + (mkx - k)) * ldc + i];
    }
    // tunables: j = 0 rpen, 1 rle, 2 rkm
    UWSHCU_HD inline D param(int j) const {
        const double v = j == 0 ? a.par.rpen : j == 1 ? a.par.rle : a.par.rkm;
        if constexpr (kAD) return D(v, ADI_PARAMS * KS + j);
        else return v;
    }
    UWSHCU_HD inline D& E(int arr, int k) const {
        UW_CHECK(arr >= 0 && arr < ENV_ARRAYS && k >= 0 && k <= mkx && iw >= 0 && (long long)iw < (long long)wld);
        return env[((size_t)arr * KS + k) * wld + iw];
    }
    UWSHCU_HD inline R Er(int arr, int k) const { return R(E(arr, k)); }
    UWSHCU_HD inline R& W(int arr, int k) const {
        UW_CHECK(arr >= 0 && arr < WS_FIXED && k >= 0 && k <= mkx && iw >= 0 && (long long)iw < (long long)wld);
        return ws[((size_t)arr * KS + k) * wld + iw]; }
    UWSHCU_HD inline R& WT(int blk, int m, int k) const {   // blk 0..3, m 1-based
        UW_CHECK(blk >= 0 && blk < 4 && m >= 1 && m <= ncnst && k >= 0 && k <= mkx && iw >= 0 && (long long)iw < (long long)wld);
        return ws[((size_t)(WS_FIXED + blk * ncnst + (m - 1)) * KS + k) * wld + iw];
    }

    UWSHCU_HD inline R thv_of(D thj, D qvj, D qlj, D qij) const {
        return R(thj * (1.0 + C.zvir * qvj - qlj - qij));
    }

    // minmod slopes (uwshcu.F90:4831-4861) of field(k) -> E(dst, k), k = 1..mkx, in FP64
    template <class F>
    UWSHCU_HD inline void slope(F field, int dst) const {
        D below = (field(2) - field(1)) / (in(a.p0, 2) - in(a.p0, 1));
        for (int k = 2; k <= mkx; ++k) {
            const D above = (field(k) - field(k - 1)) / (in(a.p0, k) - in(a.p0, k - 1));
            E(dst, k - 1) = (above > 0.0) ? fmax(0.0, fmin(above, below)) : fmin(0.0, fmax(above, below));
            below = above;
        }
        E(dst, mkx) = E(dst, mkx - 1);
    }
    template <class F>
    UWSHCU_HD inline void slope_tr(F field, int m) const {
        R below = (R(field(2)) - R(field(1))) / (R(in(a.p0, 2)) - R(in(a.p0, 1)));
        for (int k = 2; k <= mkx; ++k) {
            const R above = (R(field(k)) - R(field(k - 1))) / (R(in(a.p0, k)) - R(in(a.p0, k - 1)));
            WT(0, m, k - 1) = (above > R(0)) ? fmax(R(0), fmin(above, below)) : fmin(R(0), fmax(above, below));
            below = above;
        }
        WT(0, m, mkx) = WT(0, m, mkx - 1);
    }

    // Environment at layer interfaces (uwshcu.F90:1275-1299, 4303-4327) for set base eb
    // (A_THL0 or B_THL0), from thl0/qt0 already stored in that set.  Returns false on a
    // conden failure.
    UWSHCU_HD UW_INLINE bool interface_env(int eb) const {
        for (int k = 1; k <= mkx; ++k)
            if (!interface_env_level(eb, k)) return false;
        return true;
    }
    UWSHCU_HD UW_INLINE bool interface_env_level(int eb, int k) const {
        const int thl = eb, qt = eb + 1, ssthl = eb + 2, ssqt = eb + 3;
        const int thvb = eb == A_THL0 ? A_THV0BOT : B_THV0BOT, thvt = thvb + 1;
        const int thvlb = eb == A_THL0 ? A_THVL0BOT : B_THVL0BOT, thvlt = thvlb + 1;
        {
            D th, qv, ql, qi;
            const D pb = in(a.ps0, k - 1), pt = in(a.ps0, k), pm = in(a.p0, k);
            const D thl0bot = E(thl, k) + E(ssthl, k) * (pb - pm);
            const D qt0bot = E(qt, k) + E(ssqt, k) * (pb - pm);
            if (T.conden_x(pb, E(A_EXNS0, k - 1), thl0bot, qt0bot, th, qv, ql, qi)) return false;
            E(thvb, k) = th * (1.0 + C.zvir * qv - ql - qi);
            E(thvlb, k) = thl0bot * (1.0 + C.zvir * qt0bot);
            const D thl0top = E(thl, k) + E(ssthl, k) * (pt - pm);
            const D qt0top = E(qt, k) + E(ssqt, k) * (pt - pm);
            if (T.conden_x(pt, E(A_EXNS0, k), thl0top, qt0top, th, qv, ql, qi)) return false;
            E(thvt, k) = th * (1.0 + C.zvir * qv - ql - qi);
            E(thvlt, k) = thl0top * (1.0 + C.zvir * qt0top);
        }
        return true;
    }

    // Saturation state of the layer-k mean environment, evaluated once per column
    // (uwshcu.F90:2390-2397 evaluate it in every buoyancy-sorting iteration with the
    // same arguments).  Returns false when the saturation adjustment fails.
    UWSHCU_HD UW_INLINE bool env_mid(int k, D& thv0j, D& qs0) const {
        D& tv = E(E_THV0MID, k);
        if (tv == 0.0) {
            D th, qv, ql, qi, es, qs;
            if (T.conden_x(in(a.p0, k), E(A_EXN0, k), E(A_THL0, k), E(A_QT0, k), th, qv, ql, qi)) {
                tv = -1.0;
            } else {
                tv = th * (1.0 + C.zvir * qv - ql - qi);
                T.qsat_es_qs(E(A_THL0, k) * E(A_EXN0, k), in(a.p0, k), es, qs);
                E(E_QS0MID, k) = qs;
            }
        }
        thv0j = tv;
        qs0 = E(E_QS0MID, k);
        return tv > 0.0;
    }

    // Zero the column's outputs and the working arrays the reference zeroes before a
    // pass (uwshcu.F90:1333-1408, 1817-1869).
    UWSHCU_HD UW_INLINE void zero_pass_state() const {
        for (int k = 0; k <= mkx; ++k) {
            out(a.umf, k) = 0.0;
            out(a.wu, k) = 0.0;
            out(a.slflx, k) = 0.0;
            out(a.qtflx, k) = 0.0;
            W(W_EMF, k) = R(0);
            W(W_UFRC, k) = R(0);
            W(W_UFLX, k) = R(0);
            W(W_VFLX, k) = R(0);
            W(W_THLU, k) = R(0);
            W(W_QTU, k) = R(0);
            W(W_UU, k) = R(0);
            W(W_VU, k) = R(0);
            W(W_THVU, k) = R(0);
            for (int m = 1; m <= ncnst; ++m) {
                WT(1, m, k) = R(0);
                WT(3, m, k) = R(0);
            }
        }
        for (int k = 1; k <= mkx; ++k) {
            out(a.qvten, k) = 0.0; out(a.qlten, k) = 0.0; out(a.qiten, k) = 0.0; out(a.sten, k) = 0.0;
            out(a.uten, k) = 0.0; out(a.vten, k) = 0.0; out(a.qrten, k) = 0.0; out(a.qsten, k) = 0.0;
            out(a.evapc, k) = 0.0; out(a.cufrc, k) = 0.0; out(a.qcu, k) = 0.0; out(a.qlu, k) = 0.0;
            out(a.qiu, k) = 0.0; out(a.qc, k) = 0.0;
            W(W_FER, k) = R(0); W(W_FDR, k) = R(0); W(W_DWTEN, k) = R(0); W(W_DITEN, k) = R(0);
            W(W_QCL, k) = R(0); W(W_QCI, k) = R(0);
            for (int m = 1; m <= ncnst; ++m) trout(m, k) = 0.0;
        }
    }

    UWSHCU_HD UW_INLINE void write_exit(int mk) const {
        // uwshcu.F90:4448-4482; flxprc1/flxsnow1 keep their initial zero (uwshcu.F90:1090-1091)
        for (int k = 0; k <= mk; ++k) {
            out(a.umf, k) = 0.0; out(a.wu, k) = 0.0; out(a.slflx, k) = 0.0; out(a.qtflx, k) = 0.0;
            out(a.flxprc1, k) = 0.0; out(a.flxsnow1, k) = 0.0;
        }
        for (int k = 1; k <= mk; ++k) {
            out(a.qvten, k) = 0.0; out(a.qlten, k) = 0.0; out(a.qiten, k) = 0.0; out(a.sten, k) = 0.0;
            out(a.uten, k) = 0.0; out(a.vten, k) = 0.0; out(a.qrten, k) = 0.0; out(a.qsten, k) = 0.0;
            out(a.evapc, k) = 0.0; out(a.cufrc, k) = 0.0; out(a.qcu, k) = 0.0; out(a.qlu, k) = 0.0;
            out(a.qiu, k) = 0.0; out(a.qc, k) = 0.0;
            for (int m = 1; m <= ncnst; ++m) trout(m, k) = 0.0;
        }
        cout(a.precip) = 0.0; cout(a.snow) = 0.0; cout(a.cush) = -1.0; cout(a.cbmf) = 0.0; cout(a.rliq) = 0.0;
        cout(a.cnt) = D(mk + 1) - 1.0;        // cnt = 1 flipped: mkx+1-1
        cout(a.cnb) = D(mk + 1) - D(mk); // cnb = mkx flipped
    }

    UWSHCU_HD void run() const { run_body<true>(); }
    template <int PART>
    UWSHCU_HD void run_part() const { run_body<true, PART>(); }
    // the column pass itself, expanded into each forward kernel so that each gets its
    // own register allocation (the two builds differ in their register budget).
    // kSplit = false leaves out what only the launches of a split large launch do (stop
    // at, and continue from, the closure, the buoyancy sorting or the second pass),
    // which the one-wavefront-per-SIMD build never runs.  PART (PART_*) fixes which part
    // of a split launch runs.
    template <bool kSplit, int PART = PART_ALL>
    UWSHCU_HD __attribute__((always_inline)) void run_body() const;

    // Level-parallel set-up (one thread per column and level), the part of S0 and of
    // the wet bulb that needs no sequential sweep.  Stage 1, level k = 0..mkx: Exner
    // functions, thl0, qt0, findsp.  Stage 2, level k = 1..mkx, after stage 1 of the
    // neighbouring levels: slopes, the set-A interface environment (a failed
    // saturation adjustment marks A_THV0BOT(k) = -1) and the layer-mean saturation
    // cache.  Every value equals the one run() computes itself without env_ready.
    UWSHCU_HD void setup_level1(int k) const {
        E(A_EXNS0, k) = upow(in(a.ps0, k) / C.p00, C.rovcp);
        E(E_THV0MID, k) = 0.0;
        if (k < 1) return;
        D qsp, enerr;
        const int doit = findsp_point(T, in_qv0(k), in_t0(k), in(a.p0, k), qsp, enerr);
        if (doit == 0) record_error(a.err, UWSHCU_ERR_FINDSP_CONV, i);
        else if (doit == 2 && enerr > 1.e-4) record_error(a.err, UWSHCU_ERR_FINDSP_ENTH, i);
        W(W_QW0, k) = R(qsp);
        E(A_EXN0, k) = upow(in(a.p0, k) / C.p00, C.rovcp);
        const D qv0 = in_qv0(k), ql0 = in(a.ql0, k), qi0 = in(a.qi0, k), t0 = in_t0(k);
        const D qt0 = (qv0 + ql0 + qi0);
        E(A_QT0, k) = qt0;
        E(A_THL0, k) = (t0 - C.xlv * ql0 / C.cp - C.xls * qi0 / C.cp) / E(A_EXN0, k);
    }
    UWSHCU_HD void setup_level2(int k) const {
        // minmod slope at level k as slope() computes it (the top level copies mkx-1)
        const int j = k < mkx ? k : mkx - 1;
        auto mm = [&](auto field) {
            const auto above = (field(j + 1) - field(j)) / (in(a.p0, j + 1) - in(a.p0, j));
            const auto below = (j == 1) ? above : (field(j) - field(j - 1)) / (in(a.p0, j) - in(a.p0, j - 1));
            using V = decltype(above + below);
            return (above > V(0)) ? fmax(V(0), fmin(above, below)) : fmin(V(0), fmax(above, below));
        };
        E(A_SSTHL0, k) = mm([&](int q) { return E(A_THL0, q); });
        E(A_SSQT0, k) = mm([&](int q) { return E(A_QT0, q); });
        E(A_SSU0, k) = mm([&](int q) { return in(a.u0, q); });
        E(A_SSV0, k) = mm([&](int q) { return in(a.v0, q); });
        for (int m = 1; m <= ncnst; ++m) {
            const auto above = (R(trin(m, j + 1)) - R(trin(m, j))) / (R(in(a.p0, j + 1)) - R(in(a.p0, j)));
            const auto below = (j == 1) ? above
                                        : (R(trin(m, j)) - R(trin(m, j - 1))) / (R(in(a.p0, j)) - R(in(a.p0, j - 1)));
            WT(0, m, k) = (above > R(0)) ? fmax(R(0), fmin(above, below)) : fmin(R(0), fmax(above, below));
        }
        if (!interface_env_level(A_THL0, k)) E(A_THV0BOT, k) = -1.0;
        D thv0j, qs0;
        env_mid(k, thv0j, qs0);
    }

    // S22 of a column's first pass, level-parallel (uwshcu.F90:4169-4329), after the
    // first pass's tendencies are in place: stage 1, level k = 1..mkx, the provisional
    // state's thl and qt (set B); stage 2, level k = 1..mkx, after stage 1 of the
    // neighbouring levels, their slopes and the set-B interface environment (a failed
    // saturation adjustment marks B_THV0BOT(k) = -1, which ends the column as the
    // sequential S22 would).  Every value equals the one run() computes in S22.
    UWSHCU_HD void setup_level_b1(int k) const {
        const D qv0s = in_qv0(k) + out(a.qvten, k) * a.dt;
        const D ql0s = in(a.ql0, k) + out(a.qlten, k) * a.dt;
        const D qi0s = in(a.qi0, k) + out(a.qiten, k) * a.dt;
        const D t0s = in_t0(k) + out(a.sten, k) * a.dt / C.cp;
        const D qt0 = (qv0s + ql0s + qi0s);
        E(B_QT0, k) = qt0;
        E(B_THL0, k) = (t0s - C.xlv * ql0s / C.cp - C.xls * qi0s / C.cp) / E(A_EXN0, k);
    }
    UWSHCU_HD void setup_level_b2(int k) const {
        const int j = k < mkx ? k : mkx - 1;
        auto mm = [&](int arr) {
            const D above = (E(arr, j + 1) - E(arr, j)) / (in(a.p0, j + 1) - in(a.p0, j));
            const D below = (j == 1) ? above : (E(arr, j) - E(arr, j - 1)) / (in(a.p0, j) - in(a.p0, j - 1));
            return (above > 0.0) ? fmax(0.0, fmin(above, below)) : fmin(0.0, fmax(above, below));
        };
        E(B_SSTHL0, k) = mm(B_THL0);
        E(B_SSQT0, k) = mm(B_QT0);
        if (!interface_env_level(B_THL0, k)) E(B_THV0BOT, k) = -1.0;
    }

    // Adjoint kernel, after run(): copy the recorded outputs' values to the caller's
    // arrays and their tape nodes to nodes[(f * (mkx+1) + k) * wld + slot].
    UWSHCU_HD UW_INLINE void ad_export(int* nodes) const {
        if constexpr (kAD) {
            auto put = [&](int f, int k, double* dst) {
                const D& v = adout(f, k);
                *dst = v.v;
                nodes[((size_t)f * KS + k) * wld + iw] = v.i;
            };
            double* const fields[] = {a.umf, a.slflx, a.qtflx, a.flxprc1, a.flxsnow1, a.wu, a.qvten, a.qlten,
                                      a.qiten, a.sten, a.uten, a.vten, a.qrten, a.qsten, a.evapc, a.cufrc, a.qcu,
                                      a.qlu, a.qiu, a.qc};
            for (int f = 0; f < ADF_TR; ++f)
                for (int k = (f <= ADF_WU ? 0 : 1); k <= mkx; ++k) put(f, k, &fields[f][row(k)]);
            for (int m = 1; m <= ncnst; ++m)
                for (int k = 1; k <= mkx; ++k)
                    put(ADF_TR + m - 1, k, &a.trten[((size_t)(m - 1) * mkx + (mkx - k)) * ldc + i]);
            double* const scal[] = {a.precip, a.snow, a.cbmf, a.rliq, a.cnt, a.cnb, a.cush};
            for (int q = 0; q < ADC_COUNT; ++q) put(ADF_TR + ncnst + q, 0, &scal[q][i]);
        } else {
            (void)nodes;
        }
    }
};

template <class R, class D, bool ZW, bool FAST>
template <bool kSplit, int PART>
UWSHCU_HD __attribute__((always_inline)) void Column<R, D, ZW, FAST>::run_body() const {
    static_assert(PART == PART_ALL || kSplit, "the parts of a split launch need the split build");
    // what this launch continues from: 0 the start, 1 the closure of the first pass,
    // 2 the second convection pass, 3 the forcedCu filter of the first pass, 4 the
    // implicit-CIN iteration
    const int resume = PART == PART_SORT ? 1 : PART == PART_REST ? 3 : PART == PART_CLOSURE ? 0
                     : PART == PART_CIN2 ? 4 : a.resume;
    if constexpr (PART == PART_REST || PART == PART_CIN2) {
        if (rstate(RS_CONT) == 0.0) return;      // the column ended in the previous part
    }
    // level loops with per-column first levels: level-aligned lanes (wave_min_level) in
    // the parts of UWSHCU_ALIGN_PARTS (bit PART), each lane's own first level elsewhere
    constexpr bool kAlign = ((UWSHCU_ALIGN_PARTS) >> PART) & 1;
    auto lvl_lo = [](int k) { return kAlign ? wave_min_level(k) : k; };
    auto lvl_hi = [](int k) { return kAlign ? wave_max_level(k) : k; };
    // uwshcu.F90:1040-1068 parameters
    const R rmaxfrac = R(0.10), rbuoy = R(1), rdrag = R(1);
    const R PGFc = R(0.7), criqc = R(0.7e-3), frc_rasn = R(1), kevp = R(2.e-6);
    const R xlv = R(C.xlv), xls = R(C.xls), cp = R(C.cp), zvir = R(C.zvir), r = R(C.r), g = R(C.g);
    const R rle = R(param(1)), rkm = R(param(2)), rpen = R(param(0));
    const R dt = R(a.dt);
    const double dtd = a.dt;
    const int ixl = C.ixnumliq, ixi = C.ixnumice;
    const bool has_nl = ixl >= 1 && ixl <= ncnst, has_ni = ixi >= 1 && ixi <= ncnst;
    auto trz = [&](int m, int k) -> R { return (m >= 1 && m <= ncnst) ? R(trin(m, k)) : R(0); };

    // Branch guard: a decision the reference takes on a quantity computed here in FP32,
    // or through the fast path's converged saturation adjustment, is flagged when the
    // quantity lies within a relative 1e-5 (FP32) or 1e-8 (fast path, whose results
    // differ from the reference's by about 1e-10) of the branch point, well above the
    // perturbation it carries.  Flagged columns are recomputed in FP64 by the
    // mixed-precision driver.  No effect on the FP64 path.
    constexpr bool kFp32 = sizeof(R) == sizeof(float);
    constexpr double kGuard = kFp32 ? 1.e-5 : 1.e-8;
    int amb = 0;
    auto near = [&](auto dist, auto scale) {
        if constexpr (kFp32 || FAST) {
            if (fabs(dist) <= kGuard * fabs(scale)) amb = 1;
        }
    };

    int exit_code = EX_NONE, path = 0, exit_iter = 0;
    int kinv = 0, klcl = 0, klfc = 0, krel = 0, kbup = 0, kpen = 0;
    int limit_ufrc = 0, limit_ppen = 0, limit_emf = 0;
    D th, qvj, qlj, qij;

    // ---- wet bulb for the evaporation limiter (uwshcu.F90:1216; findsp 5164-5454) ----
    if (!a.env_ready)
    for (int k = 1; k <= mkx; ++k) {
        D qsp, enerr;
        const int doit = findsp_point(T, in_qv0(k), in_t0(k), in(a.p0, k), qsp, enerr);
        if (doit == 0) record_error(a.err, UWSHCU_ERR_FINDSP_CONV, i);
        else if (doit == 2 && enerr > 1.e-4) record_error(a.err, UWSHCU_ERR_FINDSP_ENTH, i);
        W(W_QW0, k) = R(qsp);
    }

    // ---- S0 environment, set A (uwshcu.F90:1255-1299) ----
    if (!a.env_ready) {
    for (int k = 1; k <= mkx; ++k) E(A_EXN0, k) = upow(in(a.p0, k) / C.p00, C.rovcp);
    for (int k = 0; k <= mkx; ++k) E(A_EXNS0, k) = upow(in(a.ps0, k) / C.p00, C.rovcp);
    for (int k = 1; k <= mkx; ++k) {
        const D qv0 = in_qv0(k), ql0 = in(a.ql0, k), qi0 = in(a.qi0, k), t0 = in_t0(k);
        const D qt0 = (qv0 + ql0 + qi0);
        E(A_QT0, k) = qt0;
        E(A_THL0, k) = (t0 - C.xlv * ql0 / C.cp - C.xls * qi0 / C.cp) / E(A_EXN0, k);
    }
    slope([&](int k) { return E(A_THL0, k); }, A_SSTHL0);
    slope([&](int k) { return E(A_QT0, k); }, A_SSQT0);
    slope([&](int k) { return in(a.u0, k); }, A_SSU0);
    slope([&](int k) { return in(a.v0, k); }, A_SSV0);
    for (int k = 0; k <= mkx; ++k) E(E_THV0MID, k) = 0.0;
    for (int m = 1; m <= ncnst; ++m) slope_tr([&](int k) { return R(trin(m, k)); }, m);
    }

    // a column continued at its first pass keeps what launch 1 wrote; launch 1 of a split
    // clears a column's outputs and plume arrays only when it reaches the release level
    // (S1-S7 read neither, and every exit rewrites all outputs)
    if (PART == PART_ALL && (!kSplit || !resume)) {
        zero_pass_state();
        for (int k = 0; k <= mkx; ++k) { out(a.flxprc1, k) = 0.0; out(a.flxsnow1, k) = 0.0; }
    }

    D cush = D(a.cush_in ? a.cush_in[i] : a.cush[i]);
    const D pblh = D(a.pblh[i]);

    // iteration-1 values restored for the implicit-CIN pass (uwshcu.F90:1693-1713)
    int kinv_o = 0, klcl_o = 0, klfc_o = 0;
    D plcl_o = 0, plfc_o = 0, tkeavg_o = 0, thvlmin_o = 0, qtsrc_o = 0, thvlsrc_o = 0, thlsrc_o = 0;
    D usrc_o = 0, vsrc_o = 0, thv0lcl_o = 0, cin_i = 0, cinlcl_i = 0, ke = 0;
    D cush_s = 0;

    // locals that persist across stages
    D plcl = 0, plfc = 0, tkeavg = 0, thvlmin = 0, qtsrc = 0, thvlsrc = 0, thlsrc = 0, usrc = 0, vsrc = 0;
    D thv0lcl = 0, cin = 0, cinlcl = 0;
    R prel = 0, thv0rel = 0, cbmf = 0, wlcl = 0, ufrclcl = 0, ppen = 0, thlu_top = 0, qtu_top = 0;
    R rho0j = 0, drage = 0, bogbot = 0, bogtop = 0, wtwb = 0;
    R trsrc_tmp = 0;
    (void)trsrc_tmp;

    // The scalars live across the split of a deferred column: every function-scope
    // local that the second convection pass reads or carries to the end of the column.
    // FP32 values (the mixed path) are held exactly as doubles.
    auto resume_io = [&](auto&& io) {
        io(0, cush); io(1, cin); io(2, cinlcl); io(3, plcl); io(4, plfc); io(5, tkeavg); io(6, thvlmin);
        io(7, qtsrc); io(8, thvlsrc); io(9, thlsrc); io(10, usrc); io(11, vsrc); io(12, thv0lcl); io(13, prel);
        io(14, thv0rel); io(15, cbmf); io(16, wlcl); io(17, ufrclcl); io(18, ppen); io(19, thlu_top);
        io(20, qtu_top); io(21, rho0j); io(22, drage); io(23, bogbot); io(24, bogtop); io(25, wtwb); io(26, th);
        io(27, qvj); io(28, qlj); io(29, qij); io(30, kinv); io(31, klcl); io(32, klfc); io(33, krel);
        io(34, kbup); io(35, kpen); io(36, limit_ufrc); io(37, limit_ppen); io(38, limit_emf); io(39, amb);
        // the first pass's values that its implicit-CIN restore reads (a column continued
        // at its second pass has used them already)
        if constexpr (kSplit) {
            io(40, kinv_o); io(41, klcl_o); io(42, klfc_o); io(43, plcl_o); io(44, plfc_o); io(45, tkeavg_o);
            io(46, thvlmin_o); io(47, qtsrc_o); io(48, thvlsrc_o); io(49, thlsrc_o); io(50, usrc_o);
            io(51, vsrc_o); io(52, thv0lcl_o); io(53, cin_i); io(54, cinlcl_i); io(55, ke);
        }
    };
    auto defer_here = [&](bool mark) {   // save the scalars and list (or mark) the column
        resume_io([&](int f, auto& x) {
            if constexpr (std::is_integral<std::remove_reference_t<decltype(x)>>::value) rstate(f) = (double)x;
            else rstate(f) = (double)val(x);
        });
        if (mark) a.defer_mark[iw] = 1;
        else a.defer_list[__atomic_fetch_add(a.defer_count, 1, __ATOMIC_RELAXED)] = i;
    };
    if (resume) {
        resume_io([&](int f, auto& x) { x = decltype(+x)(rstate(f)); });
    }
    if constexpr (PART == PART_CIN2) {
        // the first pass ended with this cush; a failed saturation adjustment of the
        // set-B environment ends the column as the first pass's S22 would
        cush_s = cush;
        for (int k = 1; k <= mkx; ++k)
            if (E(B_THV0BOT, k) < 0.0) { exit_iter = 1; exit_code = EX_CONDEN; goto L333; }
    }

    if (a.env_ready) {
        for (int k = 1; k <= mkx; ++k)
            if (E(A_THV0BOT, k) < 0.0) { exit_code = EX_CONDEN; goto L333; }
    } else if (!interface_env(A_THL0)) {
        exit_code = EX_CONDEN;
        goto L333;
    }

    for (int iter = resume ? (kSplit ? (resume == 3 ? 1 : resume == 4 ? 2 : resume) : 2) : 1; iter <= 2; ++iter) {
        // this pass continues after its buoyancy sorting
        const bool after_sort = resume == 3 && iter == 1;
        exit_iter = iter;
        const int eb = (iter == 1) ? A_THL0 : B_THL0;        // environment set for S1-S6
        const int e_thl = eb, e_qt = eb + 1, e_ssthl = eb + 2, e_ssqt = eb + 3;
        const int e_thvb = (iter == 1) ? A_THV0BOT : B_THV0BOT, e_thvt = e_thvb + 1;
        const int e_thvlb = (iter == 1) ? A_THVL0BOT : B_THVL0BOT, e_thvlt = e_thvlb + 1;
        const D tscaleh = cush;
        if (!after_sort) cush = -1.0;   // continued after S11: its cush stands

        if (kSplit ? (resume == iter || after_sort) : resume != 0) {
            // a deferred column continues with the scalars it saved: at the second pass
            // those of its implicit-CIN restore, whose convection starts afresh
            if (!kSplit || iter == 2) zero_pass_state();
        } else {
        // ---- S1 kinv (uwshcu.F90:1455-1468) ----
        kinv = 1;
        for (int k = mkx - 1; k >= 1; --k) {
            if ((pblh + 5.0 - in(a.zs0, k)) * (pblh + 5.0 - in(a.zs0, k + 1)) < 0.0) { kinv = k + 1; break; }
        }
        if (kinv <= 1) { exit_code = EX_KINV1; goto L333; }

        // ---- S2 PBL averages (uwshcu.F90:1484-1499) ----
        {
            D dpsum = 0.0;
            tkeavg = 0.0;
            thvlmin = 1000.0;
            for (int k = 0; k <= kinv - 1; ++k) {
                D dpi;
                if (k == 0) dpi = in(a.ps0, 0) - in(a.p0, 1);
                else if (k == kinv - 1) dpi = in(a.p0, kinv - 1) - in(a.ps0, kinv - 1);
                else dpi = in(a.p0, k) - in(a.p0, k + 1);
                dpsum = dpsum + dpi;
                tkeavg = tkeavg + dpi * in(a.tke, k);
                if (k != 0) thvlmin = fmin(thvlmin, fmin(E(e_thvlb, k), E(e_thvlt, k)));
            }
            tkeavg = tkeavg / dpsum;
        }

        // ---- S3 source air (uwshcu.F90:1509-1516) ----
        qtsrc = E(e_qt, 1);
        thvlsrc = thvlmin;
        thlsrc = thvlsrc / (1.0 + C.zvir * qtsrc);
        usrc = in(a.u0, kinv - 1) + E(A_SSU0, kinv - 1) * (in(a.ps0, kinv - 1) - in(a.p0, kinv - 1));
        vsrc = in(a.v0, kinv - 1) + E(A_SSV0, kinv - 1) * (in(a.ps0, kinv - 1) - in(a.p0, kinv - 1));

        // ---- S4 LCL (uwshcu.F90:1533-1549) ----
        plcl = T.qsinvert(qtsrc, thlsrc, in(a.ps0, 0));
        klcl = mkx;
        for (int k = 0; k <= mkx; ++k) {
            if (in(a.ps0, k) < plcl) { klcl = k; break; }
        }
        klcl = klcl > 1 ? klcl : 1;
        if (plcl < 30000.0) { exit_code = EX_KLCLMKX; goto L333; }

        // ---- S5 thv0lcl (uwshcu.F90:1559-1567) ----
        {
            const D thl0lcl = E(e_thl, klcl) + E(e_ssthl, klcl) * (plcl - in(a.p0, klcl));
            const D qt0lcl = E(e_qt, klcl) + E(e_ssqt, klcl) * (plcl - in(a.p0, klcl));
            if (T.conden(plcl, thl0lcl, qt0lcl, th, qvj, qlj, qij)) { exit_code = EX_CONDEN; goto L333; }
            thv0lcl = th * (1.0 + C.zvir * qvj - qlj - qij);
        }

        // ---- S6 CIN (uwshcu.F90:1592-1685) ----
        cin = 0.0;
        cinlcl = 0.0;
        plfc = 0.0;
        klfc = mkx;
        if (klcl >= kinv) {
            D thvubot = 0.0, thvutop = 0.0;
            for (int k = lvl_lo(kinv); k <= mkx - 1; ++k) {
                if (k < kinv) continue;
                if (k < klcl) {
                    thvubot = thvlsrc;
                    thvutop = thvlsrc;
                    cin = cin + T.single_cin(in(a.ps0, k - 1), E(e_thvb, k), in(a.ps0, k), E(e_thvt, k), thvubot, thvutop);
                } else if (k == klcl) {
                    thvubot = thvlsrc;
                    thvutop = thvlsrc;
                    cin = cin + T.single_cin(in(a.ps0, k - 1), E(e_thvb, k), plcl, thv0lcl, thvubot, thvutop);
                    cinlcl = fmax(cin, 0.0);
                    cin = cinlcl;
                    thvubot = thvlsrc;
                    if (T.conden_x(in(a.ps0, k), E(A_EXNS0, k), thlsrc, qtsrc, th, qvj, qlj, qij)) { exit_code = EX_CONDEN; goto L333; }
                    thvutop = th * (1.0 + C.zvir * qvj - qlj - qij);
                    T.getbuoy(plcl, thv0lcl, in(a.ps0, k), E(e_thvt, k), thvubot, thvutop, plfc, cin);
                    if (plfc > 0.0) { klfc = k; break; }
                } else {
                    thvubot = thvutop;
                    if (T.conden_x(in(a.ps0, k), E(A_EXNS0, k), thlsrc, qtsrc, th, qvj, qlj, qij)) { exit_code = EX_CONDEN; goto L333; }
                    thvutop = th * (1.0 + C.zvir * qvj - qlj - qij);
                    T.getbuoy(in(a.ps0, k - 1), E(e_thvb, k), in(a.ps0, k), E(e_thvt, k), thvubot, thvutop, plfc, cin);
                    if (plfc > 0.0) { klfc = k; break; }
                }
            }
        } else {
            cinlcl = 0.0;
            for (int k = lvl_lo(kinv); k <= mkx - 1; ++k) {
                if (k < kinv) continue;
                if (T.conden_x(in(a.ps0, k - 1), E(A_EXNS0, k - 1), thlsrc, qtsrc, th, qvj, qlj, qij)) { exit_code = EX_CONDEN; goto L333; }
                const D thvubot = th * (1.0 + C.zvir * qvj - qlj - qij);
                if (T.conden_x(in(a.ps0, k), E(A_EXNS0, k), thlsrc, qtsrc, th, qvj, qlj, qij)) { exit_code = EX_CONDEN; goto L333; }
                const D thvutop = th * (1.0 + C.zvir * qvj - qlj - qij);
                T.getbuoy(in(a.ps0, k - 1), E(e_thvb, k), in(a.ps0, k), E(e_thvt, k), thvubot, thvutop, plfc, cin);
                if (plfc > 0.0) { klfc = k; break; }
            }
        }
        cin = fmax(0.0, cin);
        if (klfc >= mkx) { klfc = mkx; exit_code = EX_KLFCMKX; goto L333; }

        // ---- S7 implicit CIN (uwshcu.F90:1693-2014) ----
        if (iter == 1) {
            cin_i = cin;
            cinlcl_i = cinlcl;
            ke = 1.0 / (1.0 * tkeavg + 5.e-4);      // rbuoy / (rkfre*tkeavg + epsvarw)
            kinv_o = kinv; klcl_o = klcl; klfc_o = klfc; plcl_o = plcl; plfc_o = plfc; tkeavg_o = tkeavg;
            thvlmin_o = thvlmin; qtsrc_o = qtsrc; thvlsrc_o = thvlsrc; thlsrc_o = thlsrc; usrc_o = usrc;
            vsrc_o = vsrc; thv0lcl_o = thv0lcl;
        } else {
            const D del_CIN = cin - cin_i;
            if (cin_i + cin > 0.0) near(del_CIN, cin_i + cin);
            if (del_CIN > 0.0) {
                const D alpha = compute_alpha(del_CIN, ke);
                cin = cin_i + alpha * del_CIN;
                cinlcl = cinlcl_i;
                kinv = kinv_o; klcl = klcl_o; klfc = klfc_o; plcl = plcl_o; plfc = plfc_o; tkeavg = tkeavg_o;
                thvlmin = thvlmin_o; qtsrc = qtsrc_o; thvlsrc = thvlsrc_o; thlsrc = thlsrc_o; usrc = usrc_o;
                vsrc = vsrc_o; thv0lcl = thv0lcl_o;
                if (PART == PART_CIN2 ||
                    (PART == PART_ALL && kSplit && (a.defer_list || a.defer_mark) && !a.defer_at_closure)) {
                    // the second pass runs in a launch of its own: save, list, stop
                    cush = tscaleh;
                    defer_here(a.defer_mark != nullptr);
                    return;
                }
                zero_pass_state();
            } else {
                // iteration-1 outputs stand (uwshcu.F90:1897-2010): they are already in the
                // output arrays; flxprc1/flxsnow1 were never written and stay zero.
                cout(a.cush) = cush_s;
                path = 1;
                goto Ldone;
            }
        }
        }   // S1-S7

        // From here on the environment is set A (the input state).
        {
        R wtw;   // set by the closure, or at the start of the buoyancy sorting
        // a column continued at its first pass (resume 1) resumes at the buoyancy sorting
        if (kSplit ? !((resume == 1 && iter == 1) || after_sort) : true) {
        if constexpr (PART == PART_CLOSURE) {
            zero_pass_state();
            for (int k = 0; k <= mkx; ++k) { out(a.flxprc1, k) = 0.0; out(a.flxsnow1, k) = 0.0; }
        }
        // ---- S8 release level (uwshcu.F90:2035-2043) ----
        if (klcl < kinv) {
            krel = kinv;
            prel = R(in(a.ps0, krel - 1));
            thv0rel = Er(A_THV0BOT, krel);
        } else {
            krel = klcl;
            prel = R(plcl);
            thv0rel = R(thv0lcl);
        }

        // ---- S9 cloud-base closure (uwshcu.F90:2092-2184; zwshcu.F90:2275-2384), FP64 ----
        {
            D wcrit = sqrt(2.0 * cin * 1.0);
            D sigmaw = sqrt(1.0 * tkeavg + 5.e-4);
            D rho0inv = 0.0;
            if constexpr (ZW) {
                rho0inv = in(a.ps0, kinv - 1) / (C.r * E(A_THV0TOP, kinv - 1) * E(A_EXNS0, kinv - 1));
                if (a.zw.pertw) {
                    // zwshcu.F90:2282-2285: a warm perturbation widens the w' distribution
                    const double tp = fmin(a.tpert[i], a.zw.tpertmax);
                    tkeavg = tkeavg + zmh_ramp(tp, 0.0, 2.0) * 2.0;
                    sigmaw = sqrt(1.0 * tkeavg + 5.e-4) + zmh_ramp(tp, 0.0, 1.0) * 0.1;
                }
                if (a.zw.lf) {
                    // zwshcu.F90:2296-2300: ascent at the PBL-top layer lowers wcrit
                    UW_CHECK(kinv - 1 >= 1 && kinv - 1 <= mkx);
                    const D tmp1 = fmin(fmax(-a.omega[row(kinv - 1)], 0.0), 3.0);
                    wcrit = wcrit - tmp1 / rho0inv / 9.8;
                    wcrit = fmax(wcrit, 0.0);
                }
                wcrit = wcrit * a.zw.fwcrit;   // zwshcu.F90:2303
            }
            D mu = wcrit / sigmaw / 1.4142;
            if (mu >= 3.0) { exit_code = EX_MU3; goto L333; }
            if constexpr (!ZW)
                rho0inv = in(a.ps0, kinv - 1) / (C.r * E(A_THV0TOP, kinv - 1) * E(A_EXNS0, kinv - 1));
            D cbmf_d = (rho0inv * sigmaw / 2.5066) * uexp(-(mu * mu));
            const D cbmflimit = 0.9 * in(a.dp0, kinv - 1) / C.g / dtd;
            D mumin0 = 0.0;
            if (cbmf_d > cbmflimit) mumin0 = sqrt(-ulog(2.5066 * cbmflimit / rho0inv / sigmaw));
            mu = fmax(fmax(mu, mumin0), 0.906);
            const D mulcl = sqrt(2.0 * cinlcl * 1.0) / 1.4142 / sigmaw;
            D mulclstar;
            {
                const D e1 = uexp(-(mu * mu)) / 2.5066, ef = uerfc(mu);
                mulclstar = sqrt(fmax(0.0, 2.0 * (e1 * e1) * (1.0 / (ef * ef) - 0.25 / (0.10 * 0.10))));
            }
            if (mulcl > 1.e-8 && mulcl > mulclstar) {
                const D mumin2 = compute_mumin2<D>(mulcl, D(0.10), mu);
                if (mu > mumin2) record_error(a.err, UWSHCU_ERR_MUMIN2, i);
                mu = fmax(mu, mumin2);
                if (mu == mumin2) limit_ufrc = 1;
            }
            if (mu == 0.906) limit_ufrc = 1;
            cbmf_d = (rho0inv * sigmaw / 2.5066) * uexp(-(mu * mu));
            const D winv = sigmaw * (2.0 / 2.5066) * uexp(-(mu * mu)) / uerfc(mu);
            const D wtw_d = winv * winv - 2.0 * cinlcl * 1.0;
            if (wtw_d <= 0.0) { exit_code = EX_WTW; goto L333; }
            const D wlcl_d = sqrt(wtw_d);
            const D ufrclcl_d = cbmf_d / wlcl_d / rho0inv;
            if (ufrclcl_d <= 0.0001) { exit_code = EX_UFRC; goto L333; }
            cbmf = R(cbmf_d);
            wlcl = R(wlcl_d);
            ufrclcl = R(ufrclcl_d);
            wtw = R(wtw_d);
            for (int k = kinv - 1; k <= krel - 1; ++k) { out(a.umf, k) = cbmf_d; out(a.wu, k) = winv; }
        }
        const R wrel = wlcl;
        W(W_UFRC, krel - 1) = ufrclcl;

        // ---- S10 release-level updraft (uwshcu.F90:2205-2255) ----
        W(W_EMF, krel - 1) = R(0);
        out(a.umf, krel - 1) = D(cbmf);
        out(a.wu, krel - 1) = D(wrel);
        W(W_THLU, krel - 1) = R(thlsrc);
        W(W_QTU, krel - 1) = R(qtsrc);
        if (T.conden(prel, thlsrc, qtsrc, th, qvj, qlj, qij)) { exit_code = EX_CONDEN; goto L333; }
        W(W_THVU, krel - 1) = thv_of(th, qvj, qlj, qij);
        {
            R uplus = R(0), vplus = R(0);
            if (krel == kinv) {
                uplus = PGFc * Er(A_SSU0, kinv) * (prel - R(in(a.ps0, kinv - 1)));
                vplus = PGFc * Er(A_SSV0, kinv) * (prel - R(in(a.ps0, kinv - 1)));
            } else {
                const int kend = (krel - 1 > kinv) ? krel - 1 : kinv;
                for (int k = kinv; k <= kend; ++k) {
                    uplus = uplus + PGFc * Er(A_SSU0, k) * R(in(a.ps0, k) - in(a.ps0, k - 1));
                    vplus = vplus + PGFc * Er(A_SSV0, k) * R(in(a.ps0, k) - in(a.ps0, k - 1));
                }
                uplus = uplus + PGFc * Er(A_SSU0, krel) * (prel - R(in(a.ps0, krel - 1)));
                vplus = vplus + PGFc * Er(A_SSV0, krel) * (prel - R(in(a.ps0, krel - 1)));
            }
            W(W_UU, krel - 1) = R(usrc) + uplus;
            W(W_VU, krel - 1) = R(vsrc) + vplus;
        }
        for (int m = 1; m <= ncnst; ++m) WT(1, m, krel - 1) = R(trin(m, 1));   // trsrc = tr0(1,m)

        if (PART == PART_CLOSURE ||
            (PART == PART_ALL && kSplit && iter == 1 && a.defer_at_closure && (a.defer_mark || a.defer_list))) {
            // it convects: the plume stages run in a launch of their own, packed with the
            // other columns that got this far
            cush = tscaleh;
            defer_here(a.defer_mark != nullptr);
            return;
        }
        }   // S8-S10

        // ---- S11 buoyancy sorting (uwshcu.F90:2296-2952) ----
        if (!(kSplit && after_sort)) {   // a column continued after S11 skips it
        R scaleh = R(tscaleh);
        if (R(tscaleh) < R(0)) scaleh = R(1000);
        for (int iter_scaleh = 1; iter_scaleh <= 3; ++iter_scaleh) {
            kbup = krel;
            kpen = krel;
            wtw = wlcl * wlcl;
            R pe = R(0.5) * (prel + R(in(a.ps0, krel)));
            R dpe = prel - R(in(a.ps0, krel));
            D exe = T.exnf(D(pe));          // Exner function of pe, FP64
            R exne = R(exe);
            R thvebot = thv0rel;
            R thle = Er(A_THL0, krel) + Er(A_SSTHL0, krel) * (pe - R(in(a.p0, krel)));
            R qte = Er(A_QT0, krel) + Er(A_SSQT0, krel) * (pe - R(in(a.p0, krel)));
            R ue = R(in(a.u0, krel)) + Er(A_SSU0, krel) * (pe - R(in(a.p0, krel)));
            R ve = R(in(a.v0, krel)) + Er(A_SSV0, krel) * (pe - R(in(a.p0, krel)));
            // tre(m) = tr0(krel,m) + sstr0(krel,m)*(pe - p0(krel)) is recomputed where used
            R pe_tr = pe - R(in(a.p0, krel));
            int ktr = krel;            // layer whose (tr0, sstr0) define tre, and offset pe_tr
            bool tr_mid = false;       // after the first layer tre = tr0(k) exactly
            // The plume at interface k-1: read from the workspace for the first level, then
            // carried over from the level below, which stored the same values.
            R c_thlu = W(W_THLU, krel - 1), c_qtu = W(W_QTU, krel - 1), c_uu = W(W_UU, krel - 1);
            R c_vu = W(W_VU, krel - 1), c_thvu = W(W_THVU, krel - 1);
            R c_wu = R(out(a.wu, krel - 1)), c_umf = R(out(a.umf, krel - 1));

            for (int k = lvl_lo(krel); k <= mkx - 1; ++k) {
                if (k < krel) continue;    // level-aligned lanes (wave_min_level)
                const int km1 = k - 1;
                // Read once for both iter_xc passes: the environment of layer k, which the loop
                // body does not write, and (issued early) that of layer k+1 for the next level.
                const R thlu_km1 = c_thlu, qtu_km1 = c_qtu;
                const R uu_km1 = c_uu, vu_km1 = c_vu, thvu_km1 = c_thvu;
                const R wu_km1 = c_wu;
                const R z0k = R(in(a.z0, k)), dp0k = R(in(a.dp0, k)), thv0top_k = Er(A_THV0TOP, k);
                const R ss_thl = Er(A_SSTHL0, k), ss_qt = Er(A_SSQT0, k), ss_u = Er(A_SSU0, k), ss_v = Er(A_SSV0, k);
                const D ps0k = in(a.ps0, k), exns0k = E(A_EXNS0, k);
                const D p0_kp1 = in(a.p0, k + 1), dp0_kp1 = in(a.dp0, k + 1), exn0_kp1 = E(A_EXN0, k + 1);
                const R thv0bot_kp1 = Er(A_THV0BOT, k + 1), thl0_kp1 = Er(A_THL0, k + 1), qt0_kp1 = Er(A_QT0, k + 1);
                const D u0_kp1 = in(a.u0, k + 1), v0_kp1 = in(a.v0, k + 1);
                R thlue = thlu_km1;
                R qtue = qtu_km1;
                R wue = wu_km1;
                wtwb = wtw;
                const R umf_km1 = c_umf;
                R umfk = 0, ferk = 0, fdrk = 0;
                // Level-k results of an iter_xc pass.  Each pass overwrites the last and none
                // reads them back, so they are stored once, after the passes.
                R reik = 0, thluk = 0, qtuk = 0, uuk = 0, vuk = 0, thvuk = 0, dwtenk = 0, ditenk = 0;
                // the environment at pe does not change between the iter_xc passes
                R thv0j;
                D es, qs;
                if (k == krel) {
                    if (T.conden_x(pe, exe, thle, qte, th, qvj, qlj, qij)) { exit_code = EX_CONDEN; goto L333; }
                    thv0j = thv_of(th, qvj, qlj, qij);
                    T.qsat_es_qs(D(thle * exne), pe, es, qs);
                } else {
                    D tv;
                    if (!env_mid(k, tv, qs)) { exit_code = EX_CONDEN; goto L333; }
                    thv0j = R(tv);
                }
                rho0j = pe / (r * thv0j * exne);
                const R excess0 = qte - R(qs);
                for (int iter_xc = 1; iter_xc <= 2; ++iter_xc) {
                    wtw = wu_km1 * wu_km1;
                    if (T.conden_x(pe, exe, thlue, qtue, th, qvj, qlj, qij)) { exit_code = EX_CONDEN; goto L333; }
                    near(qlj + qij - criqc, criqc);
                    if ((R(qlj) + R(qij)) > criqc) {
                        const R qlr = R(qlj), qir = R(qij);
                        const R exql = ((qlr + qir) - criqc) * qlr / (qlr + qir);
                        const R exqi = ((qlr + qir) - criqc) * qir / (qlr + qir);
                        qtue = qtue - exql - exqi;
                        thlue = thlue + (xlv / cp / exne) * exql + (xls / cp / exne) * exqi;
                        // unchanged arguments give unchanged results, so only re-adjust here
                        if (T.conden_x(pe, exe, thlue, qtue, th, qvj, qlj, qij)) { exit_code = EX_CONDEN; goto L333; }
                    }
                    const R thvj = thv_of(th, qvj, qlj, qij);
                    T.qsat_es_qs(D(thlue * exne), pe, es, qs);
                    const R excessu = qtue - R(qs);
                    near(excessu, qtue);
                    near(excess0, qte);
                    const R cridis = rle * scaleh;
                    R xc;
                    if ((excessu <= R(0) && excess0 <= R(0)) || (excessu >= R(0) && excess0 >= R(0))) {
                        xc = fmin(R(1), fmax(R(0), R(1) - R(2) * rbuoy * g * cridis / (wue * wue) * (R(1) - thvj / thv0j)));
                    } else {
                        const R xsat = excessu / (excessu - excess0);
                        const R thlxsat = thlue + xsat * (thle - thlue);
                        const R qtxsat = qtue + xsat * (qte - qtue);
                        if (T.conden_x(pe, exe, thlxsat, qtxsat, th, qvj, qlj, qij)) { exit_code = EX_CONDEN; goto L333; }
                        const R thvxsat = thv_of(th, qvj, qlj, qij);
                        R x_cu = 0, x_en = 0;
                        for (int kk = 1; kk <= 2; ++kk) {
                            R thv_x0, thv_x1;
                            if (kk == 1) {
                                thv_x0 = thvj;
                                thv_x1 = (R(1) - R(1) / xsat) * thvj + (R(1) / xsat) * thvxsat;
                            } else {
                                thv_x1 = thv0j;
                                thv_x0 = (xsat / (xsat - R(1))) * thv0j + (R(1) / (R(1) - xsat)) * thvxsat;
                            }
                            const R aquad = wue * wue;
                            const R bquad = R(2) * rbuoy * g * cridis * (thv_x1 - thv_x0) / thv0j - R(2) * (wue * wue);
                            const R cquad = R(2) * rbuoy * g * cridis * (thv_x0 - thv0j) / thv0j + (wue * wue);
                            R xs1 = 0, xs2 = 0;
                            int status;
                            near(bquad * bquad - R(4) * aquad * cquad, bquad * bquad + fabs(R(4) * aquad * cquad));
                            if (kk == 1) {
                                if ((bquad * bquad - R(4) * aquad * cquad) >= R(0)) {
                                    roots(aquad, bquad, cquad, xs1, xs2, status);
                                    near(xsat - fmin(xs1, xs2), R(1));
                                    x_cu = fmin(R(1), fmax(R(0), fmin(xsat, fmin(xs1, xs2))));
                                } else {
                                    x_cu = xsat;
                                }
                            } else {
                                if ((bquad * bquad - R(4) * aquad * cquad) >= R(0)) {
                                    roots(aquad, bquad, cquad, xs1, xs2, status);
                                    x_en = fmin(R(1), fmax(R(0), fmax(xsat, fmin(xs1, xs2))));
                                } else {
                                    x_en = R(1);
                                }
                            }
                        }
                        xc = (x_cu == xsat) ? fmax(x_cu, x_en) : x_cu;
                    }
                    near(xc - R(0.5), R(1));
                    const R ee2 = xc * xc;
                    const R ud2 = R(1) - R(2) * xc + xc * xc;
                    reik = (R(0.5) * rkm / z0k / g / rho0j);
                    if (xc > R(0.5))
                        reik = fmin(reik, R(0.9) * ulog(dp0k / g / dt / umf_km1 + R(1)) / dpe / (R(2) * xc - R(1)));
                    ferk = reik * ee2;
                    fdrk = reik * ud2;
                    if constexpr (ZW) {
                        if (a.zw.ent) {
                            // zwshcu.F90:2818-2844: entrainment from the updraft buoyancy
                            const R bog = rbuoy * (thvu_km1 / thvebot - R(1)) * g;
                            ferk = R(a.zw.ent0) * bog / fmax(wue * wue, R(1.0e-4));
                            const R fer_min = bog > R(1.e-4) ? R(1.e-4) : R(0);
                            ferk = fmin(fmax(ferk, fer_min), R(3.0e-2));
                        }
                    }
                    umfk = umf_km1 * uexp(dpe * (ferk - fdrk));
                    if (ferk * dpe < R(1.e-4)) {
                        thluk = thlu_km1 + (thle + ss_thl * dpe / R(2) - thlu_km1) * ferk * dpe;
                        qtuk = qtu_km1 + (qte + ss_qt * dpe / R(2) - qtu_km1) * ferk * dpe;
                        uuk = uu_km1 + (ue + ss_u * dpe / R(2) - uu_km1) * ferk * dpe - PGFc * ss_u * dpe;
                        vuk = vu_km1 + (ve + ss_v * dpe / R(2) - vu_km1) * ferk * dpe - PGFc * ss_v * dpe;
                    } else {
                        const R ef = uexp(-ferk * dpe);
                        thluk = (thle + ss_thl / ferk - ss_thl * dpe / R(2)) -
                                (thle + ss_thl * dpe / R(2) - thlu_km1 + ss_thl / ferk) * ef;
                        qtuk = (qte + ss_qt / ferk - ss_qt * dpe / R(2)) -
                               (qte + ss_qt * dpe / R(2) - qtu_km1 + ss_qt / ferk) * ef;
                        uuk = (ue + (R(1) - PGFc) * ss_u / ferk - ss_u * dpe / R(2)) -
                              (ue + ss_u * dpe / R(2) - uu_km1 + (R(1) - PGFc) * ss_u / ferk) * ef;
                        vuk = (ve + (R(1) - PGFc) * ss_v / ferk - ss_v * dpe / R(2)) -
                              (ve + ss_v * dpe / R(2) - vu_km1 + (R(1) - PGFc) * ss_v / ferk) * ef;
                    }
                    if (T.conden_x(ps0k, exns0k, thluk, qtuk, th, qvj, qlj, qij)) { exit_code = EX_CONDEN; goto L333; }
                    near(qlj + qij - criqc, criqc);
                    if ((R(qlj) + R(qij)) > criqc) {
                        const R qlr = R(qlj), qir = R(qij);
                        const R exql = ((qlr + qir) - criqc) * qlr / (qlr + qir);
                        const R exqi = ((qlr + qir) - criqc) * qir / (qlr + qir);
                        qtuk = qtuk - exql - exqi;
                        thluk = thluk + (xlv / cp / R(exns0k)) * exql + (xls / cp / R(exns0k)) * exqi;
                        dwtenk = exql;
                        ditenk = exqi;
                        if (T.conden_x(ps0k, exns0k, thluk, qtuk, th, qvj, qlj, qij)) { exit_code = EX_CONDEN; goto L333; }
                    } else {
                        dwtenk = R(0);
                        ditenk = R(0);
                    }
                    thvuk = thv_of(th, qvj, qlj, qij);
                    bogbot = rbuoy * (thvu_km1 / thvebot - R(1));
                    bogtop = rbuoy * (thvuk / thv0top_k - R(1));
                    const R delbog = bogtop - bogbot;
                    drage = ferk * (R(1) + rdrag);
                    const R expfac = uexp(R(-2) * drage * dpe);
                    wtwb = wtw;
                    if (drage * dpe > R(1.e-3))
                        wtw = wtw * expfac + (delbog + (R(1) - expfac) * (bogbot + delbog / (R(-2) * drage * dpe))) / (rho0j * drage);
                    else
                        wtw = wtw + dpe * (bogbot + bogtop) / rho0j;
                    near(wtw, wtwb);
                    near(bogtop, R(1));
                    if (wtw > R(0)) {
                        thlue = R(0.5) * (thlu_km1 + thluk);
                        qtue = R(0.5) * (qtu_km1 + qtuk);
                        wue = R(0.5) * sqrt(fmax(wtwb + wtw, R(0)));
                    } else {
                        break;   // go to 111
                    }
                }
                // 111 continue
                W(W_REI, k) = reik;
                W(W_FER, k) = ferk;
                W(W_FDR, k) = fdrk;
                W(W_EMF, k) = R(0);
                W(W_DWTEN, k) = dwtenk;
                W(W_DITEN, k) = ditenk;
                W(W_THLU, k) = thluk;
                W(W_QTU, k) = qtuk;
                W(W_UU, k) = uuk;
                W(W_VU, k) = vuk;
                out(a.umf, k) = D(umfk);
                W(W_THVU, k) = thvuk;
                // Updraft tracers (uwshcu.F90:2592-2614).  Nothing in an iter_xc pass reads
                // them and each pass overwrites the last, so they are formed once, from the
                // fer(k) of the pass that ended the loop.
                if (ferk * dpe < R(1.e-4)) {
                    for (int m = 1; m <= ncnst; ++m) {
                        const R tre = tr_mid ? R(trin(m, ktr)) : R(trin(m, ktr)) + WT(0, m, ktr) * pe_tr;
                        WT(1, m, k) = WT(1, m, km1) + (tre + WT(0, m, k) * dpe / R(2) - WT(1, m, km1)) * ferk * dpe;
                    }
                } else {
                    const R ef = uexp(-ferk * dpe);
                    for (int m = 1; m <= ncnst; ++m) {
                        const R tre = tr_mid ? R(trin(m, ktr)) : R(trin(m, ktr)) + WT(0, m, ktr) * pe_tr;
                        const R sst = WT(0, m, k);
                        WT(1, m, k) = (tre + sst / ferk - sst * dpe / R(2)) -
                                      (tre + sst * dpe / R(2) - WT(1, m, km1) + sst / ferk) * ef;
                    }
                }
                if (umfk == R(0)) wtw = R(-1);
                if (bogtop > R(0) && wtw > R(0)) kbup = k;
                if (wtw <= R(0)) { kpen = k; break; }   // go to 45
                {
                    const R wuk = sqrt(wtw);
                    out(a.wu, k) = D(wuk);
                    if (wuk > R(100)) { exit_code = EX_WU; goto L333; }
                    const R rhos0j = R(ps0k) / (r * R(0.5) * (thv0bot_kp1 + thv0top_k) * R(exns0k));
                    R ufrck = umfk / (rhos0j * wuk);
                    near(ufrck - rmaxfrac, rmaxfrac);
                    if (ufrck > rmaxfrac) {
                        limit_ufrc = 1;
                        ufrck = rmaxfrac;
                        umfk = rmaxfrac * rhos0j * wuk;
                        out(a.umf, k) = D(umfk);
                        W(W_FDR, k) = ferk - ulog(umfk / umf_km1) / dpe;
                    }
                    W(W_UFRC, k) = ufrck;
                    c_wu = R(D(wuk));
                    c_umf = R(D(umfk));
                }
                c_thlu = thluk;
                c_qtu = qtuk;
                c_uu = uuk;
                c_vu = vuk;
                c_thvu = thvuk;
                pe = R(p0_kp1);
                dpe = R(dp0_kp1);
                exne = R(exn0_kp1);
                if constexpr (sizeof(R) == sizeof(double)) exe = exn0_kp1;   // = exnf(pe), pe = p0(k+1)
                else exe = T.exnf(D(pe));
                thvebot = thv0bot_kp1;
                thle = thl0_kp1;
                qte = qt0_kp1;
                ue = R(u0_kp1);
                ve = R(v0_kp1);
                ktr = k + 1;
                tr_mid = true;
            }
            // 45 continue
            if (drage == R(0)) {
                const R aquad = (bogtop - bogbot) / R(in(a.ps0, kpen) - in(a.ps0, kpen - 1));
                const R bquad = R(2) * bogbot;
                const R wuk = R(out(a.wu, kpen - 1));
                const R cquad = -(wuk * wuk) * rho0j;
                R xc1 = 0, xc2 = 0;
                int status;
                roots(aquad, bquad, cquad, xc1, xc2, status);
                const R dpk = R(in(a.dp0, kpen));
                if (status == 0) {
                    if (xc1 <= R(0) && xc2 <= R(0)) {
                        ppen = fmax(xc1, xc2);
                        ppen = fmin(R(0), fmax(-dpk, ppen));
                    } else if (xc1 > R(0) && xc2 > R(0)) {
                        ppen = -dpk;
                    } else {
                        ppen = fmin(xc1, xc2);
                        ppen = fmin(R(0), fmax(-dpk, ppen));
                    }
                } else {
                    ppen = -dpk;
                }
            } else {
                ppen = compute_ppen(wtwb, drage, bogbot, bogtop, rho0j, R(in(a.dp0, kpen)));
            }
            if (ppen == -R(in(a.dp0, kpen)) || ppen == R(0)) limit_ppen = 1;
            {
                const R ferp = W(W_FER, kpen), thl0p = Er(A_THL0, kpen), ssthlp = Er(A_SSTHL0, kpen);
                const R qt0p = Er(A_QT0, kpen), ssqtp = Er(A_SSQT0, kpen);
                if (ferp * (-ppen) < R(1.e-4)) {
                    thlu_top = W(W_THLU, kpen - 1) + (thl0p + ssthlp * (-ppen) / R(2) - W(W_THLU, kpen - 1)) * ferp * (-ppen);
                    qtu_top = W(W_QTU, kpen - 1) + (qt0p + ssqtp * (-ppen) / R(2) - W(W_QTU, kpen - 1)) * ferp * (-ppen);
                } else {
                    const R ef = uexp(-ferp * (-ppen));
                    thlu_top = (thl0p + ssthlp / ferp - ssthlp * (-ppen) / R(2)) -
                               (thl0p + ssthlp * (-ppen) / R(2) - W(W_THLU, kpen - 1) + ssthlp / ferp) * ef;
                    qtu_top = (qt0p + ssqtp / ferp - ssqtp * (-ppen) / R(2)) -
                              (qt0p + ssqtp * (-ppen) / R(2) - W(W_QTU, kpen - 1) + ssqtp / ferp) * ef;
                }
            }
            {
                const R ptop = R(in(a.ps0, kpen - 1)) + ppen;
                if (T.conden(ptop, thlu_top, qtu_top, th, qvj, qlj, qij)) { exit_code = EX_CONDEN; goto L333; }
                near(qlj + qij - criqc, criqc);
                const R exntop = R(upow(D(ptop) / C.p00, C.rovcp));
                if ((R(qlj) + R(qij)) > criqc) {
                    const R qlr = R(qlj), qir = R(qij);
                    const R dw = ((qlr + qir) - criqc) * qlr / (qlr + qir);
                    const R di = ((qlr + qir) - criqc) * qir / (qlr + qir);
                    W(W_DWTEN, kpen) = dw;
                    W(W_DITEN, kpen) = di;
                    qtu_top = qtu_top - dw - di;
                    thlu_top = thlu_top + (xlv / cp / exntop) * dw + (xls / cp / exntop) * di;
                } else {
                    W(W_DWTEN, kpen) = R(0);
                    W(W_DITEN, kpen) = R(0);
                }
                const R rhos0j = R(in(a.ps0, kpen - 1)) / (r * R(0.5) * (Er(A_THV0BOT, kpen) + Er(A_THV0TOP, kpen - 1)) * Er(A_EXNS0, kpen - 1));
                cush = D(R(in(a.zs0, kpen - 1)) - ppen / rhos0j / g);
                scaleh = R(cush);
            }
        }
        if constexpr (PART == PART_SORT) {
            // the rest of the pass runs in the next launch: save what S11 set (the other
            // saved scalars still hold the closure's values) and stop
            resume_io([&](int f, auto& x) {
                const bool set_here = f == 0 || (f >= 18 && f <= 29) || f == 34 || f == 35 || f == 36 ||
                                      f == 37 || f == 39;
                if (!set_here) return;
                if constexpr (std::is_integral<std::remove_reference_t<decltype(x)>>::value) rstate(f) = (double)x;
                else rstate(f) = (double)val(x);
            });
            rstate(RS_CONT) = 1.0;
            return;
        }
        }   // S11

        // ---- S12 forcedCu filter (uwshcu.F90:2964-2992) ----
        if (kbup == krel) { exit_code = EX_CUFILTER; goto L333; }
        for (int k = lvl_lo(kpen); k <= mkx; ++k) {
            if (k < kpen) continue;
            out(a.umf, k) = 0.0; W(W_EMF, k) = R(0); W(W_UFRC, k) = R(0);
            if (k == kpen) continue;
            W(W_DWTEN, k) = R(0); W(W_DITEN, k) = R(0); W(W_FER, k) = R(0); W(W_FDR, k) = R(0);
        }

        // ---- S13 penetrative entrainment (uwshcu.F90:3055-3183) ----
        for (int k = 0; k <= mkx; ++k) {
            W(W_THLUE, k) = W(W_THLU, k);
            W(W_QTUE, k) = W(W_QTU, k);
            W(W_UUE, k) = W(W_UU, k);
            W(W_VUE, k) = W(W_VU, k);
            for (int m = 1; m <= ncnst; ++m) WT(2, m, k) = WT(1, m, k);
        }
        for (int k = lvl_hi(kpen - 1); k >= kbup; --k) {
            if (k > kpen - 1) continue;
            const R rhos0j = R(in(a.ps0, k)) / (r * R(0.5) * (Er(A_THV0BOT, k + 1) + Er(A_THV0TOP, k)) * Er(A_EXNS0, k));
            const R umfk = R(out(a.umf, k));
            if (k == kpen - 1) {
                const R x = umfk * ppen * W(W_REI, kpen) * rpen;
                const R lim2 = R(-0.9) * R(in(a.dp0, kpen)) / g / dt;
                if (x < R(-0.1) * rhos0j) limit_emf = 1;
                if (x < lim2) limit_emf = 1;
                W(W_EMF, k) = fmax(fmax(x, R(-0.1) * rhos0j), lim2);
                const R dpk = R(in(a.ps0, k) - in(a.p0, kpen));
                W(W_THLUE, k) = Er(A_THL0, kpen) + Er(A_SSTHL0, kpen) * dpk;
                W(W_QTUE, k) = Er(A_QT0, kpen) + Er(A_SSQT0, kpen) * dpk;
                W(W_UUE, k) = R(in(a.u0, kpen)) + Er(A_SSU0, kpen) * dpk;
                W(W_VUE, k) = R(in(a.v0, kpen)) + Er(A_SSV0, kpen) * dpk;
                for (int m = 1; m <= ncnst; ++m) WT(2, m, k) = R(trin(m, kpen)) + WT(0, m, kpen) * dpk;
            } else {
                const R emf1 = W(W_EMF, k + 1);
                const R x = emf1 - umfk * R(in(a.dp0, k + 1)) * W(W_REI, k + 1) * rpen;
                const R lim2 = R(-0.9) * R(in(a.dp0, k + 1)) / g / dt;
                if (x < R(-0.1) * rhos0j) limit_emf = 1;
                if (x < lim2) limit_emf = 1;
                const R emfk = fmax(fmax(x, R(-0.1) * rhos0j), lim2);
                W(W_EMF, k) = emfk;
                near(fabs(emfk) - fabs(emf1), fabs(emf1));
                if (fabs(emfk) > fabs(emf1)) {
                    W(W_THLUE, k) = (W(W_THLUE, k + 1) * emf1 + Er(A_THL0, k + 1) * (emfk - emf1)) / emfk;
                    W(W_QTUE, k) = (W(W_QTUE, k + 1) * emf1 + Er(A_QT0, k + 1) * (emfk - emf1)) / emfk;
                    W(W_UUE, k) = (W(W_UUE, k + 1) * emf1 + R(in(a.u0, k + 1)) * (emfk - emf1)) / emfk;
                    W(W_VUE, k) = (W(W_VUE, k + 1) * emf1 + R(in(a.v0, k + 1)) * (emfk - emf1)) / emfk;
                    for (int m = 1; m <= ncnst; ++m)
                        WT(2, m, k) = (WT(2, m, k + 1) * emf1 + R(trin(m, k + 1)) * (emfk - emf1)) / emfk;
                } else {
                    W(W_THLUE, k) = Er(A_THL0, k + 1);
                    W(W_QTUE, k) = Er(A_QT0, k + 1);
                    W(W_UUE, k) = R(in(a.u0, k + 1));
                    W(W_VUE, k) = R(in(a.v0, k + 1));
                    for (int m = 1; m <= ncnst; ++m) WT(2, m, k) = R(trin(m, k + 1));
                }
            }
        }

        // ---- S14 fluxes (uwshcu.F90:3255-3401) ----
        {
            // fluxbelowinv (uwshcu.F90:5030-5094) for one scalar; writes xflx(0..kinv-1) via put
            auto fluxbelowinv = [&](R xsrc, R xmean, R xtopin, R xbotin, auto put) {
                const R dp = R(in(a.ps0, kinv - 1) - in(a.ps0, kinv));
                R xbot, xtop;
                if (fabs(xbotin - xtopin) <= R(1.e-13)) { xbot = xbotin - R(1.e-13); xtop = xtopin + R(1.e-13); }
                else { xbot = xbotin; xtop = xtopin; }
                const R xtop_ori = xtop, xbot_ori = xbot;
                const R rcbmf = (cbmf * g * dt) / dp;
                R rpeff = (xmean - xtop) / (xbot - xtop);
                rpeff = fmin(fmax(R(0), rpeff), R(1));
                if (rpeff == R(0) || rpeff == R(1)) { xbot = xmean; xtop = xmean; }
                const R rr = rpeff / rcbmf;
                const R pinv = R(in(a.ps0, kinv - 1)) - rpeff * dp;
                const R ps00 = R(in(a.ps0, 0));
                for (int k = 0; k <= kinv - 1; ++k) {
                    R x = cbmf * (xsrc - xbot) * (ps00 - R(in(a.ps0, k))) / (ps00 - pinv);
                    if (k == kinv - 1 && rr <= R(1)) x = x - (R(1) - rr) * cbmf * (xtop_ori - xbot_ori);
                    put(k, x);
                }
            };
            const R dtop = R(in(a.ps0, kinv) - in(a.p0, kinv + 1));
            const R dbot = R(in(a.ps0, kinv - 1) - in(a.p0, kinv - 1));
            fluxbelowinv(R(qtsrc), Er(A_QT0, kinv), Er(A_QT0, kinv + 1) + Er(A_SSQT0, kinv + 1) * dtop,
                         Er(A_QT0, kinv - 1) + Er(A_SSQT0, kinv - 1) * dbot,
                         [&](int k, R x) { out(a.qtflx, k) = D(x); });
            fluxbelowinv(R(thlsrc), Er(A_THL0, kinv), Er(A_THL0, kinv + 1) + Er(A_SSTHL0, kinv + 1) * dtop,
                         Er(A_THL0, kinv - 1) + Er(A_SSTHL0, kinv - 1) * dbot,
                         [&](int k, R x) { out(a.slflx, k) = D(cp * Er(A_EXNS0, k) * x); });
            fluxbelowinv(R(usrc), R(in(a.u0, kinv)), R(in(a.u0, kinv + 1)) + Er(A_SSU0, kinv + 1) * dtop,
                         R(in(a.u0, kinv - 1)) + Er(A_SSU0, kinv - 1) * dbot,
                         [&](int k, R x) { W(W_UFLX, k) = x; });
            fluxbelowinv(R(vsrc), R(in(a.v0, kinv)), R(in(a.v0, kinv + 1)) + Er(A_SSV0, kinv + 1) * dtop,
                         R(in(a.v0, kinv - 1)) + Er(A_SSV0, kinv - 1) * dbot,
                         [&](int k, R x) { W(W_VFLX, k) = x; });
            for (int m = 1; m <= ncnst; ++m) {
                fluxbelowinv(R(trin(m, 1)), R(trin(m, kinv)), R(trin(m, kinv + 1)) + WT(0, m, kinv + 1) * dtop,
                             R(trin(m, kinv - 1)) + WT(0, m, kinv - 1) * dbot,
                             [&](int k, R x) { WT(3, m, k) = x; });
            }
        }
        {
            R uplus = R(0), vplus = R(0);
            for (int k = lvl_lo(kinv); k <= krel - 1; ++k) {
                if (k < kinv) continue;
                const int kp1 = k + 1;
                const R d = R(in(a.ps0, k) - in(a.p0, kp1));
                out(a.qtflx, k) = D(cbmf * (R(qtsrc) - (Er(A_QT0, kp1) + Er(A_SSQT0, kp1) * d)));
                out(a.slflx, k) = D(cbmf * (R(thlsrc) - (Er(A_THL0, kp1) + Er(A_SSTHL0, kp1) * d)) * cp * Er(A_EXNS0, k));
                uplus = uplus + PGFc * Er(A_SSU0, k) * R(in(a.ps0, k) - in(a.ps0, k - 1));
                vplus = vplus + PGFc * Er(A_SSV0, k) * R(in(a.ps0, k) - in(a.ps0, k - 1));
                W(W_UFLX, k) = cbmf * (R(usrc) + uplus - (R(in(a.u0, kp1)) + Er(A_SSU0, kp1) * d));
                W(W_VFLX, k) = cbmf * (R(vsrc) + vplus - (R(in(a.v0, kp1)) + Er(A_SSV0, kp1) * d));
                for (int m = 1; m <= ncnst; ++m)
                    WT(3, m, k) = cbmf * (R(trin(m, 1)) - (R(trin(m, kp1)) + WT(0, m, kp1) * d));
            }
        }
        for (int k = lvl_lo(krel); k <= kbup - 1; ++k) {
            if (k < krel) continue;
            const int kp1 = k + 1;
            const R d = R(in(a.ps0, k) - in(a.p0, kp1));
            const R umfk = R(out(a.umf, k));
            out(a.slflx, k) = D(cp * Er(A_EXNS0, k) * umfk * (W(W_THLU, k) - (Er(A_THL0, kp1) + Er(A_SSTHL0, kp1) * d)));
            out(a.qtflx, k) = D(umfk * (W(W_QTU, k) - (Er(A_QT0, kp1) + Er(A_SSQT0, kp1) * d)));
            W(W_UFLX, k) = umfk * (W(W_UU, k) - (R(in(a.u0, kp1)) + Er(A_SSU0, kp1) * d));
            W(W_VFLX, k) = umfk * (W(W_VU, k) - (R(in(a.v0, kp1)) + Er(A_SSV0, kp1) * d));
            for (int m = 1; m <= ncnst; ++m)
                WT(3, m, k) = umfk * (WT(1, m, k) - (R(trin(m, kp1)) + WT(0, m, kp1) * d));
        }
        for (int k = lvl_lo(kbup); k <= kpen - 1; ++k) {
            if (k < kbup) continue;
            const R d = R(in(a.ps0, k) - in(a.p0, k));
            const R emfk = W(W_EMF, k);
            out(a.slflx, k) = D(cp * Er(A_EXNS0, k) * emfk * (W(W_THLUE, k) - (Er(A_THL0, k) + Er(A_SSTHL0, k) * d)));
            out(a.qtflx, k) = D(emfk * (W(W_QTUE, k) - (Er(A_QT0, k) + Er(A_SSQT0, k) * d)));
            W(W_UFLX, k) = emfk * (W(W_UUE, k) - (R(in(a.u0, k)) + Er(A_SSU0, k) * d));
            W(W_VFLX, k) = emfk * (W(W_VUE, k) - (R(in(a.v0, k)) + Er(A_SSV0, k) * d));
            for (int m = 1; m <= ncnst; ++m)
                WT(3, m, k) = emfk * (WT(2, m, k) - (R(trin(m, k)) + WT(0, m, k) * d));
        }

        // ---- S15 subsidence condensate sinks (uwshcu.F90:3416-3476) ----
        {
            const R ps00 = R(in(a.ps0, 0)), psinv = R(in(a.ps0, kinv - 1));
            auto uemf = [&](int k) -> R {   // uwshcu.F90:3416-3422
                if (k <= kinv - 2) return cbmf * (ps00 - R(in(a.ps0, k))) / (ps00 - psinv);
                if (k <= krel - 1) return cbmf;
                if (k <= kbup - 1) return R(out(a.umf, k));
                if (k <= kpen - 1) return W(W_EMF, k);
                return R(0);
            };
            R uemf_km1 = uemf(0);
            for (int k = 1; k <= kpen; ++k) {
                const R uemf_k = uemf(k);
                const R comsub = R(0.5) * (uemf_k + uemf_km1);
                uemf_km1 = uemf_k;
                R thlten_sub, qtten_sub, qlten_sub, qiten_sub, nlten_sub, niten_sub;
                if (comsub >= R(0)) {
                    if (k == mkx) {
                        thlten_sub = qtten_sub = qlten_sub = qiten_sub = nlten_sub = niten_sub = R(0);
                    } else {
                        const R dp = R(in(a.p0, k) - in(a.p0, k + 1));
                        thlten_sub = g * comsub * (Er(A_THL0, k + 1) - Er(A_THL0, k)) / dp;
                        qtten_sub = g * comsub * (Er(A_QT0, k + 1) - Er(A_QT0, k)) / dp;
                        qlten_sub = g * comsub * (R(in(a.ql0, k + 1)) - R(in(a.ql0, k))) / dp;
                        qiten_sub = g * comsub * (R(in(a.qi0, k + 1)) - R(in(a.qi0, k))) / dp;
                        nlten_sub = g * comsub * (trz(ixl, k + 1) - trz(ixl, k)) / dp;
                        niten_sub = g * comsub * (trz(ixi, k + 1) - trz(ixi, k)) / dp;
                    }
                } else {
                    if (k == 1) {
                        thlten_sub = qtten_sub = qlten_sub = qiten_sub = nlten_sub = niten_sub = R(0);
                    } else {
                        const R dp = R(in(a.p0, k - 1) - in(a.p0, k));
                        thlten_sub = g * comsub * (Er(A_THL0, k) - Er(A_THL0, k - 1)) / dp;
                        qtten_sub = g * comsub * (Er(A_QT0, k) - Er(A_QT0, k - 1)) / dp;
                        qlten_sub = g * comsub * (R(in(a.ql0, k)) - R(in(a.ql0, k - 1))) / dp;
                        qiten_sub = g * comsub * (R(in(a.qi0, k)) - R(in(a.qi0, k - 1))) / dp;
                        nlten_sub = g * comsub * (trz(ixl, k) - trz(ixl, k - 1)) / dp;
                        niten_sub = g * comsub * (trz(ixi, k) - trz(ixi, k - 1)) / dp;
                    }
                }
                const R thl_prog = Er(A_THL0, k) + thlten_sub * dt;
                const R qt_prog = fmax(Er(A_QT0, k) + qtten_sub * dt, R(1.e-12));
                if (T.conden_x(in(a.p0, k), E(A_EXN0, k), thl_prog, qt_prog, th, qvj, qlj, qij)) { exit_code = EX_SUBCONDEN; goto L333; }
                W(W_QLSINK, k) = fmax(qlten_sub, -R(in(a.ql0, k)) / dt);
                W(W_QISINK, k) = fmax(qiten_sub, -R(in(a.qi0, k)) / dt);
                W(W_NLSINK, k) = fmax(nlten_sub, -trz(ixl, k) / dt);
                W(W_NISINK, k) = fmax(niten_sub, -trz(ixi, k) / dt);
            }
        }

        // ---- S16 tendencies (uwshcu.F90:3488-3794) ----
        for (int k = 1; k <= kpen; ++k) {
            const R gdp = g / R(in(a.dp0, k));
            out(a.uten, k) = D((W(W_UFLX, k - 1) - W(W_UFLX, k)) * g / R(in(a.dp0, k)));
            out(a.vten, k) = D((W(W_VFLX, k - 1) - W(W_VFLX, k)) * g / R(in(a.dp0, k)));
            (void)gdp;
        }
        R rliq = R(0), rainflx = R(0), snowflx = R(0);
        R qlr_base = R(0), qir_base = R(0);   // updraft condensate at the release level
        {
            auto uf = [&](int k) -> R { return R(in(a.u0, k)) + R(out(a.uten, k)) * dt; };
            auto vf = [&](int k) -> R { return R(in(a.v0, k)) + R(out(a.vten, k)) * dt; };
            R qlubelow = R(0), qiubelow = R(0);
            for (int k = 1; k <= kpen; ++k) {
                const int km1 = k - 1;
                const R dp0k = R(in(a.dp0, k));
                const R umf_km1 = R(out(a.umf, k - 1)), umf_k = R(out(a.umf, k));
                const R dwk = W(W_DWTEN, k) * R(0.5) * (umf_km1 + umf_k) * g / dp0k;
                const R dik = W(W_DITEN, k) * R(0.5) * (umf_km1 + umf_k) * g / dp0k;
                W(W_DWTEN, k) = dwk;
                W(W_DITEN, k) = dik;
                const R qrk = frc_rasn * dwk, qsk = frc_rasn * dik;
                out(a.qrten, k) = D(qrk);
                out(a.qsten, k) = D(qsk);
                rainflx = rainflx + qrk * dp0k / g;
                snowflx = snowflx + qsk * dp0k / g;
                R slten = (R(out(a.slflx, km1)) - R(out(a.slflx, k))) * g / dp0k;
                const R u0k = R(in(a.u0, k)), v0k = R(in(a.v0, k));
                if (k == 1) {
                    slten = slten - g / R(4) / dp0k *
                                        (W(W_UFLX, k) * (uf(k + 1) - uf(k) + R(in(a.u0, k + 1)) - u0k) +
                                         W(W_VFLX, k) * (vf(k + 1) - vf(k) + R(in(a.v0, k + 1)) - v0k));
                } else if (k >= 2 && k <= kpen - 1) {
                    slten = slten - g / R(4) / dp0k *
                                        (W(W_UFLX, k) * (uf(k + 1) - uf(k) + R(in(a.u0, k + 1)) - u0k) +
                                         W(W_UFLX, k - 1) * (uf(k) - uf(k - 1) + u0k - R(in(a.u0, k - 1))) +
                                         W(W_VFLX, k) * (vf(k + 1) - vf(k) + R(in(a.v0, k + 1)) - v0k) +
                                         W(W_VFLX, k - 1) * (vf(k) - vf(k - 1) + v0k - R(in(a.v0, k - 1))));
                } else if (k == kpen) {
                    slten = slten - g / R(4) / dp0k *
                                        (W(W_UFLX, k - 1) * (uf(k) - uf(k - 1) + u0k - R(in(a.u0, k - 1))) +
                                         W(W_VFLX, k - 1) * (vf(k) - vf(k - 1) + v0k - R(in(a.v0, k - 1))));
                }
                const R qtten = (R(out(a.qtflx, km1)) - R(out(a.qtflx, k))) * g / dp0k;

                R qlu_mid, qiu_mid, qlr, qir;
                if (k < krel || k > kpen) {
                    qlu_mid = R(0); qiu_mid = R(0); qlr = R(0); qir = R(0);
                } else if (k == krel) {
                    if (T.conden(prel, W(W_THLU, krel - 1), W(W_QTU, krel - 1), th, qvj, qlj, qij)) { exit_code = EX_CONDEN; goto L333; }
                    qlubelow = R(qlj);
                    qiubelow = R(qij);
                    qlr_base = qlubelow;
                    qir_base = qiubelow;
                    if (T.conden_x(in(a.ps0, k), E(A_EXNS0, k), W(W_THLU, k), W(W_QTU, k), th, qvj, qlj, qij)) { exit_code = EX_CONDEN; goto L333; }
                    qlr = R(qlj); qir = R(qij);
                    const R fr = (prel - R(in(a.ps0, k))) / R(in(a.ps0, k - 1) - in(a.ps0, k));
                    qlu_mid = R(0.5) * (qlubelow + qlr) * (prel - R(in(a.ps0, k))) / R(in(a.ps0, k - 1) - in(a.ps0, k));
                    qiu_mid = R(0.5) * (qiubelow + qir) * (prel - R(in(a.ps0, k))) / R(in(a.ps0, k - 1) - in(a.ps0, k));
                    (void)fr;
                } else if (k == kpen) {
                    if (T.conden(R(in(a.ps0, k - 1)) + ppen, thlu_top, qtu_top, th, qvj, qlj, qij)) { exit_code = EX_CONDEN; goto L333; }
                    qlr = R(qlj); qir = R(qij);
                    qlu_mid = R(0.5) * (qlubelow + qlr) * (-ppen) / R(in(a.ps0, k - 1) - in(a.ps0, k));
                    qiu_mid = R(0.5) * (qiubelow + qir) * (-ppen) / R(in(a.ps0, k - 1) - in(a.ps0, k));
                } else {
                    if (T.conden_x(in(a.ps0, k), E(A_EXNS0, k), W(W_THLU, k), W(W_QTU, k), th, qvj, qlj, qij)) { exit_code = EX_CONDEN; goto L333; }
                    qlr = R(qlj); qir = R(qij);
                    qlu_mid = R(0.5) * (qlubelow + qlr);
                    qiu_mid = R(0.5) * (qiubelow + qir);
                }
                qlubelow = qlr;
                qiubelow = qir;
                if (k >= krel) { W(W_QLR, k) = qlr; W(W_QIR, k) = qir; }

                R qc_l = (R(1) - frc_rasn) * dwk;
                R qc_i = (R(1) - frc_rasn) * dik;
                R qc_lm, qc_im, nc_lm, nc_im;
                const R ql0k = R(in(a.ql0, k)), qi0k = R(in(a.qi0, k));
                if (k <= kbup) {
                    const R fdrk = W(W_FDR, k);
                    qc_l = qc_l + g * R(0.5) * (umf_km1 + umf_k) * fdrk * qlu_mid;
                    qc_i = qc_i + g * R(0.5) * (umf_km1 + umf_k) * fdrk * qiu_mid;
                    qc_lm = -g * R(0.5) * (umf_km1 + umf_k) * fdrk * ql0k;
                    qc_im = -g * R(0.5) * (umf_km1 + umf_k) * fdrk * qi0k;
                    nc_lm = -g * R(0.5) * (umf_km1 + umf_k) * fdrk * trz(ixl, k);
                    nc_im = -g * R(0.5) * (umf_km1 + umf_k) * fdrk * trz(ixi, k);
                } else {
                    qc_lm = R(0); qc_im = R(0); nc_lm = R(0); nc_im = R(0);
                }
                if (k == kbup) {
                    const R dps = R(in(a.ps0, k - 1) - in(a.ps0, k));
                    qc_l = qc_l + g * umf_k * qlr / dps;
                    qc_i = qc_i + g * umf_k * qir / dps;
                    qc_lm = qc_lm - g * umf_k * ql0k / dps;
                    qc_im = qc_im - g * umf_k * qi0k / dps;
                    nc_lm = nc_lm - g * umf_k * trz(ixl, k) / dps;
                    nc_im = nc_im - g * umf_k * trz(ixi, k) / dps;
                    D ql_e, qi_e;
                    if (T.conden_x(in(a.p0, k), E(A_EXN0, k), W(W_THLUE, k), W(W_QTUE, k), th, qvj, ql_e, qi_e)) { exit_code = EX_KBUPCONDEN; goto L333; }
                    const R nl_e = (ql_e > 0.0 && has_nl) ? WT(2, ixl, k) : R(0);
                    const R ni_e = (qi_e > 0.0 && has_ni) ? WT(2, ixi, k) : R(0);
                    const R emfk = W(W_EMF, k);
                    qc_lm = qc_lm - g * emfk * (R(ql_e) - ql0k) / dps;
                    qc_im = qc_im - g * emfk * (R(qi_e) - qi0k) / dps;
                    nc_lm = nc_lm - g * emfk * (nl_e - trz(ixl, k)) / dps;
                    nc_im = nc_im - g * emfk * (ni_e - trz(ixi, k)) / dps;
                }
                const R qltenk = qc_l + frc_rasn * dwk + (fmax(R(0), ql0k + (qc_lm + W(W_QLSINK, k)) * dt) - ql0k) / dt;
                const R qitenk = qc_i + frc_rasn * dik + (fmax(R(0), qi0k + (qc_im + W(W_QISINK, k)) * dt) - qi0k) / dt;
                if (has_nl) trout(ixl, k) = D(fmax(nc_lm + W(W_NLSINK, k), -trz(ixl, k) / dt));
                if (has_ni) trout(ixi, k) = D(fmax(nc_im + W(W_NISINK, k), -trz(ixi, k) / dt));
                out(a.qlten, k) = D(qltenk);
                out(a.qiten, k) = D(qitenk);
                out(a.qvten, k) = D(qtten - qltenk - qitenk);
                out(a.sten, k) = D(slten + xlv * qltenk + xls * qitenk);
                W(W_QCL, k) = qc_l;
                W(W_QCI, k) = qc_i;
                const R qck = qc_l + qc_i;
                out(a.qc, k) = D(qck);
                rliq = rliq + qck * dp0k / g / R(1000);
            }
        }

        // ---- S17 precipitation evaporation (uwshcu.F90:3808-3934) ----
        R flxrain_k = R(0), flxsnow_k = R(0);   // fluxes at the top interface of layer k
        {
            R evpint_rain = R(0), evpint_snow = R(0);
            const bool final_pass = (iter == 2);
            if (final_pass) { out(a.flxprc1, mkx) = 0.0; out(a.flxsnow1, mkx) = 0.0; }
            for (int k = mkx; k >= 1; --k) {
                const R dp0k = R(in(a.dp0, k));
                const R t0k = R(in_t0(k)), qv0k = R(in_qv0(k));
                R snowmlt = (t0k > R(273.16)) ? fmax(R(0), flxsnow_k * g / dp0k) : R(0);
                D es, qs;
                T.qsat_es_qs(in_t0(k), in(a.p0, k), es, qs);
                const R subsat = fmax((R(1) - qv0k / R(qs)), R(0));
                R evprain = kevp * subsat * sqrt(flxrain_k + snowmlt * dp0k / g);
                R evpsnow = kevp * subsat * sqrt(fmax(flxsnow_k - snowmlt * dp0k / g, R(0)));
                const R evplimit = fmax(R(0), (W(W_QW0, k) - qv0k) / dt);
                R evplimit_rain = fmin(evplimit, (flxrain_k + snowmlt * dp0k / g) * g / dp0k);
                evplimit_rain = fmin(evplimit_rain, (rainflx - evpint_rain) * g / dp0k);
                evprain = fmax(R(0), fmin(evplimit_rain, evprain));
                R evplimit_snow = fmin(evplimit, fmax(flxsnow_k - snowmlt * dp0k / g, R(0)) * g / dp0k);
                evplimit_snow = fmin(evplimit_snow, (snowflx - evpint_snow) * g / dp0k);
                evpsnow = fmax(R(0), fmin(evplimit_snow, evpsnow));
                if ((evprain + evpsnow) > evplimit) {
                    const R tmp1 = evprain * evplimit / (evprain + evpsnow);
                    const R tmp2 = evpsnow * evplimit / (evprain + evpsnow);
                    evprain = tmp1;
                    evpsnow = tmp2;
                }
                out(a.evapc, k) = D(evprain + evpsnow);
                evpint_rain = evpint_rain + evprain * dp0k / g;
                evpint_snow = evpint_snow + evpsnow * dp0k / g;
                const R qrk = R(out(a.qrten, k)), qsk = R(out(a.qsten, k));
                const R ntraprd = qrk - evprain + snowmlt;
                const R ntsnprd = qsk - evpsnow - snowmlt;
                R flxrain_km1 = flxrain_k + ntraprd * dp0k / g;
                R flxsnow_km1 = flxsnow_k + ntsnprd * dp0k / g;
                flxrain_km1 = fmax(flxrain_km1, R(0));
                flxsnow_km1 = fmax(flxsnow_km1, R(0));
                flxrain_k = flxrain_km1;
                flxsnow_k = flxsnow_km1;
                if (final_pass) {
                    out(a.flxprc1, k - 1) = D(flxrain_km1 + flxsnow_km1);
                    out(a.flxsnow1, k - 1) = D(flxsnow_km1);
                }
                const R qlt = R(out(a.qlten, k)) - qrk;
                const R qit = R(out(a.qiten, k)) - qsk;
                const R qvt = R(out(a.qvten, k)) + evprain + evpsnow;
                out(a.qlten, k) = D(qlt);
                out(a.qiten, k) = D(qit);
                out(a.qvten, k) = D(qvt);
                out(a.sten, k) = D(R(out(a.sten, k)) - xlv * evprain - xls * evpsnow - (xls - xlv) * snowmlt);
            }
        }
        const R precip = (flxrain_k + flxsnow_k) / R(1000);
        const R snow = flxsnow_k / R(1000);

        // ---- S18 reserved condensate (uwshcu.F90:3966-3981) ----
        for (int k = 1; k <= kpen; ++k) {
            out(a.qlten, k) = D(R(out(a.qlten, k)) - W(W_QCL, k));
            out(a.qiten, k) = D(R(out(a.qiten, k)) - W(W_QCI, k));
            out(a.sten, k) = D(R(out(a.sten, k)) - (xls - xlv) * W(W_QCI, k));
        }

        // ---- S19 positive moisture, FP64 (uwshcu.F90:3995-4001, 5096-5162) ----
        {
            const double qvmin = C.qmin[1], qlmin = C.qmin[2], qimin = C.qmin[3];
            const double xlvd = C.xlv, xlsd = C.xls;
            D dqv = 0.0;
            D carry_qv = 0.0, carry_ten = 0.0;   // adjustments of qv(k-1) and qvten(k-1) from level k
            bool have_carry = false;
            for (int k = mkx; k >= 1; --k) {
                D qvten = out(a.qvten, k), qlten = out(a.qlten, k), qiten = out(a.qiten, k), sten = out(a.sten, k);
                D qv = in_qv0(k) + qvten * dtd;
                D ql = in(a.ql0, k) + qlten * dtd;
                D qi = in(a.qi0, k) + qiten * dtd;
                if (have_carry) {
                    qv = qv - carry_qv;
                    qvten = qvten - carry_ten;
                }
                const D dql = fmax(0.0, 1.0 * qlmin - ql);
                const D dqi = fmax(0.0, 1.0 * qimin - qi);
                qlten = qlten + dql / dtd;
                qiten = qiten + dqi / dtd;
                qvten = qvten - (dql + dqi) / dtd;
                sten = sten + xlvd * (dql / dtd) + xlsd * (dqi / dtd);
                ql = ql + dql;
                qi = qi + dqi;
                qv = qv - dql - dqi;
                dqv = fmax(0.0, 1.0 * qvmin - qv);
                qvten = qvten + dqv / dtd;
                qv = qv + dqv;
                if (k != 1) {
                    const D dpk = in(a.dp0, k), dpkm1 = in(a.dp0, k - 1);
                    carry_qv = dqv * dpk / dpkm1;
                    carry_ten = dqv * dpk / dpkm1 / dtd;
                    have_carry = true;
                }
                qv = fmax(qv, qvmin);
                out(a.qvten, k) = qvten;
                out(a.qlten, k) = qlten;
                out(a.qiten, k) = qiten;
                out(a.sten, k) = sten;
                E(E_QVSTAR, k) = qv;
            }
            if (dqv > 1.e-20) {
                D sum = 0.0;
                for (int k = 1; k <= mkx; ++k) {
                    const D qv = E(E_QVSTAR, k);
                    if (qv > 2.0 * qvmin) sum = sum + qv * in(a.dp0, k);
                }
                const D aa = dqv * in(a.dp0, 1) / fmax(1.e-20, sum);
                if (aa < 0.5) {
                    for (int k = 1; k <= mkx; ++k) {
                        const D qv = E(E_QVSTAR, k);
                        if (qv > 2.0 * qvmin) {
                            const D dum = aa * qv;
                            out(a.qvten, k) = out(a.qvten, k) - dum / dtd;
                        }
                    }
                }
            }
        }

        // ---- S20 constituent limiter (uwshcu.F90:4007-4076) ----
        for (int m = 4; m <= ncnst; ++m) {
            if (m == ixl || m == ixi) continue;
            const R trmin = R(C.qmin[m]);
            const bool wet = C.wet[m] != 0;
            auto pdelx = [&](int k) -> R { return wet ? R(in(a.dp0, k)) : R(in(a.dpdry0, k)); };
            W(W_TRFD, 0) = R(0);
            for (int k = 1; k <= mkx - 1; ++k) {
                const R dum = (R(trin(m, k)) - trmin) * pdelx(k) / g / dt + WT(3, m, k - 1) - WT(3, m, k) + W(W_TRFD, k - 1);
                W(W_TRFD, k) = fmin(R(0), dum);
            }
            W(W_TRFD, mkx) = R(0);
            W(W_TRFU, mkx) = R(0);
            W(W_TRFU, 0) = R(0);
            for (int k = mkx; k >= 2; --k) {
                const R dum = (R(trin(m, k)) - trmin) * pdelx(k) / g / dt + WT(3, m, k - 1) - WT(3, m, k) +
                              W(W_TRFD, k - 1) - W(W_TRFD, k) - W(W_TRFU, k);
                W(W_TRFU, k - 1) = fmax(R(0), -dum);
            }
            for (int k = 1; k <= mkx; ++k) {
                trout(m, k) = D((WT(3, m, k - 1) - WT(3, m, k) + W(W_TRFD, k - 1) - W(W_TRFD, k) +
                                        W(W_TRFU, k - 1) - W(W_TRFU, k)) * g / pdelx(k));
            }
        }

        // ---- S21 in-cloud diagnostics (uwshcu.F90:4092-4158) ----
        {
            // the tendency stage adjusted the same updraft states: reuse its condensate
            R qcubelow = qlr_base + qir_base, qlubelow = qlr_base, qiubelow = qir_base;
            for (int k = lvl_lo(krel); k <= kpen; ++k) {
                if (k < krel) continue;
                const R qlr = W(W_QLR, k), qir = W(W_QIR, k);
                R qcuk = R(0.5) * (qcubelow + qlr + qir);
                R qluk = R(0.5) * (qlubelow + qlr);
                R qiuk = R(0.5) * (qiubelow + qir);
                R tmp2 = R(0), tmp3 = R(1);
                if constexpr (ZW) {
                    if (a.zw.cloud_sh) {
                        // zwshcu.F90:4412-4429: cumulus fraction widened by detrainment
                        tmp3 = R(0.5) * R(a.zw.ufrc0);
                        R tmp1 = (R(in(a.z0, k)) - R(in(a.z0, krel))) / (R(in(a.z0, kpen)) - R(in(a.z0, krel)) + R(10));
                        tmp1 = R(1) + R(a.zw.ufrcz) * tmp1;
                        tmp3 = tmp3 * tmp1;
                        tmp2 = W(W_FDR, k) * R(out(a.umf, k)) * R(in(a.zs0, k) - in(a.zs0, k - 1)) * dt /
                               R(in(a.dp0, k)) * R(9.8) * R(a.zw.ufrcd);
                        tmp2 = fmin(R(1) - rmaxfrac, tmp2);
                    }
                }
                R cufrck = (W(W_UFRC, k - 1) + W(W_UFRC, k));
                if (k == krel) {
                    cufrck = (ufrclcl + W(W_UFRC, k)) * (prel - R(in(a.ps0, k))) / R(in(a.ps0, k - 1) - in(a.ps0, k));
                } else if (k == kpen) {
                    cufrck = (W(W_UFRC, k - 1) + R(0)) * (-ppen) / R(in(a.ps0, k - 1) - in(a.ps0, k));
                    if ((qlr + qir) > criqc) {
                        qcuk = R(0.5) * (qcubelow + criqc);
                        qluk = R(0.5) * (qlubelow + criqc * qlr / (qlr + qir));
                        qiuk = R(0.5) * (qiubelow + criqc * qir / (qlr + qir));
                    }
                }
                out(a.qcu, k) = D(qcuk);
                out(a.qlu, k) = D(qluk);
                out(a.qiu, k) = D(qiuk);
                if constexpr (ZW) cufrck = tmp3 * cufrck + tmp2;   // zwshcu.F90:4449
                out(a.cufrc, k) = D(cufrck);
                qcubelow = qlr + qir;
                qlubelow = qlr;
                qiubelow = qir;
            }
        }

        // column scalars of this pass
        cout(a.precip) = D(precip);
        cout(a.snow) = D(snow);
        cout(a.cbmf) = D(cbmf);
        cout(a.rliq) = D(rliq);
        cout(a.cnt) = D(mkx + 1 - kpen);
        cout(a.cnb) = D(mkx + 1 - (krel - 1));

        if constexpr (PART == PART_REST) {
            // S22 runs level-parallel (setup_level_b1/b2) and the implicit-CIN iteration in
            // a launch of its own: save what S12-S21 set and stop
            resume_io([&](int f, auto& x) {
                if (!((f >= 26 && f <= 29) || f == 38 || f == 39)) return;
                if constexpr (std::is_integral<std::remove_reference_t<decltype(x)>>::value) rstate(f) = (double)x;
                else rstate(f) = (double)val(x);
            });
            rstate(RS_CONT) = 1.0;
            return;
        }

        // ---- S22 provisional state and environment set B (uwshcu.F90:4169-4329) ----
        if (iter == 1) {
            cush_s = cush;
            for (int k = 1; k <= mkx; ++k) {
                const D qv0s = in_qv0(k) + out(a.qvten, k) * dtd;
                const D ql0s = in(a.ql0, k) + out(a.qlten, k) * dtd;
                const D qi0s = in(a.qi0, k) + out(a.qiten, k) * dtd;
                const D t0s = in_t0(k) + out(a.sten, k) * dtd / C.cp;
                const D qt0 = (qv0s + ql0s + qi0s);
                E(B_QT0, k) = qt0;
                E(B_THL0, k) = (t0s - C.xlv * ql0s / C.cp - C.xls * qi0s / C.cp) / E(A_EXN0, k);
            }
            slope([&](int k) { return E(B_THL0, k); }, B_SSTHL0);
            slope([&](int k) { return E(B_QT0, k); }, B_SSQT0);
            if (!interface_env(B_THL0)) { exit_code = EX_CONDEN; goto L333; }
        }
        }
    }

    // successful implicit-CIN pass
    cout(a.cush) = cush;
    path = 2;
    goto Ldone;

L333:
    write_exit(mkx);
    path = 0;
    if constexpr (PART == PART_SORT || PART == PART_REST) rstate(RS_CONT) = 0.0;

Ldone:
    if (a.amb_flag) a.amb_flag[i] = amb;
    if (a.diag) {
        const size_t L = (size_t)a.ldc;
        int* d = a.diag + i;
        d[0 * L] = (path == 0) ? exit_code : 0;
        d[1 * L] = kinv; d[2 * L] = klcl; d[3 * L] = klfc; d[4 * L] = krel; d[5 * L] = kbup; d[6 * L] = kpen;
        d[7 * L] = path; d[8 * L] = limit_ufrc; d[9 * L] = limit_ppen; d[10 * L] = limit_emf; d[11 * L] = exit_iter;
    }
}

}  // namespace uwshcu
#endif
