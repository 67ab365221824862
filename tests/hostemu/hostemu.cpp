// hostemu.cpp -- TEST HARNESS: runs the product's column code (uwshcu_column.h) on the
// CPU, one column at a time, so that its logic can be compared with the oracle in a
// container without a GPU.  It is built only by the test suite and never shipped;
// the library the product loads contains no CPU path.
#include <cstdlib>
#include <cstring>
#include <vector>

#include "uwshcu_adjoint.h"
#include "uwshcu_check.h"

using namespace uwshcu;

// 1: run the level-parallel set-up (setup_level1/2 over every column and level)
// before the columns, as the GPU pipeline does; 0: the columns do it themselves.
static int g_staged = 0;
// 1: as the GPU pipeline does for large launches, stop the triggered columns at the
// closure of their first pass and continue them from a list, then the same for the
// columns that need a second convection pass
static int g_defer = 0;
static int g_deferred = 0;   // columns the last deferring run continued at their second pass
static int g_triggered = 0;  // columns the last deferring run continued at their closure
extern "C" void hostemu_set_defer(int on) { g_defer = on; }
extern "C" int hostemu_last_deferred() { return g_deferred; }
extern "C" int hostemu_last_triggered() { return g_triggered; }
// backward-sweep window (entries) used by hostemu_adjoint; 0 = none
static int g_window = 0;
extern "C" void hostemu_set_window(int w) { g_window = w; }
extern "C" void hostemu_set_staged(int on) { g_staged = on; }
// hostemu_adjoint records paged tapes (the GPU's layout: pages of kTapePage nodes from a
// pool, here one pool per thread) instead of one contiguous tape per column
static int g_paged = 0;
extern "C" void hostemu_set_paged(int on) { g_paged = on; }
// the modified scheme (Column<double, double, true>) for the next hostemu_forward
// calls while set; zw == nullptr returns to the UW scheme
static const uwshcu_zw_params* g_zw = nullptr;
static const double *g_omega = nullptr, *g_tpert = nullptr, *g_qpert = nullptr;
extern "C" void hostemu_set_zw(const uwshcu_zw_params* zw, const double* omega, const double* tpert,
                               const double* qpert) {
    g_zw = zw; g_omega = omega; g_tpert = tpert; g_qpert = qpert;
}

template <class R, bool ZW = false, bool FAST = false>
static void run_columns(FwdArgs& a, const ThermoT<double, FAST>& T, int ncol, int pver) {
    using Col = Column<R, double, ZW, FAST>;
    if (g_staged) {
        a.env_ready = 1;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < ncol; ++i)
            for (int k = 0; k <= pver; ++k) Col(a, T, i, i).setup_level1(k);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < ncol; ++i)
            for (int k = 1; k <= pver; ++k) Col(a, T, i, i).setup_level2(k);
    }
    // the GPU pipeline's split of large launches: (1) every column up to the
    // cloud-base closure of its first pass, listing the triggered ones; (2) those from
    // there, listing the columns that need a second convection pass; (3) those
    std::vector<int> listed((size_t)ncol), second((size_t)ncol);
    std::vector<double> state((size_t)NRESUME * (size_t)a.wld, -777.0);
    int nlisted = 0, nsecond = 0;
    if (g_defer) {
        a.defer_list = listed.data();
        a.defer_count = &nlisted;
        a.defer_state = state.data();
        a.defer_at_closure = 1;
    }
    if (g_defer) {
#pragma omp parallel for schedule(dynamic, 16)
        for (int i = 0; i < ncol; ++i) Col(a, T, i, i).template run_part<PART_CLOSURE>();
    } else {
#pragma omp parallel for schedule(dynamic, 16)
        for (int i = 0; i < ncol; ++i) Col(a, T, i, i).run();
    }
    a.env_ready = 0;
    a.defer_list = nullptr;
    a.defer_count = nullptr;
    a.defer_at_closure = 0;
    if (g_defer) {
        // continue the listed columns in their own slots, whose set-A environment the
        // first launch (or the level-parallel set-up) left in place: the buoyancy
        // sorting, then the rest of the first pass, listing second passes
        a.env_ready = 1;
        a.slot0 = 0;
        a.resume = 1;
#pragma omp parallel for schedule(dynamic, 4)
        for (int j = 0; j < nlisted; ++j) Col(a, T, listed[j], listed[j]).template run_part<PART_SORT>();
        a.resume = 3;
#pragma omp parallel for schedule(dynamic, 4)
        for (int j = 0; j < nlisted; ++j) Col(a, T, listed[j], listed[j]).template run_part<PART_REST>();
        // the set-B environment level by level, then the implicit-CIN iteration, listing
        // second passes
#pragma omp parallel for schedule(static)
        for (int j = 0; j < nlisted; ++j) {
            const Col c(a, T, listed[j], listed[j]);
            if (c.rstate(RS_CONT) == 0.0) continue;
            for (int k = 1; k <= pver; ++k) c.setup_level_b1(k);
            for (int k = 1; k <= pver; ++k) c.setup_level_b2(k);
        }
        a.resume = 4;
        a.defer_list = second.data();
        a.defer_count = &nsecond;
#pragma omp parallel for schedule(dynamic, 4)
        for (int j = 0; j < nlisted; ++j) Col(a, T, listed[j], listed[j]).template run_part<PART_CIN2>();
        a.defer_list = nullptr;
        a.defer_count = nullptr;
        a.resume = 2;
#pragma omp parallel for schedule(dynamic, 4)
        for (int j = 0; j < nsecond; ++j) Col(a, T, second[j], second[j]).run();
        g_deferred = nsecond;
        g_triggered = nlisted;
        a.resume = 0;
        a.env_ready = 0;
        a.defer_state = nullptr;
    }
}

extern "C" int hostemu_forward(const uwshcu_consts* c, const double* estbl, int precision, const uwshcu_params* par,
                               int pcols, int ncol, int pver, int ncnst, double dt, const double* ps0,
                               const double* zs0, const double* p0, const double* z0, const double* dp0,
                               const double* u0, const double* v0, const double* qv0, const double* ql0,
                               const double* qi0, const double* t0, const double* s0, const double* tr0,
                               const double* tke, const double* pblh, double* cush, double* umf, double* slflx,
                               double* qtflx, double* flxprc1, double* flxsnow1, double* qvten, double* qlten,
                               double* qiten, double* sten, double* uten, double* vten, double* trten,
                               double* qrten, double* qsten, double* precip, double* snow, double* evapc,
                               double* cufrc, double* qcu, double* qlu, double* qiu, double* cbmf, double* qc,
                               double* rliq, double* cnt, double* cnb, const double* dpdry0, double* wu, int* diag,
                               int* bad_col, int* nfix) {
    FwdArgs a{};
    a.ncol = ncol; a.mkx = pver; a.ncnst = ncnst; a.ldc = pcols; a.col0 = 0; a.dt = dt; a.par = *par;
    a.ps0 = ps0; a.zs0 = zs0; a.p0 = p0; a.z0 = z0; a.dp0 = dp0; a.dpdry0 = dpdry0; a.u0 = u0; a.v0 = v0;
    a.qv0 = qv0; a.ql0 = ql0; a.qi0 = qi0; a.t0 = t0; a.s0 = s0; a.tr0 = tr0; a.tke = tke; a.pblh = pblh;
    a.cush = cush; a.umf = umf; a.slflx = slflx; a.qtflx = qtflx; a.flxprc1 = flxprc1; a.flxsnow1 = flxsnow1;
    a.qvten = qvten; a.qlten = qlten; a.qiten = qiten; a.sten = sten; a.uten = uten; a.vten = vten;
    a.trten = trten; a.qrten = qrten; a.qsten = qsten; a.precip = precip; a.snow = snow; a.evapc = evapc;
    a.cufrc = cufrc; a.qcu = qcu; a.qlu = qlu; a.qiu = qiu; a.cbmf = cbmf; a.qc = qc; a.rliq = rliq;
    a.cnt = cnt; a.cnb = cnb; a.wu = wu; a.diag = diag;
    int err[2] = {0, -1};
    a.err = err;
    // precision: UWSHCU_FP64, UWSHCU_MIXED (the fast FP64 pass, conservation/branch
    // check, FP64 recompute of the listed columns -- the product's mixed pipeline, run
    // serially here), or UWSHCU_MIXED_UNCHECKED (the FP32 pass alone).
    const bool fp32 = precision == UWSHCU_MIXED_UNCHECKED;
    const bool fast = precision == UWSHCU_MIXED || precision == UWSHCU_MIXED_FAST_UNCHECKED;
    a.wld = ncol;
    std::vector<char> ws(ws_bytes_per_column(pver, ncnst, sizeof(double)) * (size_t)ncol + 64);
    std::memset(ws.data(), 0xff, ws.size());   // NaN pattern: unwritten workspace reads show up
    a.ws = ws.data();
    std::vector<int> amb(ncol, 0);
    std::vector<double> cush_in(cush, cush + ncol);
    a.amb_flag = (fp32 || fast) ? amb.data() : nullptr;
    const Thermo T{c, estbl};
    const ThermoFast TF{c, estbl};
    if (g_zw) {
        if (fp32 || fast) return UWSHCU_ERR_ARG;
        a.zw = *g_zw; a.omega = g_omega; a.tpert = g_tpert; a.qpert = g_qpert;
        run_columns<double, true>(a, T, ncol, pver);
    } else if (fast) {
        run_columns<double, false, true>(a, TF, ncol, pver);
    } else if (!fp32) {
        run_columns<double>(a, T, ncol, pver);
    } else {
        run_columns<float>(a, T, ncol, pver);
    }
    int nlist = 0;
    if (precision == UWSHCU_MIXED) {
        std::vector<int> list;
        for (int i = 0; i < ncol; ++i)
            if (amb[i] || column_fails_check(a, *c, i, 1.e-4)) list.push_back(i);
        nlist = (int)list.size();
        a.amb_flag = nullptr;
        a.cush_in = cush_in.data();
        a.colmap = list.data();
        a.colcount = &nlist;
#pragma omp parallel for schedule(dynamic, 16)
        for (int j = 0; j < nlist; ++j) Column<double>(a, T, list[j], j).run();
    }
    if (nfix) *nfix = nlist;
    if (bad_col) *bad_col = err[0] ? err[1] : -1;
    return err[0];
}

// Adjoint on the host: records each column's FP64 forward pass on a tape
// (Column<AD, AD>), writes the forward outputs, seeds the outputs' adjoints with
// cot[] (ADF_TR layer/interface fields, trten, ADC_COUNT column scalars; null =
// zero) and writes the input gradients grad[] = {t0, qv0, u0, v0, tke, pblh, par}.
// tape_len[i] receives the tape length of column i (-1 if it overflowed cap).
extern "C" int hostemu_adjoint(const uwshcu_consts* c, const double* estbl, const uwshcu_params* par, int pcols,
                               int ncol, int pver, int ncnst, double dt, const double* ps0, const double* zs0,
                               const double* p0, const double* z0, const double* dp0, const double* u0,
                               const double* v0, const double* qv0, const double* ql0, const double* qi0,
                               const double* t0, const double* s0, const double* tr0, const double* tke,
                               const double* pblh, double* cush, double* const* outs, const double* dpdry0,
                               double* wu, const double* const* cot, double* const* grad, int cap, int* tape_len,
                               int* diag) {
    FwdArgs a{};
    a.ncol = ncol; a.mkx = pver; a.ncnst = ncnst; a.ldc = pcols; a.col0 = 0; a.dt = dt; a.par = *par;
    a.ps0 = ps0; a.zs0 = zs0; a.p0 = p0; a.z0 = z0; a.dp0 = dp0; a.dpdry0 = dpdry0; a.u0 = u0; a.v0 = v0;
    a.qv0 = qv0; a.ql0 = ql0; a.qi0 = qi0; a.t0 = t0; a.s0 = s0; a.tr0 = tr0; a.tke = tke; a.pblh = pblh;
    a.cush = cush;
    double** o = const_cast<double**>(outs);
    a.umf = o[0]; a.slflx = o[1]; a.qtflx = o[2]; a.flxprc1 = o[3]; a.flxsnow1 = o[4]; a.qvten = o[5];
    a.qlten = o[6]; a.qiten = o[7]; a.sten = o[8]; a.uten = o[9]; a.vten = o[10]; a.trten = o[11]; a.qrten = o[12];
    a.qsten = o[13]; a.precip = o[14]; a.snow = o[15]; a.evapc = o[16]; a.cufrc = o[17]; a.qcu = o[18];
    a.qlu = o[19]; a.qiu = o[20]; a.cbmf = o[21]; a.qc = o[22]; a.rliq = o[23]; a.cnt = o[24]; a.cnb = o[25];
    a.wu = wu;
    a.diag = diag;
    int err[2] = {0, -1};
    a.err = err;
    a.wld = ncol;
    const int KS = pver + 1;
    std::vector<char> ws(ws_bytes_per_column(pver, ncnst, sizeof(AD), sizeof(AD)) * (size_t)ncol + 64);
    a.ws = ws.data();
    std::vector<AD> adout((size_t)ad_fields(ncnst) * KS * ncol);
    std::vector<TapeCtx> tapes(ncol);
    std::vector<int> nodes((size_t)ad_fields(ncnst) * KS * ncol, -1);
    a.ad.out = adout.data();
    a.ad.tapes = tapes.data();
    AdCotangents ct{};
    for (int f = 0; f < ADF_TR; ++f) ct.f[f] = cot[f];
    ct.trten = cot[ADF_TR];
    for (int q = 0; q < ADC_COUNT; ++q) ct.scal[q] = cot[ADF_TR + 1 + q];
    AdGradients g{grad[0], grad[1], grad[2], grad[3], grad[4], grad[5], grad[6]};
    const ThermoT<AD> T{c, estbl};
    const int nin = ad_inputs(pver);
    if (g_window > 0 && g_window < kSweepBatch) return UWSHCU_ERR_ARG;   // the window holds a batch
#pragma omp parallel
    {
        const int maxp = (cap + kTapePage - 1) / kTapePage;
        std::vector<TapeNode> buf(g_paged ? (size_t)maxp * kTapePage : (size_t)cap);
        std::vector<int> pages((size_t)maxp);
        std::vector<double> adj((size_t)cap);
        TapePool pool{};
        pool.base = buf.data();
        pool.per_shard = maxp;
        pool.nshards = 1;
#pragma omp for schedule(dynamic, 4)
        for (int i = 0; i < ncol; ++i) {
            pool.next[0] = 0;
            if (g_paged) {
                ad_begin(tapes[i], tapes[i].hot, nullptr, pages.data(), &pool, 0, cap, nin);
            } else {
                ad_begin(tapes[i], tapes[i].hot, buf.data(), nullptr, nullptr, 0, cap, nin);
            }
            std::fill(adj.begin(), adj.end(), 0.0);
            const Column<AD, AD> col(a, T, i, i);
            col.run();
            col.ad_export(nodes.data());
            if (tapes[i].overflow) {
                // as on the GPU: a tape that did not fit gets no sweep (and no gradient)
            } else if (g_window > 0) {
                std::vector<double> wv((size_t)g_window + 1);
                auto win = [&](int q) -> double& { return wv[(size_t)q]; };
                ad_reverse(tapes[i], adj.data(), nodes.data(), (size_t)ncol, i, pver, ncnst, (size_t)pcols, i, ct, g,
                           win, g_window);
            } else {
                ad_reverse(tapes[i], adj.data(), nodes.data(), (size_t)ncol, i, pver, ncnst, (size_t)pcols, i, ct, g);
            }
            tape_len[i] = tapes[i].overflow ? -1 : tapes[i].hot.n;
        }
    }
    return err[0];
}
