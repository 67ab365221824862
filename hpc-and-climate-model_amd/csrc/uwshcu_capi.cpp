// uwshcu_capi.cpp -- C ABI host layer: contexts, tables, workspace, launches.
//
// Replaces the per-call allocate/copy/free sequence of the reference C entry point
// (bxcompute_uwshcu_inv.cpp:6208-6493, 6887-7043) with a persistent context:
// device constants and table, a status word and a workspace that grows on
// demand.  The level flip and the wet-bulb computation run inside the kernel.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <array>
#include <mutex>
#include <vector>

#include "uwshcu.h"
#include "uwshcu_kernels.h"

using uwshcu::FwdArgs;

struct uwshcu_ctx {
    int device = 0;
    uwshcu_consts consts{};
    double estbl[UWSHCU_ESTBL_LEN]{};
    uwshcu_consts* d_consts = nullptr;
    double* d_estbl = nullptr;
    int* d_err = nullptr;
    void* d_ws = nullptr;
    size_t ws_bytes = 0;
    size_t ws_limit = (size_t)24 << 30;
    // staging for the host-pointer entry points
    double* d_stage = nullptr;
    size_t stage_bytes = 0;
    int* d_diag = nullptr;
    size_t diag_bytes = 0;
    // mixed precision: input cush copy, branch flags, recompute list and its length
    double* d_cush_in = nullptr;
    int* d_amb = nullptr;
    int* d_list = nullptr;
    int* d_count = nullptr;
    size_t mixed_cols = 0;
    double mixed_tol = 1.e-4;
    // adjoint: tape capacity per column (0 = as many as the column needs) and a per-column
    // parameter-gradient buffer for callers that only want the sums
    long long tape_cap = 0;
    double* d_parcol = nullptr;
    size_t parcol_cols = 0;
    uwshcu::TapePool h_pool{};   // host copy of the recording pass's page-pool counters
    double tape_ratio = 0;       // previous call's mean tape over its default estimate (0: none yet)
    // columns whose second convection pass runs in a launch of its own
    int* d_defer = nullptr;
    double* d_defer_state = nullptr;   // NRESUME x defer_cols doubles: their saved scalars
    size_t defer_cols = 0;
    // chunk batching: the packed column range, its device tables and their pinned
    // host staging (reused once the event says the previous upload finished)
    double* d_packed = nullptr;
    size_t packed_bytes = 0;
    char* d_ctab = nullptr;
    size_t ctab_bytes = 0;
    char* h_ctab = nullptr;
    size_t h_ctab_bytes = 0;
    hipEvent_t ctab_event = nullptr;
    bool ctab_pending = false;
    hipStream_t stream = nullptr;
    // this device's address of the static scratch array (uwshcu_forward_device_noalloc)
    void* d_scratch = nullptr;
    std::mutex mu;
};

// Workspace of uwshcu_forward_device_noalloc: a fixed array in the code object's
// device memory, one copy per GPU the library runs on, held until the process ends.
// Host code never passes `scratch` itself to a kernel (on the host it is a stand-in,
// not a device address): uwshcu_ctx_create looks up this device's address once with
// hipGetSymbolAddress and every launch uses that.
static __device__ __align__(256) unsigned char scratch[UWSHCU_NOALLOC_SCRATCH_BYTES];

namespace {

// gffgch.F90:1-127 (Goff & Gratch 1946).  itype < 0 selects a -itype degree
// transition between water and ice below tmelt.
double goff_gratch(double t, int itype, double tmelt) {
    double tr = 0.0, eswtr = 0.0, es;
    if (itype < 0) {
        tr = std::fabs((double)itype);
        itype = 1;
    }
    const bool ice_only = (t < (tmelt - tr) && itype == 1);
    if (!ice_only) {
        const double ps = 1013.246, ts = 373.16;
        const double e1 = 11.344 * (1.0 - t / ts);
        const double e2 = -3.49149 * (ts / t - 1.0);
        const double f1 = -7.90298 * (ts / t - 1.0);
        const double f2 = 5.02808 * std::log10(ts / t);
        const double f3 = -1.3816 * (std::pow(10.0, e1) - 1.0) / 10000000.0;
        const double f4 = 8.1328 * (std::pow(10.0, e2) - 1.0) / 1000.0;
        const double f5 = std::log10(ps);
        const double f = f1 + f2 + f3 + f4 + f5;
        es = std::pow(10.0, f) * 100.0;
        eswtr = es;
        if (t >= tmelt || itype == 0) return es;
    }
    const double t0 = tmelt;
    const double term1 = 2.01889049 / (t0 / t);
    const double term2 = 3.56654 * std::log(t0 / t);
    const double term3 = 20.947031 * (t0 / t);
    es = 575.185606e10 * std::exp(-(term1 + term2 + term3));
    if (t < (tmelt - tr)) return es;
    const double weight = std::fmin((tmelt - t) / tr, 1.0);
    return weight * es + (1.0 - weight) * eswtr;
}

bool hip_ok(hipError_t e) { return e == hipSuccess; }

// Per-device address of `scratch`, looked up once per device with that device current.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxDevices] = {};

int scratch_address(int device, void** out) {
    if (device < 0 || device >= kMaxDevices) return UWSHCU_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_addr[device]) {
        void* p = nullptr;
        if (!hip_ok(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) || !p) return UWSHCU_ERR_HIP;
        g_scratch_addr[device] = p;
    }
    *out = g_scratch_addr[device];
    return UWSHCU_OK;
}

int ensure_ws(uwshcu_ctx* ctx, size_t bytes) {
    if (bytes <= ctx->ws_bytes) return UWSHCU_OK;
    if (ctx->d_ws) {
        if (!hip_ok(hipDeviceSynchronize())) return UWSHCU_ERR_HIP;
        (void)hipFree(ctx->d_ws);
        ctx->d_ws = nullptr;
        ctx->ws_bytes = 0;
    }
    if (!hip_ok(hipMalloc(&ctx->d_ws, bytes))) return UWSHCU_ERR_NOMEM;
    ctx->ws_bytes = bytes;
    return UWSHCU_OK;
}

int ensure_mixed(uwshcu_ctx* ctx, size_t ncol) {
    if (ncol <= ctx->mixed_cols && ctx->d_count) return UWSHCU_OK;
    if (ctx->d_cush_in || ctx->d_amb || ctx->d_list) {
        if (!hip_ok(hipDeviceSynchronize())) return UWSHCU_ERR_HIP;
        (void)hipFree(ctx->d_cush_in);
        (void)hipFree(ctx->d_amb);
        (void)hipFree(ctx->d_list);
        ctx->d_cush_in = nullptr;
        ctx->d_amb = ctx->d_list = nullptr;
        ctx->mixed_cols = 0;
    }
    if (!ctx->d_count && !hip_ok(hipMalloc((void**)&ctx->d_count, sizeof(int)))) return UWSHCU_ERR_NOMEM;
    if (!hip_ok(hipMalloc((void**)&ctx->d_cush_in, ncol * sizeof(double))) ||
        !hip_ok(hipMalloc((void**)&ctx->d_amb, ncol * sizeof(int))) ||
        !hip_ok(hipMalloc((void**)&ctx->d_list, ncol * sizeof(int))))
        return UWSHCU_ERR_NOMEM;
    ctx->mixed_cols = ncol;
    return UWSHCU_OK;
}

// Columns per launch: as many as the workspace limit allows, a multiple of 64, at
// least 64 and at most ncol.
long long batch_columns(const uwshcu_ctx* ctx, int ncol, size_t per_col) {
    long long nb = (long long)(ctx->ws_limit / per_col);
    if (nb > ncol) nb = ncol;
    nb = (nb / 64) * 64;
    if (nb < 64) nb = 64;
    return nb;
}

struct Chunk {
    const double *ps0, *zs0, *p0, *z0, *dp0, *u0, *v0, *qv0, *ql0, *qi0, *t0, *s0, *tr0, *tke, *pblh, *dpdry0;
    double* cush;
    double *umf, *slflx, *qtflx, *flxprc1, *flxsnow1, *qvten, *qlten, *qiten, *sten, *uten, *vten, *trten;
    double *qrten, *qsten, *precip, *snow, *evapc, *cufrc, *qcu, *qlu, *qiu, *cbmf, *qc, *rliq, *cnt, *cnb, *wu;
};

// The modified scheme's options and extra inputs (device pointers).
struct ZwInputs {
    const uwshcu_zw_params* par;
    const double *omega, *tpert, *qpert;
};

// fixed_ws (nullable): run in this caller-given workspace of fixed_bytes instead of the
// context's, allocating nothing (FP64 only, one batch, no deferred-column lists).
int run_device(uwshcu_ctx* ctx, int precision, const uwshcu_params* par, int pcols, int ncol, int pver, int ncnst,
               double dt, const Chunk& ch, int* diag, hipStream_t stream, const ZwInputs* zw = nullptr,
               void* fixed_ws = nullptr, size_t fixed_bytes = 0) {
    if (!ctx || pcols < 1 || ncol < 0 || ncol > pcols || pver < 3 || ncnst < 0 || ncnst > UWSHCU_NCNST_MAX)
        return UWSHCU_ERR_ARG;
    if (precision != UWSHCU_FP64 && precision != UWSHCU_MIXED && precision != UWSHCU_MIXED_UNCHECKED &&
        precision != UWSHCU_MIXED_FAST_UNCHECKED)
        return UWSHCU_ERR_ARG;
    if (zw && (precision != UWSHCU_FP64 || !zw->par || !zw->omega || !zw->tpert || !zw->qpert)) return UWSHCU_ERR_ARG;
    if (fixed_ws && (precision != UWSHCU_FP64 || zw)) return UWSHCU_ERR_ARG;
    const bool fp32 = precision == UWSHCU_MIXED_UNCHECKED;   // the FP32-plume pass
    const bool fast = precision == UWSHCU_MIXED || precision == UWSHCU_MIXED_FAST_UNCHECKED;   // fast FP64 pass
    const bool recompute = precision == UWSHCU_MIXED;
    const size_t per_col = uwshcu_workspace_bytes_per_column(pver, ncnst, precision);
    const size_t per_col64 = uwshcu_workspace_bytes_per_column(pver, ncnst, UWSHCU_FP64);
    // batches of columns that share one workspace; a fixed workspace holds the whole
    // chunk (its leading dimension rounded up to a wavefront) or the call is refused
    const long long nb = fixed_ws ? ((long long)ncol + 63) / 64 * 64 : batch_columns(ctx, ncol, per_col);
    if (fixed_ws && per_col * (size_t)nb > fixed_bytes) return UWSHCU_ERR_SCRATCH;
    if (ncol == 0) return UWSHCU_OK;
    if (!hip_ok(hipSetDevice(ctx->device))) return UWSHCU_ERR_HIP;
    const long long nb64 = recompute ? batch_columns(ctx, ncol, per_col64) : 0;
    size_t need = per_col * (size_t)nb;
    if (recompute && per_col64 * (size_t)nb64 > need) need = per_col64 * (size_t)nb64;
    int rc = fixed_ws ? UWSHCU_OK : ensure_ws(ctx, need);
    if (rc) return rc;
    if (recompute && (rc = ensure_mixed(ctx, (size_t)ncol))) return rc;
    // defer list (nb entries and a counter) and the deferred columns' scalars, for
    // large launches
    if (!fixed_ws && nb >= 2 * 1024 * 64 && (size_t)nb + 1 > ctx->defer_cols) {
        if (ctx->d_defer) {
            if (!hip_ok(hipDeviceSynchronize())) return UWSHCU_ERR_HIP;
            (void)hipFree(ctx->d_defer);
            (void)hipFree(ctx->d_defer_state);
            ctx->d_defer = nullptr;
            ctx->d_defer_state = nullptr;
            ctx->defer_cols = 0;
        }
        if (!hip_ok(hipMalloc((void**)&ctx->d_defer, (2 + uwshcu::defer_list_ints((size_t)nb)) * sizeof(int))))
            return UWSHCU_ERR_NOMEM;
        if (!hip_ok(hipMalloc((void**)&ctx->d_defer_state, (size_t)uwshcu::NRESUME * ((size_t)nb + 1) * sizeof(double)))) {
            (void)hipFree(ctx->d_defer);
            ctx->d_defer = nullptr;
            return UWSHCU_ERR_NOMEM;
        }
        ctx->defer_cols = (size_t)nb + 1;
    }

    FwdArgs a{};
    a.ncol = ncol;
    a.mkx = pver;
    a.ncnst = ncnst;
    a.ldc = pcols;
    a.dt = dt;
    uwshcu_default_params(&a.par);
    if (par) a.par = *par;
    a.ps0 = ch.ps0; a.zs0 = ch.zs0; a.p0 = ch.p0; a.z0 = ch.z0; a.dp0 = ch.dp0; a.dpdry0 = ch.dpdry0;
    a.u0 = ch.u0; a.v0 = ch.v0; a.qv0 = ch.qv0; a.ql0 = ch.ql0; a.qi0 = ch.qi0; a.t0 = ch.t0; a.s0 = ch.s0;
    a.tr0 = ch.tr0; a.tke = ch.tke; a.pblh = ch.pblh; a.cush = ch.cush;
    a.umf = ch.umf; a.slflx = ch.slflx; a.qtflx = ch.qtflx; a.flxprc1 = ch.flxprc1; a.flxsnow1 = ch.flxsnow1;
    a.qvten = ch.qvten; a.qlten = ch.qlten; a.qiten = ch.qiten; a.sten = ch.sten; a.uten = ch.uten;
    a.vten = ch.vten; a.trten = ch.trten; a.qrten = ch.qrten; a.qsten = ch.qsten; a.precip = ch.precip;
    a.snow = ch.snow; a.evapc = ch.evapc; a.cufrc = ch.cufrc; a.qcu = ch.qcu; a.qlu = ch.qlu; a.qiu = ch.qiu;
    a.cbmf = ch.cbmf; a.qc = ch.qc; a.rliq = ch.rliq; a.cnt = ch.cnt; a.cnb = ch.cnb; a.wu = ch.wu;
    a.diag = diag;
    a.err = ctx->d_err;
    a.ws = fixed_ws ? fixed_ws : ctx->d_ws;
    if (zw) {
        a.zw = *zw->par;
        a.omega = zw->omega;
        a.tpert = zw->tpert;
        a.qpert = zw->qpert;
    }
    if (!fixed_ws && ctx->d_defer && (size_t)nb + 1 <= ctx->defer_cols) {
        a.defer_list = ctx->d_defer + 2;   // two counts, then the lists and marks (see launch_column_pipeline)
        a.defer_count = ctx->d_defer;
        a.defer_state = ctx->d_defer_state;   // [field][slot], a.wld = nb apart
    }
    if (recompute) {
        // the FP32 pass overwrites cush; the FP64 recompute reads the input value
        if (!hip_ok(hipMemcpyAsync(ctx->d_cush_in, ch.cush, (size_t)ncol * sizeof(double), hipMemcpyDeviceToDevice,
                                   stream)))
            return UWSHCU_ERR_HIP;
        a.amb_flag = ctx->d_amb;
    }
    a.wld = nb;
    for (long long c0 = 0; c0 < ncol; c0 += nb) {
        const int n = (int)((ncol - c0) < nb ? (ncol - c0) : nb);
        a.col0 = (int)c0;
        hipError_t e = zw     ? uwshcu::launch_forward_zw(a, n, ctx->d_consts, ctx->d_estbl, stream)
                       : fp32 ? uwshcu::launch_forward_mixed(a, n, ctx->d_consts, ctx->d_estbl, stream)
                       : fast ? uwshcu::launch_forward_fast(a, n, ctx->d_consts, ctx->d_estbl, stream)
                              : uwshcu::launch_forward_fp64(a, n, ctx->d_consts, ctx->d_estbl, stream);
        if (!hip_ok(e)) return UWSHCU_ERR_HIP;
    }
    if (!recompute) return UWSHCU_OK;
    // conservation/branch check, then FP64 over the listed columns; the list length
    // stays on the device (launches past it exit at once)
    a.col0 = 0;
    if (!hip_ok(uwshcu::launch_check_columns(a, ctx->d_consts, ctx->mixed_tol, ctx->d_list, ctx->d_count, stream)))
        return UWSHCU_ERR_HIP;
    a.amb_flag = nullptr;
    a.cush_in = ctx->d_cush_in;
    a.colmap = ctx->d_list;
    a.colcount = ctx->d_count;
    a.sparse_map = 1;
    a.wld = nb64;
    for (long long c0 = 0; c0 < ncol; c0 += nb64) {
        const int n = (int)((ncol - c0) < nb64 ? (ncol - c0) : nb64);
        a.col0 = (int)c0;
        if (!hip_ok(uwshcu::launch_forward_fp64(a, n, ctx->d_consts, ctx->d_estbl, stream))) return UWSHCU_ERR_HIP;
    }
    return UWSHCU_OK;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

int run_adjoint(uwshcu_ctx* ctx, const uwshcu_params* par, int pcols, int ncol, int pver, int ncnst, double dt,
                const Chunk& ch, int* diag, const uwshcu_cotangents* cot, const uwshcu_gradients* grad,
                hipStream_t stream) {
    using namespace uwshcu;
    if (!ctx || !cot || !grad || pcols < 1 || ncol < 0 || ncol > pcols || pver < 3 || ncnst < 0 ||
        ncnst > UWSHCU_NCNST_MAX)
        return UWSHCU_ERR_ARG;
    if (ncol == 0) return UWSHCU_OK;
    if (!hip_ok(hipSetDevice(ctx->device))) return UWSHCU_ERR_HIP;
    const int KS = pver + 1;
    // tape capacity per column: the caller's, else four times the default bound
    // (2048 + 512 ncnst)(pver + 1), about ten times the longest tapes measured
    long long capn = ctx->tape_cap > 0 ? ctx->tape_cap : 4LL * (2048 + 512LL * ncnst) * KS;
    if (capn > 0x7fffffffLL - kTapePage) capn = 0x7fffffffLL - kTapePage;
    const int cap = (int)capn;
    const int maxp = (cap + kTapePage - 1) / kTapePage;
    // per-column fixed pieces: AD workspace, AD outputs, output nodes, tape header, page
    // table; per pool page: its nodes and as many adjoint entries
    const size_t b_ws = ws_bytes_per_column(pver, ncnst, sizeof(AD), sizeof(AD));
    const size_t b_out = (size_t)ad_fields(ncnst) * KS * sizeof(AD);
    const size_t b_nodes = (size_t)ad_fields(ncnst) * KS * sizeof(int);
    const size_t b_fix = b_ws + b_out + b_nodes + sizeof(TapeCtx) + (size_t)maxp * sizeof(int);
    const size_t b_page = (size_t)kTapePage * (sizeof(TapeNode) + sizeof(double));
    auto layout = [&](long long nb, size_t npages, size_t* o) {   // o: out, nodes, tapes, pages, pool, tape, adj, end
        o[0] = align256(b_ws * nb);
        o[1] = o[0] + align256(b_out * nb);
        o[2] = o[1] + align256(b_nodes * nb);
        o[3] = o[2] + align256(sizeof(TapeCtx) * nb);
        o[4] = o[3] + align256((size_t)maxp * sizeof(int) * nb);
        o[5] = o[4] + align256(sizeof(TapePool));
        o[6] = o[5] + align256(sizeof(TapeNode) * kTapePage * npages);
        o[7] = o[6] + align256(sizeof(double) * kTapePage * npages);
    };
    double* parcol = grad->par_col;
    if (!parcol && grad->par_sum) {
        if ((size_t)pcols > ctx->parcol_cols) {
            if (ctx->d_parcol) {
                if (!hip_ok(hipDeviceSynchronize())) return UWSHCU_ERR_HIP;
                (void)hipFree(ctx->d_parcol);
                ctx->d_parcol = nullptr;
                ctx->parcol_cols = 0;
            }
            if (!hip_ok(hipMalloc((void**)&ctx->d_parcol, 3 * (size_t)pcols * sizeof(double)))) return UWSHCU_ERR_NOMEM;
            ctx->parcol_cols = (size_t)pcols;
        }
        parcol = ctx->d_parcol;
    }
    AdCotangents ct;
    static_assert(sizeof(AdCotangents) == sizeof(uwshcu_cotangents), "cotangent layouts differ");
    std::memcpy(&ct, cot, sizeof(ct));
    const AdGradients g{grad->t0, grad->qv0, grad->u0, grad->v0, grad->tke, grad->pblh, parcol};

    FwdArgs a{};
    a.ncol = ncol; a.mkx = pver; a.ncnst = ncnst; a.ldc = pcols; a.dt = dt;
    uwshcu_default_params(&a.par);
    if (par) a.par = *par;
    a.ps0 = ch.ps0; a.zs0 = ch.zs0; a.p0 = ch.p0; a.z0 = ch.z0; a.dp0 = ch.dp0; a.dpdry0 = ch.dpdry0;
    a.u0 = ch.u0; a.v0 = ch.v0; a.qv0 = ch.qv0; a.ql0 = ch.ql0; a.qi0 = ch.qi0; a.t0 = ch.t0; a.s0 = ch.s0;
    a.tr0 = ch.tr0; a.tke = ch.tke; a.pblh = ch.pblh; a.cush = ch.cush;
    a.umf = ch.umf; a.slflx = ch.slflx; a.qtflx = ch.qtflx; a.flxprc1 = ch.flxprc1; a.flxsnow1 = ch.flxsnow1;
    a.qvten = ch.qvten; a.qlten = ch.qlten; a.qiten = ch.qiten; a.sten = ch.sten; a.uten = ch.uten;
    a.vten = ch.vten; a.trten = ch.trten; a.qrten = ch.qrten; a.qsten = ch.qsten; a.precip = ch.precip;
    a.snow = ch.snow; a.evapc = ch.evapc; a.cufrc = ch.cufrc; a.qcu = ch.qcu; a.qlu = ch.qlu; a.qiu = ch.qiu;
    a.cbmf = ch.cbmf; a.qc = ch.qc; a.rliq = ch.rliq; a.cnt = ch.cnt; a.cnb = ch.cnb; a.wu = ch.wu;
    a.diag = diag;
    a.err = ctx->d_err;
    // cush is read and written by the column: a launch recorded again (below) must read
    // the caller's input, so every launch reads a copy of it
    if (int e = ensure_mixed(ctx, (size_t)ncol)) return e;
    if (!hip_ok(hipMemcpyAsync(ctx->d_cush_in, ch.cush, (size_t)ncol * sizeof(double), hipMemcpyDeviceToDevice,
                               stream)))
        return UWSHCU_ERR_HIP;
    a.cush_in = ctx->d_cush_in;

    // Columns per launch and pool size: the pool is sized for a per-column estimate of
    // the tape length, plus half a page per column for their last pages.  The pool is
    // split into shards with a page counter each, one per four workgroups of 64 columns.  The estimate
    // is (600 + 150 ncnst)(pver + 1) nodes (1.7 times the mean tape of a 35-level chunk
    // that is 72 % convecting), scaled after a context's first call by 1.25 times the
    // ratio of the previous call's mean tape to its own such estimate (so that a call
    // with other levels or constituents scales it).  A launch whose columns need more finds
    // the pool empty; it is then recorded again in parts that fit, from the lengths its
    // columns counted.
    const double est0 = (double)(600 + 150 * ncnst) * KS;
    const double est = ctx->tape_ratio > 0 ? ctx->tape_ratio * est0 : est0;
    const double pages_per_col = est / kTapePage + 0.5;
    double recorded = 0;   // tape nodes of the columns recorded
    long long nb = (long long)((double)ctx->ws_limit / ((double)b_fix + pages_per_col * (double)b_page + 512.0));
    if (nb > ncol) nb = ncol;
    if (nb < 1) return UWSHCU_ERR_NOMEM;
    // pages the workspace limit leaves for a launch of `cols` columns
    auto room = [&](long long cols) {
        size_t o[8];
        layout(cols, 0, o);
        return ctx->ws_limit > o[7] + 1024 ? (ctx->ws_limit - o[7] - 1024) / b_page : (size_t)0;
    };
    auto pages_for = [&](long long cols) {
        const size_t np = (size_t)std::ceil((double)cols * pages_per_col) + 1;
        return np < room(cols) ? np : room(cols);
    };
    size_t o[8];
    layout(nb, pages_for(nb), o);
    int rc = ensure_ws(ctx, o[7]);
    if (rc) return rc;
    char* base = (char*)ctx->d_ws;
    TapePool* const hp = &ctx->h_pool;

    // record [c0, c0 + n) with a pool of npages pages; *retry receives the number of
    // columns that found the pool empty, *nadj the adjoint entries taken
    auto record = [&](long long c0, long long n, size_t npages, int nshards, int* retry, long long* nadj) -> int {
        size_t oo[8];
        layout(n, npages, oo);
        if (oo[7] > ctx->ws_bytes) {   // a part of a launch that ran out of pages (after a synchronise)
            if (int e = ensure_ws(ctx, oo[7])) return e;
            base = (char*)ctx->d_ws;
        }
        a.ws = base;
        a.wld = n;
        a.col0 = (int)c0;
        a.ad.out = (AD*)(base + oo[0]);
        a.ad.tapes = (TapeCtx*)(base + oo[2]);
        std::memset(hp, 0, sizeof(TapePool));
        hp->base = (TapeNode*)(base + oo[5]);
        hp->nshards = nshards;
        hp->per_shard = (int)(npages / (size_t)nshards);
        TapePool* dpool = (TapePool*)(base + oo[4]);
        if (!hip_ok(hipMemcpyAsync(dpool, hp, sizeof(TapePool), hipMemcpyHostToDevice, stream)) ||
            !hip_ok(launch_adjoint_record(a, (int)n, ctx->d_consts, ctx->d_estbl, dpool, (int*)(base + oo[3]), maxp,
                                          cap, (int*)(base + oo[1]), stream)) ||
            !hip_ok(hipMemcpyAsync(hp, dpool, offsetof(TapePool, next), hipMemcpyDeviceToHost, stream)) ||
            !hip_ok(hipStreamSynchronize(stream)))
            return UWSHCU_ERR_HIP;
        *retry = hp->retry;
        *nadj = (long long)hp->adj_next;
        return UWSHCU_OK;
    };
    auto sweep = [&](long long n, size_t npages, long long nadj) -> int {
        size_t oo[8];
        layout(n, npages, oo);
        if (!hip_ok(launch_adjoint_sweep(a, (int)n, (double*)(base + oo[6]), nadj, (const int*)(base + oo[1]), ct, g,
                                         stream)))
            return UWSHCU_ERR_HIP;
        return UWSHCU_OK;
    };
    for (long long c0 = 0; c0 < ncol; c0 += nb) {
        const long long n = (ncol - c0) < nb ? (ncol - c0) : nb;
        const size_t npages = pages_for(n);
        // a page counter per four workgroups (at most kTapeShards)
        const long long groups = (n + 63) / 64;
        const int nshards = (int)std::max(1LL, std::min<long long>(kTapeShards, groups / 4));
        int retry = 0;
        long long nadj = 0;
        if ((rc = record(c0, n, npages, nshards, &retry, &nadj))) return rc;
        if (retry == 0) {
            if ((rc = sweep(n, npages, nadj))) return rc;
            recorded += (double)nadj;
            continue;
        }
        // the pool ran out: every column counted its length; record again in parts
        std::vector<TapeCtx> hdr((size_t)n);
        layout(n, npages, o);
        if (!hip_ok(hipMemcpyAsync(hdr.data(), base + o[2], (size_t)n * sizeof(TapeCtx), hipMemcpyDeviceToHost,
                                   stream)) ||
            !hip_ok(hipStreamSynchronize(stream)))
            return UWSHCU_ERR_HIP;
        long long s0 = 0;
        while (s0 < n) {
            size_t need = 0;
            long long s1 = s0;
            while (s1 < n) {
                const size_t pc = ((size_t)(hdr[s1].hot.n < cap ? hdr[s1].hot.n : cap) + kTapePage - 1) / kTapePage;
                if (s1 > s0 && need + pc > room(s1 + 1 - s0)) break;
                need += pc;
                ++s1;
            }
            if (need > room(s1 - s0)) return UWSHCU_ERR_NOMEM;   // one column needs more than the limit holds
            if ((rc = record(c0 + s0, s1 - s0, need, 1, &retry, &nadj))) return rc;   // one counter: exact fit
            if (retry) return UWSHCU_ERR_HIP;                    // the same columns counted a different length
            if ((rc = sweep(s1 - s0, need, nadj))) return rc;
            recorded += (double)nadj;
            s0 = s1;
        }
    }
    ctx->tape_ratio = 1.25 * recorded / ((double)ncol * est0);
    if (grad->par_sum && !hip_ok(launch_param_sum(parcol, pcols, ncol, grad->par_sum, stream))) return UWSHCU_ERR_HIP;
    return UWSHCU_OK;
}

std::mutex g_dropin_mu;
uwshcu_ctx* g_dropin_ctx = nullptr;
uwshcu_params g_dropin_par{10.0, 0.1, 14.0};
bool g_dropin_par_set = false;   // rpen has no default: init_uwshcu stops the model if it is unset
int g_dropin_status = 0, g_dropin_badcol = -1;

}  // namespace

extern "C" {

int uwshcu_default_consts(uwshcu_consts* c) {
    if (!c) return UWSHCU_ERR_ARG;
    std::memset(c, 0, sizeof(*c));
    // CESM shr_const_mod values, combined in the order shr_const_mod combines them
    const double boltz = 1.38065e-23, avogad = 6.02214e26;
    const double rgas = avogad * boltz;
    const double mwdair = 28.966, mwwv = 18.016;
    const double rdair = rgas / mwdair, rwv = rgas / mwwv;
    c->xlv = 2.501e6;
    c->xlf = 3.337e5;
    c->xls = c->xlv + c->xlf;
    c->cp = 1.00464e3;
    c->zvir = rwv / rdair - 1.0;
    c->r = rdair;
    c->g = 9.80616;
    c->ep2 = mwwv / mwdair;
    c->p00 = 1.e5;
    c->rovcp = c->r / c->cp;
    c->tmin = 173.16;      // esinti.F90:44
    c->tmax = 375.16;      // esinti.F90:46
    c->ttrice = 20.0;      // esinti.F90:47
    c->icephs = 1;         // esinti.F90:48
    c->epsqs = mwwv / mwdair;
    c->rgasv = rwv;
    c->hlatf = 3.337e5;
    c->hlatv = 2.501e6;
    c->cpws = 1.00464e3;
    c->tmelt = 273.15;
    c->ixnumliq = 4;
    c->ixnumice = 5;
    // degree-5 fit of es(water)-es(ice), wv_saturation.F90:178-182
    c->pcf[0] = 5.04469588506e-01;
    c->pcf[1] = -5.47288442819e+00;
    c->pcf[2] = -3.67471858735e-01;
    c->pcf[3] = -8.95963532403e-03;
    c->pcf[4] = -7.78053686625e-05;
    for (int m = 0; m <= UWSHCU_NCNST_MAX; ++m) {
        c->qmin[m] = 0.0;
        c->wet[m] = 1;
    }
    c->qmin[1] = 1.e-12;   // Q (initindx.F90:75); CLDLIQ/CLDICE/NUMLIQ/NUMICE 0 (microp_driver.F90:112-120)
    return UWSHCU_OK;
}

int uwshcu_build_estbl(const uwshcu_consts* c, double* estbl) {
    if (!c || !estbl) return UWSHCU_ERR_ARG;
    // wv_saturation.F90:141-169
    const int lentbl = (int)(c->tmax - c->tmin + 2.000001);
    if (lentbl > UWSHCU_ESTBL_LEN) return UWSHCU_ERR_ARG;
    int itype;
    if (c->icephs) itype = (c->ttrice != 0.0) ? (int)(-c->ttrice) : 1;
    else itype = 0;
    double t = c->tmin - 1.0;
    for (int n = 0; n < lentbl; ++n) {
        t = t + 1.0;
        estbl[n] = goff_gratch(t, itype, c->tmelt);
    }
    for (int n = lentbl; n < UWSHCU_ESTBL_LEN; ++n) estbl[n] = -99999.0;
    return UWSHCU_OK;
}

int uwshcu_default_params(uwshcu_params* p) {
    if (!p) return UWSHCU_ERR_ARG;
    p->rpen = 10.0;
    p->rle = 0.1;
    p->rkm = 14.0;
    return UWSHCU_OK;
}

int uwshcu_zw_default_params(uwshcu_params* p, uwshcu_zw_params* z) {
    if (!p && !z) return UWSHCU_ERR_ARG;
    if (p) {
        p->rpen = 15.0;   // uwshcu_rpen (zwshcu.F90:44)
        p->rle = 0.1;
        p->rkm = 10.0;    // zmh_rkm (zwshcu.F90:43, 1251)
    }
    if (z) {
        *z = uwshcu_zw_params{};
        z->lf = 1;
        z->fwcrit = 1.0;
        z->ent0 = (double)0.015f;   // default-real literals in the reference
        z->ufrcd = 2.0;
        z->ufrc0 = 2.0;
        z->ufrcz = 2.0;
        z->tpertmax = 2.0;
        z->qpertmax = (double)1.0e-3f;
    }
    return UWSHCU_OK;
}

size_t uwshcu_workspace_bytes_per_column(int pver, int ncnst, int precision) {
    const size_t elem = (precision == UWSHCU_MIXED_UNCHECKED) ? sizeof(float) : sizeof(double);
    return uwshcu::ws_bytes_per_column(pver, ncnst, elem);
}

int uwshcu_ctx_create(const uwshcu_consts* c, const double* estbl, int device, uwshcu_ctx** out) {
    if (!out) return UWSHCU_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (!hip_ok(hipGetDeviceCount(&ndev)) || ndev < 1) return UWSHCU_ERR_NO_DEVICE;
    if (device < 0 && !hip_ok(hipGetDevice(&device))) return UWSHCU_ERR_HIP;
    if (device >= ndev) return UWSHCU_ERR_ARG;
    uwshcu_ctx* ctx = new uwshcu_ctx();
    ctx->device = device;
    if (c) ctx->consts = *c;
    else uwshcu_default_consts(&ctx->consts);
    if (estbl) std::memcpy(ctx->estbl, estbl, sizeof(ctx->estbl));
    else uwshcu_build_estbl(&ctx->consts, ctx->estbl);
    bool ok = hip_ok(hipSetDevice(device)) && hip_ok(hipMalloc((void**)&ctx->d_consts, sizeof(uwshcu_consts))) &&
              hip_ok(hipMalloc((void**)&ctx->d_estbl, sizeof(ctx->estbl))) &&
              hip_ok(hipMalloc((void**)&ctx->d_err, 2 * sizeof(int))) &&
              hip_ok(hipMemcpy(ctx->d_consts, &ctx->consts, sizeof(uwshcu_consts), hipMemcpyHostToDevice)) &&
              hip_ok(hipMemcpy(ctx->d_estbl, ctx->estbl, sizeof(ctx->estbl), hipMemcpyHostToDevice)) &&
              hip_ok(uwshcu::launch_clear_status(ctx->d_err, nullptr)) && hip_ok(hipStreamCreate(&ctx->stream)) &&
              scratch_address(device, &ctx->d_scratch) == UWSHCU_OK && hip_ok(hipDeviceSynchronize());
    if (!ok) {
        uwshcu_ctx_destroy(ctx);
        return UWSHCU_ERR_HIP;
    }
    *out = ctx;
    return UWSHCU_OK;
}

int uwshcu_ctx_destroy(uwshcu_ctx* ctx) {
    if (!ctx) return UWSHCU_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    if (ctx->d_consts) (void)hipFree(ctx->d_consts);
    if (ctx->d_estbl) (void)hipFree(ctx->d_estbl);
    if (ctx->d_err) (void)hipFree(ctx->d_err);
    if (ctx->d_ws) (void)hipFree(ctx->d_ws);
    if (ctx->d_stage) (void)hipFree(ctx->d_stage);
    if (ctx->d_diag) (void)hipFree(ctx->d_diag);
    if (ctx->d_cush_in) (void)hipFree(ctx->d_cush_in);
    if (ctx->d_amb) (void)hipFree(ctx->d_amb);
    if (ctx->d_list) (void)hipFree(ctx->d_list);
    if (ctx->d_count) (void)hipFree(ctx->d_count);
    if (ctx->d_parcol) (void)hipFree(ctx->d_parcol);
    if (ctx->d_defer) (void)hipFree(ctx->d_defer);
    if (ctx->d_defer_state) (void)hipFree(ctx->d_defer_state);
    if (ctx->d_packed) (void)hipFree(ctx->d_packed);
    if (ctx->d_ctab) (void)hipFree(ctx->d_ctab);
    if (ctx->h_ctab) (void)hipHostFree(ctx->h_ctab);
    if (ctx->ctab_event) (void)hipEventDestroy(ctx->ctab_event);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return UWSHCU_OK;
}

int uwshcu_ctx_set_mixed_tolerance(uwshcu_ctx* ctx, double tol) {
    if (!ctx || !(tol > 0.0)) return UWSHCU_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->mixed_tol = tol;
    return UWSHCU_OK;
}

int uwshcu_mixed_recomputed(uwshcu_ctx* ctx, void* stream, int* count) {
    if (!ctx || !count) return UWSHCU_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    *count = 0;
    if (!ctx->d_count) return UWSHCU_OK;
    if (!hip_ok(hipSetDevice(ctx->device))) return UWSHCU_ERR_HIP;
    if (!hip_ok(hipMemcpyAsync(count, ctx->d_count, sizeof(int), hipMemcpyDeviceToHost, (hipStream_t)stream)) ||
        !hip_ok(hipStreamSynchronize((hipStream_t)stream)))
        return UWSHCU_ERR_HIP;
    return UWSHCU_OK;
}

int uwshcu_ctx_set_tape_capacity(uwshcu_ctx* ctx, long long nodes_per_column) {
    if (!ctx || nodes_per_column < 0 || nodes_per_column > (1LL << 30)) return UWSHCU_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->tape_cap = nodes_per_column;
    return UWSHCU_OK;
}

int uwshcu_ctx_set_workspace_limit(uwshcu_ctx* ctx, size_t bytes) {
    if (!ctx || bytes == 0) return UWSHCU_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->ws_limit = bytes;
    return UWSHCU_OK;
}

int uwshcu_forward_device(uwshcu_ctx* ctx, int precision, const uwshcu_params* par, int pcols, int ncol, int pver,
                          int ncnst, double dt, const double* ps0, const double* zs0, const double* p0,
                          const double* z0, const double* dp0, const double* u0, const double* v0,
                          const double* qv0, const double* ql0, const double* qi0, const double* t0,
                          const double* s0, const double* tr0, const double* tke, const double* cldfrc,
                          const double* concldfrc, const double* pblh, double* cush, double* umf, double* slflx,
                          double* qtflx, double* flxprc1, double* flxsnow1, double* qvten, double* qlten,
                          double* qiten, double* sten, double* uten, double* vten, double* trten, double* qrten,
                          double* qsten, double* precip, double* snow, double* evapc, double* cufrc, double* qcu,
                          double* qlu, double* qiu, double* cbmf, double* qc, double* rliq, double* cnt,
                          double* cnb, const double* dpdry0, double* wu, int* diag, void* stream) {
    (void)cldfrc;
    (void)concldfrc;
    if (!ctx) return UWSHCU_ERR_ARG;
    const void* req[] = {ps0, zs0, p0, z0, dp0, u0, v0, qv0, ql0, qi0, t0, s0, tke, pblh, cush, umf, slflx,
                         qtflx, flxprc1, flxsnow1, qvten, qlten, qiten, sten, uten, vten, qrten, qsten, precip,
                         snow, evapc, cufrc, qcu, qlu, qiu, cbmf, qc, rliq, cnt, cnb, dpdry0, wu};
    for (const void* p : req)
        if (!p) return UWSHCU_ERR_ARG;
    if (ncnst > 0 && (!tr0 || !trten)) return UWSHCU_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    Chunk ch{ps0, zs0, p0, z0, dp0, u0, v0, qv0, ql0, qi0, t0, s0, tr0, tke, pblh, dpdry0, cush,
             umf, slflx, qtflx, flxprc1, flxsnow1, qvten, qlten, qiten, sten, uten, vten, trten,
             qrten, qsten, precip, snow, evapc, cufrc, qcu, qlu, qiu, cbmf, qc, rliq, cnt, cnb, wu};
    return run_device(ctx, precision, par, pcols, ncol, pver, ncnst, dt, ch, diag, (hipStream_t)stream);
}

int uwshcu_forward_device_noalloc(uwshcu_ctx* ctx, const uwshcu_params* par, int pcols, int ncol, int pver,
                                  int ncnst, double dt, const double* ps0, const double* zs0, const double* p0,
                                  const double* z0, const double* dp0, const double* u0, const double* v0,
                                  const double* qv0, const double* ql0, const double* qi0, const double* t0,
                                  const double* s0, const double* tr0, const double* tke, const double* cldfrc,
                                  const double* concldfrc, const double* pblh, double* cush, double* umf,
                                  double* slflx, double* qtflx, double* flxprc1, double* flxsnow1, double* qvten,
                                  double* qlten, double* qiten, double* sten, double* uten, double* vten,
                                  double* trten, double* qrten, double* qsten, double* precip, double* snow,
                                  double* evapc, double* cufrc, double* qcu, double* qlu, double* qiu, double* cbmf,
                                  double* qc, double* rliq, double* cnt, double* cnb, const double* dpdry0,
                                  double* wu, int* diag, void* stream) {
    (void)cldfrc;
    (void)concldfrc;
    if (!ctx || !ctx->d_scratch) return UWSHCU_ERR_ARG;
    const void* req[] = {ps0, zs0, p0, z0, dp0, u0, v0, qv0, ql0, qi0, t0, s0, tke, pblh, cush, umf, slflx,
                         qtflx, flxprc1, flxsnow1, qvten, qlten, qiten, sten, uten, vten, qrten, qsten, precip,
                         snow, evapc, cufrc, qcu, qlu, qiu, cbmf, qc, rliq, cnt, cnb, dpdry0, wu};
    for (const void* p : req)
        if (!p) return UWSHCU_ERR_ARG;
    if (ncnst > 0 && (!tr0 || !trten)) return UWSHCU_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    Chunk ch{ps0, zs0, p0, z0, dp0, u0, v0, qv0, ql0, qi0, t0, s0, tr0, tke, pblh, dpdry0, cush,
             umf, slflx, qtflx, flxprc1, flxsnow1, qvten, qlten, qiten, sten, uten, vten, trten,
             qrten, qsten, precip, snow, evapc, cufrc, qcu, qlu, qiu, cbmf, qc, rliq, cnt, cnb, wu};
    return run_device(ctx, UWSHCU_FP64, par, pcols, ncol, pver, ncnst, dt, ch, diag, (hipStream_t)stream, nullptr,
                      ctx->d_scratch, UWSHCU_NOALLOC_SCRATCH_BYTES);
}

int uwshcu_zw_forward_device(uwshcu_ctx* ctx, const uwshcu_params* par, const uwshcu_zw_params* zw, int pcols,
                             int ncol, int pver, int ncnst, double dt, const double* ps0, const double* zs0,
                             const double* p0, const double* z0, const double* dp0, const double* u0,
                             const double* v0, const double* qv0, const double* ql0, const double* qi0,
                             const double* t0, const double* s0, const double* tr0, const double* tke,
                             const double* cldfrc, const double* concldfrc, const double* pblh, double* cush,
                             const double* omega, const double* tpert, const double* qpert, double* umf,
                             double* slflx, double* qtflx, double* flxprc1, double* flxsnow1, double* qvten,
                             double* qlten, double* qiten, double* sten, double* uten, double* vten, double* trten,
                             double* qrten, double* qsten, double* precip, double* snow, double* evapc,
                             double* cufrc, double* qcu, double* qlu, double* qiu, double* cbmf, double* qc,
                             double* rliq, double* cnt, double* cnb, const double* dpdry0, double* wu, int* diag,
                             void* stream) {
    (void)cldfrc;
    (void)concldfrc;
    if (!ctx || !zw) return UWSHCU_ERR_ARG;
    const void* req[] = {ps0, zs0, p0, z0, dp0, u0, v0, qv0, ql0, qi0, t0, s0, tke, pblh, cush, omega, tpert,
                         qpert, umf, slflx, qtflx, flxprc1, flxsnow1, qvten, qlten, qiten, sten, uten, vten, qrten,
                         qsten, precip, snow, evapc, cufrc, qcu, qlu, qiu, cbmf, qc, rliq, cnt, cnb, dpdry0, wu};
    for (const void* p : req)
        if (!p) return UWSHCU_ERR_ARG;
    if (ncnst > 0 && (!tr0 || !trten)) return UWSHCU_ERR_ARG;
    uwshcu_params pz;
    uwshcu_zw_default_params(&pz, nullptr);
    std::lock_guard<std::mutex> lk(ctx->mu);
    Chunk ch{ps0, zs0, p0, z0, dp0, u0, v0, qv0, ql0, qi0, t0, s0, tr0, tke, pblh, dpdry0, cush,
             umf, slflx, qtflx, flxprc1, flxsnow1, qvten, qlten, qiten, sten, uten, vten, trten,
             qrten, qsten, precip, snow, evapc, cufrc, qcu, qlu, qiu, cbmf, qc, rliq, cnt, cnb, wu};
    const ZwInputs zin{zw, omega, tpert, qpert};
    return run_device(ctx, UWSHCU_FP64, par ? par : &pz, pcols, ncol, pver, ncnst, dt, ch, diag,
                      (hipStream_t)stream, &zin);
}

int uwshcu_forward_chunks_device(uwshcu_ctx* ctx, int precision, const uwshcu_params* par, int nchunk,
                                 const uwshcu_chunk* chunks, int pver, int ncnst, double dt, void* stream) {
    if (!ctx || nchunk < 1 || !chunks || pver < 3 || ncnst < 0 || ncnst > UWSHCU_NCNST_MAX) return UWSHCU_ERR_ARG;
    long long n = 0;
    for (int c = 0; c < nchunk; ++c) {
        const uwshcu_chunk& k = chunks[c];
        if (k.pcols < 1 || k.ncol < 0 || k.ncol > k.pcols) return UWSHCU_ERR_ARG;
        if (k.ncol == 0) continue;
        const void* req[] = {k.ps0, k.zs0, k.p0, k.z0, k.dp0, k.u0, k.v0, k.qv0, k.ql0, k.qi0, k.t0, k.s0, k.tke,
                             k.pblh, k.dpdry0, k.cush, k.umf, k.slflx, k.qtflx, k.flxprc1, k.flxsnow1, k.wu,
                             k.qvten, k.qlten, k.qiten, k.sten, k.uten, k.vten, k.qrten, k.qsten, k.evapc,
                             k.cufrc, k.qcu, k.qlu, k.qiu, k.qc, k.precip, k.snow, k.cbmf, k.rliq, k.cnt, k.cnb};
        for (const void* q : req)
            if (!q) return UWSHCU_ERR_ARG;
        if (ncnst > 0 && (!k.tr0 || !k.trten)) return UWSHCU_ERR_ARG;
        n += k.ncol;
    }
    if (n == 0) return UWSHCU_OK;
    if (n > 0x7fffffffLL) return UWSHCU_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (!hip_ok(hipSetDevice(ctx->device))) return UWSHCU_ERR_HIP;
    const hipStream_t s = (hipStream_t)stream;
    const int L = pver, KS = pver + 1, TR = pver * ncnst;

    // Fields in packed order.  Inputs (gathered): cush last, shared with the outputs.
    constexpr int nin = 17, nout = 28;
    auto in_ptrs = [](const uwshcu_chunk& k) {
        return std::array<double*, nin>{const_cast<double*>(k.ps0), const_cast<double*>(k.zs0), const_cast<double*>(k.tke),
                                    const_cast<double*>(k.p0),  const_cast<double*>(k.z0),  const_cast<double*>(k.dp0),
                                    const_cast<double*>(k.u0),  const_cast<double*>(k.v0),  const_cast<double*>(k.qv0),
                                    const_cast<double*>(k.ql0), const_cast<double*>(k.qi0), const_cast<double*>(k.t0),
                                    const_cast<double*>(k.s0),  const_cast<double*>(k.dpdry0),
                                    const_cast<double*>(k.tr0), const_cast<double*>(k.pblh), k.cush};
    };
    const std::array<int, nin> in_rows{KS, KS, KS, L, L, L, L, L, L, L, L, L, L, L, TR, 1, 1};
    // Outputs (scattered): cush again, at the in-cush packed array.
    auto out_ptrs = [](const uwshcu_chunk& k) {
        return std::array<double*, nout>{k.umf,   k.slflx, k.qtflx, k.flxprc1, k.flxsnow1, k.wu,     k.qvten,
                                    k.qlten, k.qiten, k.sten,  k.uten,    k.vten,     k.qrten,  k.qsten,
                                    k.evapc, k.cufrc, k.qcu,   k.qlu,     k.qiu,      k.qc,     k.trten,
                                    k.precip, k.snow, k.cbmf,  k.rliq,    k.cnt,      k.cnb,    k.cush};
    };
    const std::array<int, nout> out_rows{KS, KS, KS, KS, KS, KS, L, L, L, L, L, L, L, L, L, L, L, L, L, L, TR,
                                         1,  1,  1,  1,  1,  1,  1};

    // packed arrays: inputs, then the outputs other than cush
    size_t rows_total = 0;
    for (int r : in_rows) rows_total += (size_t)r;
    for (int f = 0; f < nout - 1; ++f) rows_total += (size_t)out_rows[f];
    const size_t pbytes = rows_total * (size_t)n * sizeof(double);
    if (pbytes > ctx->packed_bytes) {
        if (ctx->d_packed) {
            if (!hip_ok(hipDeviceSynchronize())) return UWSHCU_ERR_HIP;
            (void)hipFree(ctx->d_packed);
            ctx->d_packed = nullptr;
            ctx->packed_bytes = 0;
        }
        if (!hip_ok(hipMalloc((void**)&ctx->d_packed, pbytes))) return UWSHCU_ERR_NOMEM;
        ctx->packed_bytes = pbytes;
    }
    std::array<double*, nin> pin{};
    std::array<double*, nout> pout{};
    {
        double* q = ctx->d_packed;
        for (int f = 0; f < nin; ++f) { pin[f] = q; q += (size_t)in_rows[f] * (size_t)n; }
        for (int f = 0; f < nout - 1; ++f) { pout[f] = q; q += (size_t)out_rows[f] * (size_t)n; }
        pout[nout - 1] = pin[nin - 1];
    }

    // device tables, built in pinned host memory and uploaded on the stream
    auto al8 = [](size_t b) { return (b + 7) & ~(size_t)7; };
    const size_t o_off = 0, o_pcols = al8(o_off + sizeof(long long) * ((size_t)nchunk + 1));
    const size_t o_rin = al8(o_pcols + sizeof(int) * (size_t)nchunk), o_rout = al8(o_rin + sizeof(int) * nin);
    const size_t o_pin = al8(o_rout + sizeof(int) * nout), o_pout = o_pin + sizeof(double*) * nin;
    const size_t o_tin = o_pout + sizeof(double*) * nout, o_tout = o_tin + sizeof(double*) * (size_t)nin * nchunk;
    const size_t tbytes = o_tout + sizeof(double*) * (size_t)nout * nchunk;
    if (!ctx->ctab_event && !hip_ok(hipEventCreateWithFlags(&ctx->ctab_event, hipEventDisableTiming)))
        return UWSHCU_ERR_HIP;
    if (ctx->ctab_pending && !hip_ok(hipEventSynchronize(ctx->ctab_event))) return UWSHCU_ERR_HIP;
    ctx->ctab_pending = false;
    if (tbytes > ctx->h_ctab_bytes) {
        if (ctx->h_ctab) (void)hipHostFree(ctx->h_ctab);
        ctx->h_ctab = nullptr;
        ctx->h_ctab_bytes = 0;
        if (!hip_ok(hipHostMalloc((void**)&ctx->h_ctab, tbytes, hipHostMallocDefault))) return UWSHCU_ERR_NOMEM;
        ctx->h_ctab_bytes = tbytes;
    }
    if (tbytes > ctx->ctab_bytes) {
        if (ctx->d_ctab) {
            if (!hip_ok(hipDeviceSynchronize())) return UWSHCU_ERR_HIP;
            (void)hipFree(ctx->d_ctab);
            ctx->d_ctab = nullptr;
            ctx->ctab_bytes = 0;
        }
        if (!hip_ok(hipMalloc((void**)&ctx->d_ctab, tbytes))) return UWSHCU_ERR_NOMEM;
        ctx->ctab_bytes = tbytes;
    }
    char* h = ctx->h_ctab;
    long long* off = reinterpret_cast<long long*>(h + o_off);
    int* pcols = reinterpret_cast<int*>(h + o_pcols);
    off[0] = 0;
    for (int c = 0; c < nchunk; ++c) {
        off[c + 1] = off[c] + chunks[c].ncol;
        pcols[c] = chunks[c].pcols;
        const std::array<double*, nin> pi = in_ptrs(chunks[c]);
        const std::array<double*, nout> po = out_ptrs(chunks[c]);
        for (int f = 0; f < nin; ++f)
            reinterpret_cast<double**>(h + o_tin)[(size_t)f * nchunk + c] = chunks[c].ncol ? pi[f] : nullptr;
        for (int f = 0; f < nout; ++f)
            reinterpret_cast<double**>(h + o_tout)[(size_t)f * nchunk + c] = chunks[c].ncol ? po[f] : nullptr;
    }
    std::memcpy(h + o_rin, in_rows.data(), sizeof(int) * nin);
    std::memcpy(h + o_rout, out_rows.data(), sizeof(int) * nout);
    std::memcpy(h + o_pin, pin.data(), sizeof(double*) * nin);
    std::memcpy(h + o_pout, pout.data(), sizeof(double*) * nout);
    if (!hip_ok(hipMemcpyAsync(ctx->d_ctab, h, tbytes, hipMemcpyHostToDevice, s))) return UWSHCU_ERR_HIP;
    if (!hip_ok(hipEventRecord(ctx->ctab_event, s))) return UWSHCU_ERR_HIP;
    ctx->ctab_pending = true;

    char* d = ctx->d_ctab;
    const uwshcu::ChunkTable tin{nchunk, nin, n, reinterpret_cast<const long long*>(d + o_off),
                                 reinterpret_cast<const int*>(d + o_pcols), reinterpret_cast<const int*>(d + o_rin),
                                 reinterpret_cast<double* const*>(d + o_pin), reinterpret_cast<double* const*>(d + o_tin)};
    const uwshcu::ChunkTable tout{nchunk, nout, n, reinterpret_cast<const long long*>(d + o_off),
                                  reinterpret_cast<const int*>(d + o_pcols), reinterpret_cast<const int*>(d + o_rout),
                                  reinterpret_cast<double* const*>(d + o_pout),
                                  reinterpret_cast<double* const*>(d + o_tout)};
    if (!hip_ok(uwshcu::launch_chunks_gather(tin, s))) return UWSHCU_ERR_HIP;
    Chunk ch{};
    ch.ps0 = pin[0]; ch.zs0 = pin[1]; ch.tke = pin[2]; ch.p0 = pin[3]; ch.z0 = pin[4]; ch.dp0 = pin[5];
    ch.u0 = pin[6]; ch.v0 = pin[7]; ch.qv0 = pin[8]; ch.ql0 = pin[9]; ch.qi0 = pin[10]; ch.t0 = pin[11];
    ch.s0 = pin[12]; ch.dpdry0 = pin[13]; ch.tr0 = pin[14]; ch.pblh = pin[15]; ch.cush = pin[16];
    ch.umf = pout[0]; ch.slflx = pout[1]; ch.qtflx = pout[2]; ch.flxprc1 = pout[3]; ch.flxsnow1 = pout[4];
    ch.wu = pout[5]; ch.qvten = pout[6]; ch.qlten = pout[7]; ch.qiten = pout[8]; ch.sten = pout[9];
    ch.uten = pout[10]; ch.vten = pout[11]; ch.qrten = pout[12]; ch.qsten = pout[13]; ch.evapc = pout[14];
    ch.cufrc = pout[15]; ch.qcu = pout[16]; ch.qlu = pout[17]; ch.qiu = pout[18]; ch.qc = pout[19];
    ch.trten = pout[20]; ch.precip = pout[21]; ch.snow = pout[22]; ch.cbmf = pout[23]; ch.rliq = pout[24];
    ch.cnt = pout[25]; ch.cnb = pout[26];
    const int rc = run_device(ctx, precision, par, (int)n, (int)n, pver, ncnst, dt, ch, nullptr, s);
    if (rc) return rc;
    if (!hip_ok(uwshcu::launch_chunks_scatter(tout, s))) return UWSHCU_ERR_HIP;
    return UWSHCU_OK;
}

int uwshcu_merge_shallow_device(int pcols, int ncol, int pver, double latvap, const uwshcu_merge_fields* f,
                                void* stream) {
    if (pcols < 1 || ncol < 0 || ncol > pcols || pver < 1 || !f) return UWSHCU_ERR_ARG;
    const void* req[] = {f->umf,   f->slflx, f->qtflx, f->qrten, f->qsten,   f->qcu,    f->qc2,   f->rliq2,
                         f->cnt2,  f->cnb2,  f->cbmf,  f->pmid,  f->rprddp,  f->cmfmc,  f->cnt,   f->cnb,
                         f->rliq,  f->qc,    f->icwmr, f->rprdsh, f->rprdtot, f->cmfsl, f->cmflq, f->freqsh,
                         f->pcnt,  f->pcnb,  f->massflx_sh};
    for (const void* q : req)
        if (!q) return UWSHCU_ERR_ARG;
    if (!hip_ok(uwshcu::launch_merge_shallow(pcols, ncol, pver, latvap, *f, (hipStream_t)stream)))
        return UWSHCU_ERR_HIP;
    return UWSHCU_OK;
}

int uwshcu_adjoint_device(uwshcu_ctx* ctx, const uwshcu_params* par, int pcols, int ncol, int pver, int ncnst,
                          double dt, const double* ps0, const double* zs0, const double* p0, const double* z0,
                          const double* dp0, const double* u0, const double* v0, const double* qv0,
                          const double* ql0, const double* qi0, const double* t0, const double* s0,
                          const double* tr0, const double* tke, const double* cldfrc, const double* concldfrc,
                          const double* pblh, double* cush, double* umf, double* slflx, double* qtflx,
                          double* flxprc1, double* flxsnow1, double* qvten, double* qlten, double* qiten,
                          double* sten, double* uten, double* vten, double* trten, double* qrten, double* qsten,
                          double* precip, double* snow, double* evapc, double* cufrc, double* qcu, double* qlu,
                          double* qiu, double* cbmf, double* qc, double* rliq, double* cnt, double* cnb,
                          const double* dpdry0, double* wu, int* diag, const uwshcu_cotangents* cot,
                          const uwshcu_gradients* grad, void* stream) {
    (void)cldfrc;
    (void)concldfrc;
    if (!ctx) return UWSHCU_ERR_ARG;
    const void* req[] = {ps0, zs0, p0, z0, dp0, u0, v0, qv0, ql0, qi0, t0, s0, tke, pblh, cush, umf, slflx,
                         qtflx, flxprc1, flxsnow1, qvten, qlten, qiten, sten, uten, vten, qrten, qsten, precip,
                         snow, evapc, cufrc, qcu, qlu, qiu, cbmf, qc, rliq, cnt, cnb, dpdry0, wu};
    for (const void* p : req)
        if (!p) return UWSHCU_ERR_ARG;
    if (ncnst > 0 && (!tr0 || !trten)) return UWSHCU_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    Chunk ch{ps0, zs0, p0, z0, dp0, u0, v0, qv0, ql0, qi0, t0, s0, tr0, tke, pblh, dpdry0, cush,
             umf, slflx, qtflx, flxprc1, flxsnow1, qvten, qlten, qiten, sten, uten, vten, trten,
             qrten, qsten, precip, snow, evapc, cufrc, qcu, qlu, qiu, cbmf, qc, rliq, cnt, cnb, wu};
    return run_adjoint(ctx, par, pcols, ncol, pver, ncnst, dt, ch, diag, cot, grad, (hipStream_t)stream);
}

int uwshcu_status(uwshcu_ctx* ctx, void* stream, int* bad_col) {
    if (!ctx) return UWSHCU_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (!hip_ok(hipSetDevice(ctx->device))) return UWSHCU_ERR_HIP;
    int h[2] = {0, -1};
    if (!hip_ok(hipMemcpyAsync(h, ctx->d_err, sizeof(h), hipMemcpyDeviceToHost, (hipStream_t)stream)) ||
        !hip_ok(hipStreamSynchronize((hipStream_t)stream)))
        return UWSHCU_ERR_HIP;
    if (!hip_ok(uwshcu::launch_clear_status(ctx->d_err, (hipStream_t)stream))) return UWSHCU_ERR_HIP;
    if (bad_col) *bad_col = h[0] ? h[1] : -1;
    return h[0];
}

// Host-array forward: stage the inputs to the device, run, copy the outputs back.
// zwp != nullptr selects the modified scheme, whose host inputs omega_h (pcols, pver)
// and tpert_h, qpert_h (pcols) are staged too.
static int forward_host_impl(uwshcu_ctx* ctx, int precision, const uwshcu_params* par, int pcols, int ncol,
                             int pver, int ncnst, double dt, const double* ps0, const double* zs0, const double* p0,
                             const double* z0, const double* dp0, const double* u0, const double* v0,
                             const double* qv0, const double* ql0, const double* qi0, const double* t0,
                             const double* s0, const double* tr0, const double* tke, const double* pblh,
                             double* cush, double* umf, double* slflx, double* qtflx, double* flxprc1,
                             double* flxsnow1, double* qvten, double* qlten, double* qiten, double* sten,
                             double* uten, double* vten, double* trten, double* qrten, double* qsten,
                             double* precip, double* snow, double* evapc, double* cufrc, double* qcu, double* qlu,
                             double* qiu, double* cbmf, double* qc, double* rliq, double* cnt, double* cnb,
                             const double* dpdry0, double* wu, int* diag, int* bad_col,
                             const uwshcu_zw_params* zwp, const double* omega_h, const double* tpert_h,
                             const double* qpert_h) {
    if (!ctx || pcols < 1 || ncol < 0 || ncol > pcols || pver < 3 || ncnst < 0 || ncnst > UWSHCU_NCNST_MAX)
        return UWSHCU_ERR_ARG;
    if (zwp && (!omega_h || !tpert_h || !qpert_h)) return UWSHCU_ERR_ARG;
    if (bad_col) *bad_col = -1;
    if (ncol == 0) return UWSHCU_OK;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (!hip_ok(hipSetDevice(ctx->device))) return UWSHCU_ERR_HIP;
    const size_t L = (size_t)pcols * pver, I = (size_t)pcols * (pver + 1), S = (size_t)pcols;
    const size_t T = (size_t)pcols * pver * ncnst;
    // staging layout: inputs then outputs, each a full (pcols, ...) array
    const size_t n_in = 3 * I + 11 * L + T + 2 * S + (zwp ? L + 2 * S : 0);
    const size_t n_out = 6 * I + 14 * L + T + 6 * S;
    const size_t bytes = (n_in + n_out) * sizeof(double);
    if (bytes > ctx->stage_bytes) {
        if (ctx->d_stage) (void)hipFree(ctx->d_stage);
        ctx->d_stage = nullptr;
        ctx->stage_bytes = 0;
        if (!hip_ok(hipMalloc((void**)&ctx->d_stage, bytes))) return UWSHCU_ERR_NOMEM;
        ctx->stage_bytes = bytes;
    }
    int* ddiag = nullptr;
    if (diag) {
        const size_t db = (size_t)UWSHCU_NDIAG * pcols * sizeof(int);
        if (db > ctx->diag_bytes) {
            if (ctx->d_diag) (void)hipFree(ctx->d_diag);
            ctx->d_diag = nullptr;
            ctx->diag_bytes = 0;
            if (!hip_ok(hipMalloc((void**)&ctx->d_diag, db))) return UWSHCU_ERR_NOMEM;
            ctx->diag_bytes = db;
        }
        ddiag = ctx->d_diag;
    }
    double* base = ctx->d_stage;
    size_t off = 0;
    auto take = [&](size_t n) { double* p = base + off; off += n; return p; };
    hipStream_t st = ctx->stream;
    bool ok = true;
    auto h2d = [&](const double* h, size_t n) {
        double* d = take(n);
        if (n) ok = ok && hip_ok(hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, st));
        return (const double*)d;
    };
    Chunk ch{};
    ch.ps0 = h2d(ps0, I); ch.zs0 = h2d(zs0, I); ch.tke = h2d(tke, I);
    ch.p0 = h2d(p0, L); ch.z0 = h2d(z0, L); ch.dp0 = h2d(dp0, L); ch.dpdry0 = h2d(dpdry0, L);
    ch.u0 = h2d(u0, L); ch.v0 = h2d(v0, L); ch.qv0 = h2d(qv0, L); ch.ql0 = h2d(ql0, L); ch.qi0 = h2d(qi0, L);
    ch.t0 = h2d(t0, L); ch.s0 = h2d(s0, L);
    ch.tr0 = h2d(tr0, T);
    ch.pblh = h2d(pblh, S);
    ch.cush = const_cast<double*>(h2d(cush, S));
    ZwInputs zin{zwp, nullptr, nullptr, nullptr};
    if (zwp) {
        zin.omega = h2d(omega_h, L);
        zin.tpert = h2d(tpert_h, S);
        zin.qpert = h2d(qpert_h, S);
    }
    ch.umf = take(I); ch.slflx = take(I); ch.qtflx = take(I); ch.flxprc1 = take(I); ch.flxsnow1 = take(I);
    ch.wu = take(I);
    ch.qvten = take(L); ch.qlten = take(L); ch.qiten = take(L); ch.sten = take(L); ch.uten = take(L);
    ch.vten = take(L); ch.qrten = take(L); ch.qsten = take(L); ch.evapc = take(L); ch.cufrc = take(L);
    ch.qcu = take(L); ch.qlu = take(L); ch.qiu = take(L); ch.qc = take(L);
    ch.trten = take(T);
    ch.precip = take(S); ch.snow = take(S); ch.cbmf = take(S); ch.rliq = take(S); ch.cnt = take(S); ch.cnb = take(S);
    if (!ok) return UWSHCU_ERR_HIP;
    int rc = run_device(ctx, precision, par, pcols, ncol, pver, ncnst, dt, ch, ddiag, st, zwp ? &zin : nullptr);
    if (rc) return rc;
    // outputs: the first ncol columns of each row only (the caller's padding columns
    // ncol..pcols-1 stay as they were, as in the reference, which writes 1:iend)
    const size_t rows_i = (size_t)pver + 1, rows_l = (size_t)pver, rows_t = (size_t)pver * ncnst;
    auto d2h = [&](void* h, const void* d, size_t rows, size_t elem) {
        if (rows)
            ok = ok && hip_ok(hipMemcpy2DAsync(h, (size_t)pcols * elem, d, (size_t)pcols * elem, (size_t)ncol * elem,
                                               rows, hipMemcpyDeviceToHost, st));
    };
    const size_t e8 = sizeof(double);
    d2h(cush, ch.cush, 1, e8);
    d2h(umf, ch.umf, rows_i, e8); d2h(slflx, ch.slflx, rows_i, e8); d2h(qtflx, ch.qtflx, rows_i, e8);
    d2h(flxprc1, ch.flxprc1, rows_i, e8); d2h(flxsnow1, ch.flxsnow1, rows_i, e8); d2h(wu, ch.wu, rows_i, e8);
    d2h(qvten, ch.qvten, rows_l, e8); d2h(qlten, ch.qlten, rows_l, e8); d2h(qiten, ch.qiten, rows_l, e8);
    d2h(sten, ch.sten, rows_l, e8); d2h(uten, ch.uten, rows_l, e8); d2h(vten, ch.vten, rows_l, e8);
    d2h(qrten, ch.qrten, rows_l, e8); d2h(qsten, ch.qsten, rows_l, e8); d2h(evapc, ch.evapc, rows_l, e8);
    d2h(cufrc, ch.cufrc, rows_l, e8); d2h(qcu, ch.qcu, rows_l, e8); d2h(qlu, ch.qlu, rows_l, e8);
    d2h(qiu, ch.qiu, rows_l, e8); d2h(qc, ch.qc, rows_l, e8);
    d2h(trten, ch.trten, rows_t, e8);
    d2h(precip, ch.precip, 1, e8); d2h(snow, ch.snow, 1, e8); d2h(cbmf, ch.cbmf, 1, e8); d2h(rliq, ch.rliq, 1, e8);
    d2h(cnt, ch.cnt, 1, e8); d2h(cnb, ch.cnb, 1, e8);
    if (diag) d2h(diag, ddiag, UWSHCU_NDIAG, sizeof(int));
    int h[2] = {0, -1};
    ok = ok && hip_ok(hipMemcpyAsync(h, ctx->d_err, sizeof(h), hipMemcpyDeviceToHost, st)) &&
         hip_ok(hipStreamSynchronize(st)) && hip_ok(uwshcu::launch_clear_status(ctx->d_err, st));
    if (!ok) return UWSHCU_ERR_HIP;
    if (bad_col) *bad_col = h[0] ? h[1] : -1;
    return h[0];
}

int uwshcu_forward_host(uwshcu_ctx* ctx, int precision, const uwshcu_params* par, int pcols, int ncol, int pver,
                        int ncnst, double dt, const double* ps0, const double* zs0, const double* p0,
                        const double* z0, const double* dp0, const double* u0, const double* v0, const double* qv0,
                        const double* ql0, const double* qi0, const double* t0, const double* s0,
                        const double* tr0, const double* tke, const double* cldfrc, const double* concldfrc,
                        const double* pblh, double* cush, double* umf, double* slflx, double* qtflx,
                        double* flxprc1, double* flxsnow1, double* qvten, double* qlten, double* qiten,
                        double* sten, double* uten, double* vten, double* trten, double* qrten, double* qsten,
                        double* precip, double* snow, double* evapc, double* cufrc, double* qcu, double* qlu,
                        double* qiu, double* cbmf, double* qc, double* rliq, double* cnt, double* cnb,
                        const double* dpdry0, double* wu, int* diag, int* bad_col) {
    (void)cldfrc;
    (void)concldfrc;
    return forward_host_impl(ctx, precision, par, pcols, ncol, pver, ncnst, dt, ps0, zs0, p0, z0, dp0, u0, v0, qv0,
                             ql0, qi0, t0, s0, tr0, tke, pblh, cush, umf, slflx, qtflx, flxprc1, flxsnow1, qvten,
                             qlten, qiten, sten, uten, vten, trten, qrten, qsten, precip, snow, evapc, cufrc, qcu,
                             qlu, qiu, cbmf, qc, rliq, cnt, cnb, dpdry0, wu, diag, bad_col, nullptr, nullptr,
                             nullptr, nullptr);
}

int uwshcu_zw_forward_host(uwshcu_ctx* ctx, const uwshcu_params* par, const uwshcu_zw_params* zw, int pcols,
                           int ncol, int pver, int ncnst, double dt, const double* ps0, const double* zs0,
                           const double* p0, const double* z0, const double* dp0, const double* u0, const double* v0,
                           const double* qv0, const double* ql0, const double* qi0, const double* t0,
                           const double* s0, const double* tr0, const double* tke, const double* cldfrc,
                           const double* concldfrc, const double* pblh, double* cush, const double* omega,
                           const double* tpert, const double* qpert, double* umf, double* slflx, double* qtflx,
                           double* flxprc1, double* flxsnow1, double* qvten, double* qlten, double* qiten,
                           double* sten, double* uten, double* vten, double* trten, double* qrten, double* qsten,
                           double* precip, double* snow, double* evapc, double* cufrc, double* qcu, double* qlu,
                           double* qiu, double* cbmf, double* qc, double* rliq, double* cnt, double* cnb,
                           const double* dpdry0, double* wu, int* diag, int* bad_col) {
    (void)cldfrc;
    (void)concldfrc;
    if (!zw) return UWSHCU_ERR_ARG;
    uwshcu_params pz;
    uwshcu_zw_default_params(&pz, nullptr);
    return forward_host_impl(ctx, UWSHCU_FP64, par ? par : &pz, pcols, ncol, pver, ncnst, dt, ps0, zs0, p0, z0, dp0,
                             u0, v0, qv0, ql0, qi0, t0, s0, tr0, tke, pblh, cush, umf, slflx, qtflx, flxprc1,
                             flxsnow1, qvten, qlten, qiten, sten, uten, vten, trten, qrten, qsten, precip, snow,
                             evapc, cufrc, qcu, qlu, qiu, cbmf, qc, rliq, cnt, cnb, dpdry0, wu, diag, bad_col, zw,
                             omega, tpert, qpert);
}

void bxcompute_uwshcu_inv_(int* pcols, int* pver, int* ncol, int* pcnst, double* ztodt, double* ps0_inv,
                           double* zs0_inv, double* p0_inv, double* z0_inv, double* dp0_inv, double* u0_inv,
                           double* v0_inv, double* qv0_inv, double* ql0_inv, double* qi0_inv, double* t0_inv,
                           double* s0_inv, double* tr0_inv, double* tke_inv, double* cldfrct_inv,
                           double* concldfrct_inv, double* pblh, double* cush, double* umf_inv,
                           double* slflx_inv, double* qtflx_inv, double* flxprc1_inv, double* flxsnow1_inv,
                           double* qvten_inv, double* qlten_inv, double* qiten_inv, double* sten_inv,
                           double* uten_inv, double* vten_inv, double* trten_inv, double* qrten_inv,
                           double* qsten_inv, double* precip, double* snow, double* evapc_inv, double* cufrc_inv,
                           double* qcu_inv, double* qlu_inv, double* qiu_inv, double* cbmf, double* qc_inv,
                           double* rliq, double* cnt_inv, double* cnb_inv, int* lchnk, double* dpdry0_inv,
                           double* wu_inv) {
    std::lock_guard<std::mutex> lk(g_dropin_mu);
    if (!g_dropin_par_set) {
        // uwshcu.F90:241-242: the reference stops the model when uwshcu_rpen is unset
        std::fprintf(stderr, "bxcompute_uwshcu_inv_: uwshcu_rpen is not set; call uwshcu_set_dropin_params first "
                             "(chunk %d not computed)\n", lchnk ? *lchnk : -1);
        g_dropin_status = UWSHCU_ERR_PARAMS;
        g_dropin_badcol = -1;
        return;
    }
    if (!g_dropin_ctx) {
        int rc = uwshcu_ctx_create(nullptr, nullptr, -1, &g_dropin_ctx);
        if (rc) {
            std::fprintf(stderr, "bxcompute_uwshcu_inv_: context creation failed (%d)\n", rc);
            g_dropin_status = rc;
            return;
        }
    }
    int bad = -1;
    int rc = uwshcu_forward_host(g_dropin_ctx, UWSHCU_FP64, &g_dropin_par, *pcols, *ncol, *pver, *pcnst, *ztodt,
                                 ps0_inv, zs0_inv, p0_inv, z0_inv, dp0_inv, u0_inv, v0_inv, qv0_inv, ql0_inv,
                                 qi0_inv, t0_inv, s0_inv, tr0_inv, tke_inv, cldfrct_inv, concldfrct_inv, pblh, cush,
                                 umf_inv, slflx_inv, qtflx_inv, flxprc1_inv, flxsnow1_inv, qvten_inv, qlten_inv,
                                 qiten_inv, sten_inv, uten_inv, vten_inv, trten_inv, qrten_inv, qsten_inv, precip,
                                 snow, evapc_inv, cufrc_inv, qcu_inv, qlu_inv, qiu_inv, cbmf, qc_inv, rliq, cnt_inv,
                                 cnb_inv, dpdry0_inv, wu_inv, nullptr, &bad);
    g_dropin_status = rc;
    g_dropin_badcol = bad;
    if (rc)
        std::fprintf(stderr, "bxcompute_uwshcu_inv_: chunk %d failed with code %d at column %d\n",
                     lchnk ? *lchnk : -1, rc, bad);
}

int uwshcu_set_dropin_params(const uwshcu_params* par) {
    std::lock_guard<std::mutex> lk(g_dropin_mu);
    if (!par) {
        g_dropin_par_set = false;
        return UWSHCU_OK;
    }
    if (!(par->rpen == par->rpen) || !(par->rle == par->rle) || !(par->rkm == par->rkm)) return UWSHCU_ERR_ARG;
    g_dropin_par = *par;
    g_dropin_par_set = true;
    return UWSHCU_OK;
}

int uwshcu_dropin_last_status(int* bad_col) {
    std::lock_guard<std::mutex> lk(g_dropin_mu);
    if (bad_col) *bad_col = g_dropin_badcol;
    return g_dropin_status;
}

}  // extern "C"
