// uwshcu_adjoint.h -- tape set-up and reverse sweep for one column (host and device).
//
// The adjoint of compute_uwshcu_inv is taken by recording the column's forward pass
// on a tape (Column<AD, AD>, see uwshcu_ad.h) and sweeping it backwards.  The
// independent variables are t0, qv0, u0, v0 (layers), tke (interfaces) and the
// tunables rpen, rle, rkm; the first ad_inputs(mkx) tape nodes are their leaves.
// pblh enters the scheme only through the integer inversion level kinv
// (uwshcu.F90:1455-1468), so its derivative is zero wherever it exists.
#ifndef UWSHCU_ADJOINT_H
#define UWSHCU_ADJOINT_H

#include <type_traits>

#include "uwshcu_column.h"

namespace uwshcu {

// Start a column's tape: contiguous in nodes (cap nodes; nullptr = count only), or,
// with pages != nullptr, paged from shard `shard` of pool with the column's page table
// in pages.  The first nin nodes are the leaves (the independent variables).
UWSHCU_HD inline void ad_begin(TapeCtx& t, TapeHot& h, TapeNode* nodes, int* pages, TapePool* pool, int shard, int cap,
                               int nin) {
    t.pages = pages;
    t.pool = pool;
    t.adj = 0;
    t.cap = cap;
    t.overflow = 0;
    t.pad = 0;
    h.n = 0;
    if (pages) {
        h.nodes = nullptr;
        h.limit = 0;
        t.ctr = &pool->next[shard * 32];
        t.shard_base = shard * pool->per_shard;
    } else {
        h.nodes = nodes ? nodes + (cap - 1) : nullptr;
        h.limit = cap;
        t.ctr = nullptr;
        t.shard_base = 0;
    }
    for (int q = 0; q < nin; ++q) {
        if (q >= h.limit && !tape_grow(t, h, q)) break;
        h.n = q + 1;
        if (TapeNode* top = h.nodes) tape_store(top - q, 0.0, 0.0, -1, -1);
    }
    t.hot = h;
}

// The sweep's accesses to a node's adjoint in the tape.  On the GPU they are agent-scope
// atomics: an addition is a no-return FP64 atomic add (the thread does not wait for it),
// and a read is a coherent load that, being sequenced after the thread's earlier
// additions to the same node, returns their sum.  The additions to one node are made by
// one thread, in program order, so each adjoint still receives its contributions in a
// fixed order and rounds as the host's plain additions do.
UWSHCU_HD inline void adj_add(UWSHCU_GMEM double* p, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
    __hip_atomic_fetch_add(p, c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#else
    *p += c;
#endif
}
UWSHCU_HD inline double adj_load(const UWSHCU_GMEM double* p) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#else
    return *p;
#endif
}

// Cotangents of the outputs, in the caller's layout (nullable per field).
struct AdCotangents {
    const double* f[ADF_TR];         // umf .. qc, as AdField
    const double* trten;
    const double* scal[ADC_COUNT];   // precip, snow, cbmf, rliq, cnt, cnb, cush
};

// Gradients of the inputs, in the caller's layout (nullable per field); par is
// [3][ldc]: the column's contribution to d/d(rpen, rle, rkm).
struct AdGradients {
    double *t0, *qv0, *u0, *v0, *tke, *pblh, *par;
};

// Reverse sweep of one column: seed the outputs' adjoints from the cotangents,
// propagate back through the tape, write the input gradients.  adjv[q] is node q's
// adjoint (n entries, zero on entry).
//
// With a window of S >= kSweepBatch entries (win(q % S), and a spare entry win(S)), the
// adjoints of recent nodes are kept in it: 82 % of the operands a node refers to lie
// within 64 nodes of it (35-level columns).  The sweep works kSweepBatch (K) nodes at a
// time, in batches aligned to K (so that a batch never straddles a tape page); when a
// batch starts at node q0 the window holds the adjoints of nodes q0 - S + 1 .. q0.  A
// contribution to one of those goes to the window, one to an older node to adjv (a far
// contribution).  When the batch ends, its K nodes leave the window and nodes
// q0 - S - K + 1 .. q0 - S take their entries, carrying what far contributions they
// received.  Every adjoint still receives its contributions in decreasing node order,
// so the result is the same, bit for bit, for every S, and as without a window.
//
// The memory operations of a batch are issued together at its end: its far
// contributions (no-return atomic adds, in step order), then the loads of the adjoints
// that enter the window and of the next batch's tape nodes (indices clamped into the
// batch's page), so that one memory latency is paid per batch and no step waits on
// memory.  (A far contribution can wait to the batch's end: its node enters the window
// at a later batch's end.)  The page-table entry of a batch is loaded at the end of the
// batch two before it, so that it has arrived when its nodes are loaded.  Within a
// step, the node's adjoint and its operands' window entries are
// read together and written back unconditionally (an operand outside the window uses
// the spare entry), so that a step waits for the window once.
#ifndef UWSHCU_SWEEP_BATCH
#define UWSHCU_SWEEP_BATCH 16
#endif
constexpr int kSweepBatch = UWSHCU_SWEEP_BATCH;
static_assert(kTapePage % kSweepBatch == 0, "a sweep batch lies within one tape page");

template <class Win>
UWSHCU_HD inline void ad_reverse(const TapeCtx& t, double* adjv_, const int* nodes, size_t wld, int slot, int mkx,
                                 int ncnst, size_t ldc, int col, const AdCotangents& cot, const AdGradients& g, Win win,
                                 int S) {
    const int KS = mkx + 1;
    // the header's fields in registers: the sweep's stores could alias the header
    const TapeCtx th = t;
    const int n_nodes = th.hot.n;
    UWSHCU_GMEM double* const adjv = (UWSHCU_GMEM double*)adjv_;
    auto P = [&](int q) { return adjv + q; };
    const bool paged = th.pages != nullptr;
    const UWSHCU_GMEM int* const pages = (const UWSHCU_GMEM int*)th.pages;
    const UWSHCU_GMEM TapeNode* const pool = paged ? (const UWSHCU_GMEM TapeNode*)th.pool->base : nullptr;
    const UWSHCU_GMEM TapeNode* const contig = (const UWSHCU_GMEM TapeNode*)th.hot.nodes;   // biased: node q at contig - q
    // node q of a batch whose page (paged tapes) is pg lives at top(pg) - (q & qmask)
    const int qmask = paged ? kTapePage - 1 : 0x7fffffff;
    auto top = [&](int pg) { return paged ? pool + (size_t)pg * kTapePage + (kTapePage - 1) : contig; };
    auto page_of = [&](int q) { return paged ? pages[q >> kTapePageLog2] : 0; };
    auto node = [&](int f, int k) { return nodes[((size_t)f * KS + k) * wld + slot]; };
    auto row = [&](int k) { return (size_t)(mkx - k) * ldc + col; };
    for (int f = 0; f < ADF_TR; ++f) {
        if (!cot.f[f]) continue;
        for (int k = (f <= ADF_WU ? 0 : 1); k <= mkx; ++k) {
            const int nd = node(f, k);
            if (nd >= 0) adj_add(P(nd), cot.f[f][row(k)]);
        }
    }
    if (cot.trten)
        for (int m = 1; m <= ncnst; ++m)
            for (int k = 1; k <= mkx; ++k) {
                const int nd = node(ADF_TR + m - 1, k);
                if (nd >= 0) adj_add(P(nd), cot.trten[((size_t)(m - 1) * mkx + (mkx - k)) * ldc + col]);
            }
    for (int q = 0; q < ADC_COUNT; ++q) {
        if (!cot.scal[q]) continue;
        const int nd = node(ADF_TR + ncnst + q, 0);
        if (nd >= 0) adj_add(P(nd), cot.scal[q][col]);
    }
    const int nin = ad_inputs(mkx);
    const int n = n_nodes;
    // window entry of node x >= 0 (unsigned: a mask when S is a power of two)
    auto wslot = [&](int x) { return (int)((unsigned)x % (unsigned)S); };
    if (S > 0) {
        constexpr int K = kSweepBatch;
        double da[K], db[K];
        int ia[K], ib[K];
        // the nodes of the batch at q0, clamped into [nin, n - 1] (the same page)
        auto load_batch = [&](int q0, const UWSHCU_GMEM TapeNode* tb) {
#pragma unroll
            for (int u = 0; u < K; ++u) {
                int q = q0 - u < n - 1 ? q0 - u : n - 1;
                q = q > nin ? q : nin;
                const UWSHCU_GMEM TapeNode* p = tb - (q & qmask);
                da[u] = p->da;
                db[u] = p->db;
                ia[u] = p->a;
                ib[u] = p->b;
            }
        };
        int q0 = (n - 1) | (K - 1);
        for (int q = n - 1; q >= 0 && q > q0 - S; --q) win(wslot(q)) = adj_load(P(q));
        if (n > nin) load_batch(q0, top(page_of(n - 1)));
        // page of the next batch, loaded a batch ahead of its use (clamped to node 0)
        int pg_next = page_of(q0 - K >= 0 ? q0 - K : 0);
        for (; q0 >= nin; q0 -= K) {
            const int lo = q0 - S;     // the window holds the nodes above lo
            const int lo1 = lo > -1 ? lo : -1;   // an operand a is in the window iff a > lo1
            int fx[2 * K];             // far contributions of the batch: node (or -1), value
            double fc[2 * K];
            // the steps of the batch; kAll: every node of the batch lies in [nin, n - 1]
            auto steps = [&](auto all) {
                constexpr bool kAll = decltype(all)::value;
#pragma unroll
                for (int u = 0; u < K; ++u) {
                    // a step outside [nin, n - 1] still runs, on the spare entry, so that
                    // every path consumes the batch's loads
                    const int q = q0 - u;
                    const bool act = kAll || (q >= nin && q < n);
                    const int a = ia[u], b = ib[u];
                    const bool nearA = act && a > lo1, nearB = act && b > lo1;
                    const int sa = nearA ? wslot(a) : S, sb = nearB ? wslot(b) : S;
                    const double w = win(act ? wslot(q) : S), va = win(sa), vb = win(sb);
                    const bool live = act && w != 0.0;
                    const double ca = w * da[u], cb = w * db[u];
                    fx[2 * u] = live && a >= 0 && !nearA ? a : -1;
                    fx[2 * u + 1] = live && b >= 0 && !nearB ? b : -1;
                    fc[2 * u] = ca;
                    fc[2 * u + 1] = cb;
                    const double va1 = live && nearA ? va + ca : va;
                    const double vb0 = nearA && sb == sa ? va1 : vb;    // a == b, both in the window
                    const double vb1 = live && nearB ? vb0 + cb : vb;
                    win(sa) = va1;
                    win(sb) = vb1;
                }
            };
            if (q0 - (K - 1) >= nin && q0 < n) steps(std::true_type{});
            else steps(std::false_type{});
            // the batch's far contributions, in step order, then the batch's nodes leave the
            // window and nodes lo - u enter, into their entries
#pragma unroll
            for (int v = 0; v < 2 * K; ++v)
                if (fx[v] >= 0) adj_add(P(fx[v]), fc[v]);
            double enter[K];
#pragma unroll
            for (int u = 0; u < K; ++u) enter[u] = adj_load(P(lo - u > 0 ? lo - u : 0));
            const int pg_after = page_of(q0 - 2 * K >= 0 ? q0 - 2 * K : 0);
            if (q0 - K >= nin) load_batch(q0 - K, top(pg_next));
            pg_next = pg_after;
            // written unconditionally (to the spare entry when nothing enters), so that no
            // load is left pending on some paths into the next batch
#pragma unroll
            for (int u = 0; u < K; ++u) win(q0 - u >= nin && lo - u >= 0 ? wslot(q0 - u) : S) = enter[u];
        }
    } else {
        for (int q = n - 1; q >= nin; --q) {
            const double w = adj_load(P(q));
            if (w == 0.0) continue;
            const UWSHCU_GMEM TapeNode& nd = *(top(page_of(q)) - (q & qmask));
            if (nd.a >= 0) adj_add(P(nd.a), w * nd.da);
            if (nd.b >= 0) adj_add(P(nd.b), w * nd.db);
        }
    }
    // the inputs' adjoints: the window holds nodes nin - S .. nin - 1, adjv the rest
    auto A = [&](int q) { return (S > 0 && q >= nin - S) ? win(wslot(q)) : adj_load(P(q)); };
    for (int k = 1; k <= mkx; ++k) {
        if (g.t0) g.t0[row(k)] = A(ADI_T0 * KS + k);
        if (g.qv0) g.qv0[row(k)] = A(ADI_QV0 * KS + k);
        if (g.u0) g.u0[row(k)] = A(ADI_U0 * KS + k);
        if (g.v0) g.v0[row(k)] = A(ADI_V0 * KS + k);
    }
    if (g.tke)
        for (int k = 0; k <= mkx; ++k) g.tke[row(k)] = A(ADI_TKE * KS + k);
    if (g.pblh) g.pblh[col] = 0.0;
    if (g.par)
        for (int j = 0; j < 3; ++j) g.par[(size_t)j * ldc + col] = A(ADI_PARAMS * KS + j);
}

// The window of a sweep without one (S = 0): never read, but the call must bind.
struct NoWindow {
    mutable double unused = 0.0;
    UWSHCU_HD double& operator()(int) const { return unused; }
};
UWSHCU_HD inline void ad_reverse(const TapeCtx& t, double* adjv, const int* nodes, size_t wld, int slot, int mkx,
                                 int ncnst, size_t ldc, int col, const AdCotangents& cot, const AdGradients& g) {
    ad_reverse(t, adjv, nodes, wld, slot, mkx, ncnst, ldc, col, cot, g, NoWindow{}, 0);
}

}  // namespace uwshcu
#endif
