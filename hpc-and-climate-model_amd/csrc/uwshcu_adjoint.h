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

#include "uwshcu_column.h"

namespace uwshcu {

UWSHCU_HD inline void ad_begin(TapeCtx& t, TapeNode* nodes, int cap, int nin, int stride = 1) {
    t.nodes = nodes;
    t.cap = cap;
    t.stride = stride;
    t.overflow = nin > cap ? 1 : 0;
    t.n = nin < cap ? nin : cap;
    if (!nodes) return;
    for (int q = 0; q < t.n; ++q) {
        TapeNode nd;
        nd.da = 0.0;
        nd.db = 0.0;
        nd.a = -1;
        nd.b = -1;
        nd.adj = 0.0;
        ((UWSHCU_GMEM TapeNode*)nodes)[tape_slot(t, q)] = nd;
    }
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
// propagate back through the tape, write the input gradients.  Node q's adjoint is the
// adj field of its tape node (zero when recorded).
//
// With a window of S >= kSweepBatch entries (win(q % S), and a spare entry win(S)), the
// adjoints of recent nodes are kept in it: 82 % of the operands a node refers to lie
// within 64 nodes of it (35-level columns).  The sweep works kSweepBatch (K) nodes at a
// time; when a batch starts at node q0 the window holds the adjoints of nodes
// q0 - S + 1 .. q0.  A contribution to one of those goes to the window, one to an
// older node to the tape (a far contribution).  When the batch ends, its K nodes leave
// the window and nodes q0 - S - K + 1 .. q0 - S take their entries, carrying what far
// contributions they received.  Every adjoint still receives its contributions in
// decreasing node order, so the result is the same, bit for bit, for every S, and as
// without a window.
//
// The memory operations of a batch are issued together at its end: its far
// contributions (no-return atomic adds, in step order), then the loads of the adjoints
// that enter the window and of the next batch's tape nodes (indices clamped into the
// tape), so that one memory latency is paid per batch and no step waits on memory.
// (A far contribution can wait to the batch's end: its node enters the window at a
// later batch's end.)  Within a step, the node's
// adjoint and its operands' window entries are read together and written back
// unconditionally (an operand outside the window uses the spare entry), so that a
// step waits for the window once.
constexpr int kSweepBatch = 16;

template <class Win>
UWSHCU_HD inline void ad_reverse(const TapeCtx& t, const int* nodes, size_t wld, int slot, int mkx, int ncnst,
                                 size_t ldc, int col, const AdCotangents& cot, const AdGradients& g, Win win, int S) {
    const int KS = mkx + 1;
    // the header's fields in registers: the sweep's stores could alias the header
    const TapeCtx th = t;
    UWSHCU_GMEM TapeNode* const tp = (UWSHCU_GMEM TapeNode*)th.nodes;
    auto P = [&](int q) { return &tp[tape_slot(th, q)].adj; };
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
    const int n = th.n;
    if (S > 0) {
        constexpr int K = kSweepBatch;
        double da[K], db[K];
        int ia[K], ib[K];
        auto load_batch = [&](int q0) {
#pragma unroll
            for (int u = 0; u < K; ++u) {
                const UWSHCU_GMEM TapeNode* p = &tp[tape_slot(th, q0 - u > nin ? q0 - u : nin)];
                da[u] = p->da;
                db[u] = p->db;
                ia[u] = p->a;
                ib[u] = p->b;
            }
        };
        for (int q = n - 1; q >= 0 && q > n - 1 - S; --q) win(q % S) = adj_load(P(q));
        if (n > nin) load_batch(n - 1);
        for (int q0 = n - 1; q0 >= nin; q0 -= K) {
            const int lo = q0 - S;     // the window holds the nodes above lo
            int fx[2 * K];             // far contributions of the batch: node (or -1), value
            double fc[2 * K];
#pragma unroll
            for (int u = 0; u < K; ++u) {
                // past the last node (q < nin) a step still runs, on the spare entry, so that
                // every path consumes the batch's loads
                const int q = q0 - u;
                const bool act = q >= nin;
                const int a = ia[u], b = ib[u];
                const bool nearA = act && a >= 0 && a > lo, nearB = act && b >= 0 && b > lo;
                const int sa = nearA ? a % S : S, sb = nearB ? b % S : S;
                const double w = win(act ? q % S : S), va = win(sa), vb = win(sb);
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
            // the batch's far contributions, in step order, then the batch's nodes leave the
            // window and nodes lo - u enter, into their entries
#pragma unroll
            for (int v = 0; v < 2 * K; ++v)
                if (fx[v] >= 0) adj_add(P(fx[v]), fc[v]);
            double enter[K];
#pragma unroll
            for (int u = 0; u < K; ++u) enter[u] = adj_load(P(lo - u > 0 ? lo - u : 0));
            load_batch(q0 - K);
            // written unconditionally (to the spare entry when nothing enters), so that no
            // load is left pending on some paths into the next batch
#pragma unroll
            for (int u = 0; u < K; ++u) win(q0 - u >= nin && lo - u >= 0 ? (q0 - u) % S : S) = enter[u];
        }
    } else {
        for (int q = n - 1; q >= nin; --q) {
            const double w = adj_load(P(q));
            if (w == 0.0) continue;
            const UWSHCU_GMEM TapeNode& nd = tp[tape_slot(th, q)];
            if (nd.a >= 0) adj_add(P(nd.a), w * nd.da);
            if (nd.b >= 0) adj_add(P(nd.b), w * nd.db);
        }
    }
    // the inputs' adjoints: the window holds nodes nin - S .. nin - 1, the tape the rest
    auto A = [&](int q) { return (S > 0 && q >= nin - S) ? win(q % S) : adj_load(P(q)); };
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

struct NoWindow {
    UWSHCU_HD double& operator()(int) const { return *(double*)nullptr; }
};
UWSHCU_HD inline void ad_reverse(const TapeCtx& t, const int* nodes, size_t wld, int slot, int mkx, int ncnst,
                                 size_t ldc, int col, const AdCotangents& cot, const AdGradients& g) {
    ad_reverse(t, nodes, wld, slot, mkx, ncnst, ldc, col, cot, g, NoWindow{}, 0);
}

}  // namespace uwshcu
#endif
