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
        nd.a = -1;
        nd.b = -1;
        nd.da = 0.0;
        nd.db = 0.0;
        nodes[tape_slot(t, q)] = nd;
    }
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
// propagate back through the tape, write the input gradients.  adj is laid out like
// the tape's nodes (adj[tape_slot(t, q)] is the adjoint of node q).
//
// With a window (win != nullptr, W entries, win[q % W]) the adjoints of the W most
// recent nodes are kept in it: 82 % of the operands a node refers to lie within 64
// nodes of it (35-level columns), so most of the sweep's read-modify-writes stay out
// of HBM.  A node's adjoint enters the window when the sweep comes within W of it,
// carrying what farther nodes already added in adj; later contributions go to the
// window.  Every adjoint still receives its contributions in decreasing node order,
// so the result is the same, bit for bit, as without the window.
//
// The windowed sweep works kSweepAhead nodes at a time.  It first issues, together,
// the loads of the batch's tape nodes and of the adjoints that enter the window during
// the batch (adj of node q - W for each node q), so that their latencies overlap
// instead of adding up node by node.  A far contribution the batch itself makes to one
// of those entering adjoints is added to the early-loaded copy too, before the copy
// enters the window, so every adjoint still receives its contributions in the same
// order and the result stays bit for bit the same.
template <class Win>
UWSHCU_HD inline void ad_reverse(const TapeCtx& t, double* adj, const int* nodes, size_t wld, int slot, int mkx,
                                 int ncnst, size_t ldc, int col, const AdCotangents& cot, const AdGradients& g,
                                 Win win, int W) {
    const int KS = mkx + 1;
    for (int q = 0; q < t.n; ++q) adj[tape_slot(t, q)] = 0.0;
    auto A = [&](int q) -> double& { return adj[tape_slot(t, q)]; };
    auto node = [&](int f, int k) { return nodes[((size_t)f * KS + k) * wld + slot]; };
    auto row = [&](int k) { return (size_t)(mkx - k) * ldc + col; };
    for (int f = 0; f < ADF_TR; ++f) {
        if (!cot.f[f]) continue;
        for (int k = (f <= ADF_WU ? 0 : 1); k <= mkx; ++k) {
            const int nd = node(f, k);
            if (nd >= 0) A(nd) += cot.f[f][row(k)];
        }
    }
    if (cot.trten)
        for (int m = 1; m <= ncnst; ++m)
            for (int k = 1; k <= mkx; ++k) {
                const int nd = node(ADF_TR + m - 1, k);
                if (nd >= 0) A(nd) += cot.trten[((size_t)(m - 1) * mkx + (mkx - k)) * ldc + col];
            }
    for (int q = 0; q < ADC_COUNT; ++q) {
        if (!cot.scal[q]) continue;
        const int nd = node(ADF_TR + ncnst + q, 0);
        if (nd >= 0) A(nd) += cot.scal[q][col];
    }
    const int nin = ad_inputs(mkx);
    if (W > 0) {
        constexpr int kSweepAhead = 8;
        for (int q = t.n - 1; q >= 0 && q >= t.n - W; --q) win(q % W) = A(q);
        for (int q0 = t.n - 1; q0 >= nin; q0 -= kSweepAhead) {
            TapeNode buf[kSweepAhead];
            double enter[kSweepAhead];   // enter[u]: adj of node q0 - u - W, entering at step u
#pragma unroll
            for (int u = 0; u < kSweepAhead; ++u) {
                if (q0 - u >= nin) buf[u] = t.nodes[tape_slot(t, q0 - u)];
                enter[u] = (q0 - u >= nin && q0 - u - W >= 0) ? A(q0 - u - W) : 0.0;
            }
            // a far contribution c to node x made at step u: x enters at step q0 - W - x
            auto far = [&](int x, double c, int u) {
                A(x) += c;
                const int r = q0 - W - x;
#pragma unroll
                for (int v = 0; v < kSweepAhead; ++v)
                    if (v == r && v >= u) enter[v] += c;
            };
#pragma unroll
            for (int u = 0; u < kSweepAhead; ++u) {
                const int q = q0 - u;
                if (q < nin) break;
                const double w = win(q % W);
                if (w != 0.0) {
                    const TapeNode& nd = buf[u];
                    if (nd.a >= 0) {
                        if (q - nd.a < W) win(nd.a % W) += w * nd.da;
                        else far(nd.a, w * nd.da, u);
                    }
                    if (nd.b >= 0) {
                        if (q - nd.b < W) win(nd.b % W) += w * nd.db;
                        else far(nd.b, w * nd.db, u);
                    }
                }
                if (q - W >= 0) win(q % W) = enter[u];      // node q - W enters the window
            }
        }
        for (int q = nin - 1; q >= 0 && q >= nin - W; --q) A(q) = win(q % W);
    } else {
        for (int q = t.n - 1; q >= nin; --q) {
            const double w = A(q);
            if (w == 0.0) continue;
            const TapeNode nd = t.nodes[tape_slot(t, q)];
            if (nd.a >= 0) A(nd.a) += w * nd.da;
            if (nd.b >= 0) A(nd.b) += w * nd.db;
        }
    }
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
UWSHCU_HD inline void ad_reverse(const TapeCtx& t, double* adj, const int* nodes, size_t wld, int slot, int mkx,
                                 int ncnst, size_t ldc, int col, const AdCotangents& cot, const AdGradients& g) {
    ad_reverse(t, adj, nodes, wld, slot, mkx, ncnst, ldc, col, cot, g, NoWindow{}, 0);
}

}  // namespace uwshcu
#endif
