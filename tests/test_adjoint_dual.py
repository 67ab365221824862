"""The adjoint against forward-mode derivatives of the oracle (CPU only).

oracle.jvp runs the portable-math oracle over dual numbers (oracle/orc_generic.cpp):
the same restatement, values bit-identical to the FP64 oracle, each value carrying its
derivative along a chosen direction in (t0, qv0, u0, v0, tke) and in the tunables
(rpen, rle, rkm).  For a cotangent c of the outputs and any direction v, the adjoint's
gradient g = J^T c must satisfy <g, v> = <c, J v> column by column.  The adjoint here
is the product's column code compiled for the host (tests/hostemu), the same code the
GPU adjoint kernel runs; tests/test_gpu_adjoint.py checks the GPU against the same
identity at 65 536 columns.

The tape records the saturation adjustment's derivative at its converged root (the
implicit-function derivative) where the dual numbers differentiate the reference's ten
Newton steps; the two agree to the steps' convergence: 1e-13 typically, 1e-7 in the rare
column whose adjustment has not converged in ten steps.
"""
import numpy as np
import pytest

import hostemu
import oracle
from uwshcu_amd.generator import make_chunk


def _dot_test(ch, seed):
    rng = np.random.default_rng(seed)
    L, n = ch.pver, ch.ncol
    cot = {"sten": rng.normal(0, 1, (L, n)), "qvten": rng.normal(0, 1e3, (L, n)),
           "umf": rng.normal(0, 1, (L + 1, n)), "precip": rng.normal(0, 1e3, n),
           "trten": rng.normal(0, 1e3, (ch.ncnst, L, n))}
    _, g, _, _ = hostemu.adjoint(ch, cot)
    v = {"t0": rng.normal(0, 1, (L, n)), "qv0": rng.normal(0, 1e-4, (L, n)), "u0": rng.normal(0, 1, (L, n)),
         "v0": rng.normal(0, 1, (L, n)), "tke": rng.normal(0, 0.1, (L + 1, n))}
    dpar = (1.0, 0.01, 1.0)
    out, tout, diag = oracle.jvp(ch, v, dpar=dpar)
    ref, rdiag, _, _ = oracle.run(ch, nthreads=8, portable=True)
    for k in ref:
        if k != "cush":
            np.testing.assert_array_equal(out[k], ref[k], err_msg=k)     # the dual values are the oracle's
    np.testing.assert_array_equal(diag, rdiag)
    lhs = sum((g[k] * v[k]).sum(axis=0) for k in v) + (g["par"] * np.array(dpar)[:, None]).sum(axis=0)
    rhs = sum((cot[k] * tout[k]).reshape(-1, n).sum(axis=0) for k in cot)
    conv = diag[:, 7] > 0
    rel = np.abs(lhs - rhs) / np.maximum(np.abs(lhs) + np.abs(rhs), 1e-300)
    return rel[conv], lhs[~conv], rhs[~conv]


def _kink(l0, lp, lm, lp2, lm2, h):
    """|(fw - bw) - 2 ((fw - fw/2) + (bw/2 - bw))| of the one-sided differences at steps
    h and h/2: O(h^2) for a smooth function, the slope jump for a kink at the point,
    where the centred differences at both steps agree with each other but not with
    either one-sided derivative."""
    fw, bw = (lp - l0) / h, (l0 - lm) / h
    fw2, bw2 = (lp2 - l0) / (h / 2), (l0 - lm2) / (h / 2)
    return np.abs((fw - bw) - 2.0 * ((fw - fw2) + (bw2 - bw)))


@pytest.mark.parametrize("ncol,pver,ncnst,seed", [(256, 35, 3, 7), (64, 72, 10, 8)])
def test_adjoint_equals_dual_number_derivative(ncol, pver, ncnst, seed):
    ch = make_chunk(ncol, pver, ncnst, seed=seed, inversion_frac=0.3, ptop=100.0 if pver > 60 else 200.0)
    rel, lhs0, rhs0 = _dot_test(ch, seed)
    assert len(rel) > ncol // 3
    assert np.median(rel) <= 1e-12 and np.quantile(rel, 0.95) <= 1e-9 and rel.max() <= 1e-6, \
        (np.median(rel), np.quantile(rel, 0.95), rel.max())
    assert not lhs0.any() and not rhs0.any()      # columns that do not convect have no derivative


def test_dual_derivative_matches_centred_differences():
    """The dual numbers themselves against centred differences of the FP64 oracle,
    along one direction, on the columns where the differences at two step sizes agree
    (no branch switch or table knot inside the step)."""
    ch = make_chunk(192, 35, 3, seed=9, inversion_frac=0.3)
    rng = np.random.default_rng(2)
    v = {"t0": rng.normal(0, 1, ch["t0"].shape), "qv0": rng.normal(0, 1e-4, ch["qv0"].shape)}
    _, tout, diag = oracle.jvp(ch, v, dpar=(1.0, 0.0, 0.0))

    def at(s):
        c2 = ch.copy()
        for k in v:
            c2.fields[k] = ch[k] + s * v[k]
        return oracle.run(c2, nthreads=8, portable=True, consts=oracle.default_consts(rpen=10.0 + s))[0]

    h = 1e-6
    base, up, dn, up2, dn2 = at(0.0), at(h), at(-h), at(h / 2), at(-h / 2)
    n_ok = 0
    for k in ("sten", "qvten"):
        c1 = (up[k] - dn[k]) / (2 * h)
        c2 = (up2[k] - dn2[k]) / h
        scale = np.abs(tout[k]).max(axis=0) + 1e-300
        # away from branch switches: the centred differences at two steps agree, and the
        # one-sided ones are consistent with a smooth function
        smooth = ((np.abs(c1 - c2).max(axis=0) <= 1e-7 * scale)
                  & (_kink(base[k], up[k], dn[k], up2[k], dn2[k], h).max(axis=0) <= 1e-6 * scale)
                  & (diag[:, 7] > 0))
        err = np.abs(c2 - tout[k]).max(axis=0) / scale
        assert np.all(err[smooth] <= 1e-6), (k, err[smooth].max())
        n_ok += smooth.sum()
    assert n_ok > 50
