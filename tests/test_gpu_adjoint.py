"""GPU adjoint: the device tape and sweep against the same code on the host, the
torch.autograd binding, batching and the tape-capacity error."""
import numpy as np
import pytest

import hostemu
import oracle
from uwshcu_amd.generator import make_chunk

pytestmark = pytest.mark.gpu

COT = ("sten", "qvten", "umf", "precip", "qlten")


def _cot(ch, seed):
    rng = np.random.default_rng(seed)
    ref, _, _, _ = oracle.run(ch, nthreads=8, portable=True)
    return {k: rng.normal(0.0, 1.0, np.shape(ref[k])) for k in COT}


def _gpu_adjoint(ch, cot, workspace_limit=None, cap=None):
    import torch
    from uwshcu_amd import device
    own = bool(workspace_limit or cap)
    ctx = device.Context(0, workspace_limit=workspace_limit) if own else device.default_context()
    if cap:
        ctx.set_tape_capacity(cap)
    inp = device.chunk_to_device(ch, "cuda:0")
    pver, n = ch.pver, ch.ncol
    out = ctx.alloc_outputs(n, pver, ch.ncnst, diag=True)
    cush = inp["cush"].clone()
    dev = torch.device("cuda:0")
    grads = {k: torch.zeros((pver, n), dtype=torch.float64, device=dev) for k in ("t0", "qv0", "u0", "v0")}
    grads["tke"] = torch.zeros((pver + 1, n), dtype=torch.float64, device=dev)
    grads["pblh"] = torch.ones(n, dtype=torch.float64, device=dev)
    grads["par_col"] = torch.zeros((3, n), dtype=torch.float64, device=dev)
    grads["par_sum"] = torch.zeros(3, dtype=torch.float64, device=dev)
    cd = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in cot.items()}
    ctx.adjoint_into(inp, out, cush, cd, grads, dt=ch.dt)
    rc, bad = ctx.status()
    res = {k: v.cpu().numpy() for k, v in out.items()}
    res["cush"] = cush.cpu().numpy()
    g = {k: v.cpu().numpy() for k, v in grads.items()}
    if own:
        ctx.close()
    return rc, res, g


@pytest.mark.parametrize("ncol,pver,ncnst,recipe", [(1024, 35, 3, "pbl"), (128, 72, 10, "pbl"), (1024, 30, 3, "baseline")])
def test_device_adjoint_equals_host_adjoint(ncol, pver, ncnst, recipe):
    ch = make_chunk(ncol, pver, ncnst, seed=60 + pver, ptop=100.0 if pver > 60 else 200.0, recipe=recipe)
    cot = _cot(ch, 1)
    print(f"\n[adjoint {ncol}x{pver}x{ncnst} {recipe}] device run", flush=True)
    rc, out, g = _gpu_adjoint(ch, cot)
    assert rc == 0
    ref, rdiag, _, _ = oracle.run(ch, nthreads=8, portable=True)
    np.testing.assert_array_equal(out["diag"].T, rdiag)
    for k in ref:
        np.testing.assert_array_equal(out[k], ref[k], err_msg=k)
    _, hg, tlen, _ = hostemu.adjoint(ch, cot)
    print(f"[adjoint {ncol}x{pver}x{ncnst}] tape nodes max {tlen.max()} median {int(np.median(tlen))}", flush=True)
    for k in ("t0", "qv0", "u0", "v0", "tke", "pblh"):
        np.testing.assert_array_equal(g[k], hg[k], err_msg=k)
    np.testing.assert_array_equal(g["par_col"], hg["par"])
    np.testing.assert_allclose(g["par_sum"], hg["par"].sum(axis=1), rtol=1e-12, atol=0)


def test_batched_adjoint_gives_the_same_gradients():
    from uwshcu_amd import _capi  # noqa: F401
    ch = make_chunk(700, 35, 3, seed=62)
    cot = _cot(ch, 2)
    _, _, g1 = _gpu_adjoint(ch, cot)
    _, _, g2 = _gpu_adjoint(ch, cot, workspace_limit=400 << 20)   # a few hundred columns per batch
    for k in g1:
        np.testing.assert_array_equal(g1[k], g2[k], err_msg=k)


def _adjoint_on(ctx, c, cc):
    """Gradients of one chunk's adjoint on the context `ctx`."""
    import torch
    from uwshcu_amd import device
    dev = torch.device("cuda:0")
    inp = device.chunk_to_device(c, "cuda:0")
    out = ctx.alloc_outputs(c.ncol, c.pver, c.ncnst, diag=True)
    grads = {k: torch.zeros((c.pver, c.ncol), dtype=torch.float64, device=dev) for k in ("t0", "qv0", "u0", "v0")}
    grads["tke"] = torch.zeros((c.pver + 1, c.ncol), dtype=torch.float64, device=dev)
    grads["pblh"] = torch.ones(c.ncol, dtype=torch.float64, device=dev)
    grads["par_col"] = torch.zeros((3, c.ncol), dtype=torch.float64, device=dev)
    grads["par_sum"] = torch.zeros(3, dtype=torch.float64, device=dev)
    cd = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in cc.items()}
    print(f"\n[adjoint on one context] {c.ncol}x{c.pver}x{c.ncnst}", flush=True)
    ctx.adjoint_into(inp, out, inp["cush"].clone(), cd, grads, dt=c.dt)
    rc, _ = ctx.status()
    assert rc == 0
    return {k: v.cpu().numpy() for k, v in grads.items()}


def test_adjoint_records_again_when_the_tape_pool_runs_out():
    """The recording takes tape pages from a pool sized from the previous call's mean
    tape.  After a chunk that hardly convects, a convecting chunk under a small
    workspace limit overruns the pool; its columns, which counted their lengths, are
    recorded again in parts that fit, with the same gradients."""
    from uwshcu_amd import device
    quiet = make_chunk(512, 35, 3, seed=64, inversion_frac=0.95)
    ch = make_chunk(700, 35, 3, seed=62, inversion_frac=0.1)
    cot = _cot(ch, 2)
    _, _, g1 = _gpu_adjoint(ch, cot)
    ctx = device.Context(0, workspace_limit=600 << 20)
    try:
        _adjoint_on(ctx, quiet, _cot(quiet, 4))
        g = _adjoint_on(ctx, ch, cot)
        for k in g1:
            np.testing.assert_array_equal(g[k], g1[k], err_msg=k)
    finally:
        ctx.close()


def test_adjoint_pool_estimate_follows_the_chunk_shape():
    """One context takes chunks of 35 levels and 3 constituents, then 72 levels and 10,
    then 35 and 3 again: the pool's estimate, learnt on one shape, is scaled to the
    next one's levels and constituents, and every chunk's gradients equal those of a
    fresh context."""
    from uwshcu_amd import device
    small = make_chunk(600, 35, 3, seed=66, inversion_frac=0.1)
    big = make_chunk(160, 72, 10, seed=67, ptop=100.0, inversion_frac=0.1)
    cs, cb = _cot(small, 5), _cot(big, 6)
    fresh = {}
    for name, c, cc in (("small", small, cs), ("big", big, cb)):
        _, _, fresh[name] = _gpu_adjoint(c, cc, workspace_limit=2 << 30)
    ctx = device.Context(0, workspace_limit=2 << 30)
    try:
        for name, c, cc in (("small", small, cs), ("big", big, cb), ("small", small, cs)):
            g = _adjoint_on(ctx, c, cc)
            for k in g:
                np.testing.assert_array_equal(g[k], fresh[name][k], err_msg=f"{name} {k}")
    finally:
        ctx.close()


def test_atomic_fp64_add_rounds_like_the_addition():
    """The sweep adds far contributions with a no-return FP64 atomic add (adj_add).
    Its result must be the bits of x + y: normal operands of any relative size,
    operands a half ulp apart, subnormal results, and exact cancellations."""
    import ctypes as C
    import os
    L = C.CDLL(os.path.join(os.path.dirname(__file__), "hostemu", "libmathprobe.so"))
    rng = np.random.default_rng(21)
    n = 1 << 20
    e1, e2 = rng.integers(-1074, 60, n), rng.integers(-1074, 60, n)
    x = np.ldexp(rng.uniform(-1, 1, n), e1)
    y = np.ldexp(rng.uniform(-1, 1, n), e2)
    k = rng.integers(0, 4, n)
    half = k == 0                          # y near half an ulp of x
    y[half] = np.ldexp(rng.uniform(-1, 1, half.sum()), np.maximum(-1074, np.frexp(x[half])[1] - 54))
    sub = k == 1                           # subnormal sums
    x[sub] = np.ldexp(rng.uniform(-1, 1, sub.sum()), -1060)
    y[sub] = np.ldexp(rng.uniform(-1, 1, sub.sum()), -1070)
    y[k == 2] = -x[k == 2]                 # exact cancellation
    got = np.zeros(n)
    pd = lambda a: a.ctypes.data_as(C.c_void_p)
    assert L.probe_atomic_add_gpu(n, pd(x), pd(y), pd(got)) == 0
    want = x + y
    diff = np.flatnonzero(got.view(np.int64) != want.view(np.int64))
    assert (np.abs(want[sub]) < np.finfo(float).tiny).sum() > 1000
    assert diff.size == 0, (diff.size, x[diff[:3]], y[diff[:3]])


def test_tape_capacity_overflow_is_an_error_not_a_fault():
    ch = make_chunk(256, 35, 3, seed=63)
    rc, _, _ = _gpu_adjoint(ch, _cot(ch, 3), cap=4000)
    assert rc == -14


def test_autograd_backward_matches_the_adjoint():
    import torch
    from uwshcu_amd import device
    from uwshcu_amd.autograd import uwshcu
    ch = make_chunk(512, 35, 3, seed=64)
    dev = torch.device("cuda:0")
    inp = device.chunk_to_device(ch, "cuda:0")
    for k in ("t0", "qv0", "u0", "v0", "tke", "pblh"):
        inp[k].requires_grad_(True)
    params = torch.tensor([10.0, 0.1, 14.0], dtype=torch.float64, requires_grad=True)
    rng = np.random.default_rng(4)
    tau_t = torch.from_numpy(rng.normal(0, 1e-5, (35, 512))).to(dev)
    tau_q = torch.from_numpy(rng.normal(0, 1e-8, (35, 512))).to(dev)
    out = uwshcu(inp, params, dt=ch.dt)
    loss = ((out["sten"] / 1.00464e3 - tau_t) ** 2).sum() + ((out["qvten"] - tau_q) ** 2).sum()
    loss.backward()
    cot = {"sten": (2 * (out["sten"] / 1.00464e3 - tau_t) / 1.00464e3).detach().cpu().numpy(),
           "qvten": (2 * (out["qvten"] - tau_q)).detach().cpu().numpy()}
    _, hg, _, _ = hostemu.adjoint(ch, cot)
    for k in ("t0", "qv0", "u0", "v0", "tke"):
        np.testing.assert_allclose(inp[k].grad.cpu().numpy(), hg[k], rtol=1e-12, atol=1e-300, err_msg=k)
    np.testing.assert_allclose(params.grad.numpy(), hg["par"].sum(axis=1), rtol=1e-10)
    assert not inp["pblh"].grad.any()


def test_autograd_rejects_inputs_changed_in_place():
    """Inputs are kept through save_for_backward: an in-place change between forward and
    backward raises autograd's version error instead of giving gradients of other values."""
    import torch
    from uwshcu_amd import device
    from uwshcu_amd.autograd import uwshcu
    ch = make_chunk(64, 35, 3, seed=65)
    for name in ("t0", "p0"):
        inp = device.chunk_to_device(ch, "cuda:0")
        inp["t0"].requires_grad_(True)
        params = torch.tensor([10.0, 0.1, 14.0], dtype=torch.float64, requires_grad=True)
        out = uwshcu(inp, params, dt=ch.dt)
        loss = out["sten"].sum()
        with torch.no_grad():
            inp[name].add_(0.5)
        with pytest.raises(RuntimeError, match="inplace"):
            loss.backward()


def test_c4_gradients_at_65536_columns():
    """The 65 536 x 35 x 3 calibration workload: the loss sum (sten/cp - tau_T)^2 +
    (qvten - tau_q)^2 with tau_T ~ N(0, 1e-5) K/s and tau_q ~ N(0, 1e-8) 1/s, its
    gradient by the GPU adjoint with respect to t0, qv0, u0, v0, tke and (rpen, rle, rkm).
    On 512 sampled columns, along a random direction v: <grad, v> equals <cot, J v> from
    the dual-number oracle (median 1e-12, 95th percentile 1e-9), and equals centred
    differences of the FP64 oracle's loss to 1e-6 on the columns away from branch
    switches (differences at two steps agree, one-sided ones are consistent with a smooth loss)."""
    import torch
    from uwshcu_amd import device
    cp = 1.00464e3
    n, L = 65536, 35
    ch = make_chunk(n, L, 3, seed=70, inversion_frac=0.3)
    rng = np.random.default_rng(71)
    tau_t, tau_q = rng.normal(0, 1e-5, (L, n)), rng.normal(0, 1e-8, (L, n))
    ctx = device.default_context()
    inp = device.chunk_to_device(ch, "cuda:0")
    out = ctx.compute_uwshcu_inv(inp, dt=ch.dt)
    sten, qvten = out["sten"].cpu().numpy(), out["qvten"].cpu().numpy()
    cot = {"sten": 2.0 * (sten / cp - tau_t) / cp, "qvten": 2.0 * (qvten - tau_q)}
    print(f"\n[c4] GPU adjoint over {n}x{L}x3", flush=True)
    rc, _, g = _gpu_adjoint(ch, cot)
    assert rc == 0
    idx = np.sort(rng.choice(n, 512, replace=False))
    sub = oracle.column_chunk(ch, 0)
    sub.ncol = len(idx)
    sub.fields = {k: np.ascontiguousarray(v[..., idx]) for k, v in ch.fields.items()}
    m = len(idx)
    v = {"t0": rng.normal(0, 1, (L, m)), "qv0": rng.normal(0, 1e-4, (L, m)), "u0": rng.normal(0, 1, (L, m)),
         "v0": rng.normal(0, 1, (L, m)), "tke": rng.normal(0, 0.1, (L + 1, m))}
    dpar = np.array([1.0, 0.01, 1.0])
    lhs = sum((g[k][:, idx] * v[k]).sum(axis=0) for k in v) + (g["par_col"][:, idx] * dpar[:, None]).sum(axis=0)
    _, tout, diag = oracle.jvp(sub, v, dpar=tuple(dpar))
    subcot = {k: c[:, idx] for k, c in cot.items()}
    rhs = sum((subcot[k] * tout[k]).sum(axis=0) for k in subcot)
    conv = diag[:, 7] > 0
    rel = np.abs(lhs - rhs) / np.maximum(np.abs(lhs) + np.abs(rhs), 1e-300)
    print(f"[c4] adjoint vs dual numbers on {conv.sum()} convecting columns: median {np.median(rel[conv]):.1e} "
          f"p95 {np.quantile(rel[conv], 0.95):.1e} max {rel[conv].max():.1e}", flush=True)
    assert np.median(rel[conv]) <= 1e-12 and np.quantile(rel[conv], 0.95) <= 1e-9
    assert not lhs[~conv].any()

    def loss(s):
        c2 = sub.copy()
        for k in v:
            c2.fields[k] = sub[k] + s * v[k]
        r = oracle.run(c2, nthreads=16, portable=True,
                       consts=oracle.default_consts(rpen=10.0 + s * dpar[0], rle=0.1 + s * dpar[1],
                                                    rkm=14.0 + s * dpar[2]))[0]
        return (((r["sten"] / cp - tau_t[:, idx]) ** 2).sum(axis=0) + ((r["qvten"] - tau_q[:, idx]) ** 2).sum(axis=0))

    h = 1e-6
    l0, lp, lm, lp2, lm2 = loss(0.0), loss(h), loss(-h), loss(h / 2), loss(-h / 2)
    c1, c2 = (lp - lm) / (2 * h), (lp2 - lm2) / h
    # one-sided differences at h and h/2: fw - bw = 2 ((fw - fw/2) + (bw/2 - bw)) + O(h^2)
    # for a smooth loss; a kink at the point breaks it by its slope jump (the centred
    # differences at both steps then agree with each other, not with the derivative)
    fw, bw, fw2, bw2 = (lp - l0) / h, (l0 - lm) / h, (lp2 - l0) / (h / 2), (l0 - lm2) / (h / 2)
    kink = np.abs((fw - bw) - 2.0 * ((fw - fw2) + (bw2 - bw)))
    scale = np.abs(lhs) + 1e-300
    smooth = conv & (np.abs(c1 - c2) <= 1e-7 * scale) & (kink <= 1e-6 * scale)
    err = np.abs(lhs - c2) / scale
    print(f"[c4] adjoint vs centred differences on {smooth.sum()} smooth columns: median {np.median(err[smooth]):.1e} "
          f"max {err[smooth].max():.1e}", flush=True)
    assert smooth.sum() > 100
    assert np.all(err[smooth] <= 1e-6)
