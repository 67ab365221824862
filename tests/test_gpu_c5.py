"""The parameter-calibration workload: 4 194 304 columns x 72 levels x 10 constituents
(top 1 hPa), mixed-precision forward and FP64 forward + adjoint.

The chunk is 262 144 distinct columns repeated 16 times on the GPU (columns are
independent, so this changes no per-column work).  Checks:
* the mixed forward closes the column water budget in every column;
* the adjoint's forward outputs (the FP64 path) equal the portable-math oracle bit for
  bit on a sample of distinct columns, and are the same in every repeat of a column;
* the parameter gradients d/d(rpen, rle, rkm) reduced over all 4 194 304 columns on the
  device in FP64 equal the host's sum of the per-column contributions, and the
  per-column contributions are the same in every repeat of a column.
Inputs and outputs take about 140 GB of HBM; the adjoint's workspace is limited to
64 GiB (it runs in column batches)."""
import numpy as np
import pytest

import oracle
from uwshcu_amd.generator import make_chunk

pytestmark = pytest.mark.gpu

NCOL, TILE, PVER, NCNST = 4194304, 16, 72, 10
GRAV = 9.80616


def test_c5_forward_and_parameter_gradients():
    import torch
    from uwshcu_amd import _capi, device
    nd = NCOL // TILE
    ch = make_chunk(nd, PVER, NCNST, seed=50, inversion_frac=0.3, ptop=100.0)
    inp = device.chunk_to_device(ch, "cuda:0")
    inp = {k: v.repeat(*([1] * (v.dim() - 1)), TILE).contiguous() for k, v in inp.items()}
    torch.cuda.empty_cache()
    ctx = device.Context(0, workspace_limit=64 << 30)
    out = ctx.alloc_outputs(NCOL, PVER, NCNST, diag=True)
    cush = inp["cush"].clone()
    print(f"\n[c5] mixed forward over {NCOL}x{PVER}x{NCNST}", flush=True)
    ctx.forward_into(inp, out, cush, dt=ch.dt, precision=_capi.MIXED)
    rc, bad = ctx.status()
    assert rc == 0, (rc, bad)
    conv = float((out["diag"][7] > 0).float().mean())
    w = ((out["qvten"] + out["qlten"] + out["qiten"]) * inp["dp0"]).sum(0) / GRAV + \
        (out["precip"] + out["rliq"]) * 1000.0
    scale = out["qtflx"].abs().max(0).values + 1e-30
    worst = float((w.abs() / scale).max())
    print(f"[c5] convecting {conv:.3f}, water budget residual max {worst:.1e}, "
          f"recomputed in FP64 {ctx.mixed_recomputed()}", flush=True)
    assert worst < 1e-10
    # FP64 forward + adjoint of a loss on sten and qvten
    rng = np.random.default_rng(51)
    dev = torch.device("cuda:0")
    cot = {"sten": torch.from_numpy(rng.normal(0, 1e-3, (PVER, nd))).to(dev).repeat(1, TILE).contiguous(),
           "qvten": torch.from_numpy(rng.normal(0, 1e3, (PVER, nd))).to(dev).repeat(1, TILE).contiguous()}
    grads = {"par_col": torch.zeros((3, NCOL), dtype=torch.float64, device=dev),
             "par_sum": torch.zeros(3, dtype=torch.float64, device=dev)}
    cush.copy_(inp["cush"])
    print(f"[c5] FP64 forward + adjoint over {NCOL}x{PVER}x{NCNST}", flush=True)
    ctx.adjoint_into(inp, out, cush, cot, grads, dt=ch.dt)
    rc, bad = ctx.status()
    assert rc == 0, (rc, bad)
    pc = grads["par_col"].cpu().numpy()
    ps = grads["par_sum"].cpu().numpy()
    host = pc.sum(axis=1)
    print(f"[c5] d/d(rpen, rle, rkm) summed on the device {ps}, on the host {host}", flush=True)
    np.testing.assert_allclose(ps, host, rtol=1e-12, atol=0)
    assert np.abs(ps).min() > 0
    # repeats of a column give the same per-column gradients and outputs
    pcr = pc.reshape(3, TILE, nd)
    assert np.array_equal(pcr, np.broadcast_to(pcr[:, :1], pcr.shape))
    sten = out["sten"].reshape(PVER, TILE, nd)
    assert torch.equal(sten, sten[:, :1].expand_as(sten))
    # the FP64 outputs against the oracle on a sample of distinct columns
    idx = np.sort(rng.choice(nd, 512, replace=False))
    sub = oracle.column_chunk(ch, 0)
    sub.ncol = len(idx)
    sub.fields = {k: np.ascontiguousarray(v[..., idx]) for k, v in ch.fields.items()}
    ref, rdiag, orc, _ = oracle.run(sub, nthreads=16, portable=True)
    assert orc == 0
    ti = torch.from_numpy(idx).to(dev)
    for k in ref:
        if k == "cush":
            got = cush.index_select(0, ti).cpu().numpy()
        else:
            got = out[k].index_select(out[k].dim() - 1, ti).cpu().numpy()
        np.testing.assert_array_equal(got, ref[k], err_msg=k)
