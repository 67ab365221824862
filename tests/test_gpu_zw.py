"""GPU parity of the modified scheme compute_zwshcu_inv (uwshcu_zw_forward_device).

The HIP kernel (Column<double, double, true>, uwshcu_zw.hip) must equal the
portable-math oracle's compute_zwshcu_inv bit for bit: every output and every
integer diagnostic, for each option of the scheme, at the reference's chunk shape
and at a launch large enough for the two-wavefronts-per-SIMD kernel with deferred
second passes.  With every option off it must also equal the UW kernel.
"""
import numpy as np
import pytest

import oracle
from uwshcu_amd.generator import add_zw_inputs, make_chunk

pytestmark = pytest.mark.gpu

OPTIONS = [dict(), dict(ent=1), dict(cloud_sh=1), dict(pertt=1, pertw=1),
           dict(ent=1, cloud_sh=1, pertt=1, pertw=1, fwcrit=0.7, tpertmax=1.5, ufrcz=1.0)]


def _to_host(out, ncnst):
    import torch
    torch.cuda.synchronize()
    res = {k: v.cpu().numpy() for k, v in out.items()}
    res["diag"] = np.ascontiguousarray(res["diag"].T)
    res["trten"] = res["trten"][:ncnst]           # allocated with at least one constituent
    return res


def _gpu_zw(chunk, opts):
    from uwshcu_amd import _capi, device
    ctx = device.default_context()
    par, z = _capi.zw_default_params(**opts)
    out = ctx.compute_zwshcu_inv(device.chunk_to_device(chunk), z, dt=chunk.dt, params=par, diag=True)
    return _to_host(out, chunk.ncnst)


def _check(chunk, opts, label):
    print(f"\n[{label}] {chunk.ncol}x{chunk.pver}x{chunk.ncnst} options {opts}", flush=True)
    got = _gpu_zw(chunk, opts)
    consts = oracle.default_consts(rpen=15.0, rkm=10.0)
    ref, rdiag, rc, _ = oracle.run_zw(chunk, oracle.zw_defaults(**opts), consts, portable=True, nthreads=16)
    assert rc == 0
    np.testing.assert_array_equal(got["diag"], rdiag)
    differing = [k for k in ref if not np.array_equal(got[k], ref[k])]
    assert not differing, f"[{label}] not bit-identical to the oracle: {differing}"
    print(f"[{label}] bit-identical; convecting {(rdiag[:, 7] > 0).mean():.3f}", flush=True)


@pytest.mark.parametrize("opts", OPTIONS, ids=lambda o: "-".join(o) or "defaults")
def test_zw_weak_inversion_30_levels(opts):
    ch = add_zw_inputs(make_chunk(2048, 30, 3, seed=21, inversion_frac=1.0, inversion_dt=1.0), 21)
    _check(ch, opts, "weak inversion")


@pytest.mark.parametrize("opts", [OPTIONS[0], OPTIONS[-1]], ids=["defaults", "all"])
def test_zw_35_and_72_levels(opts):
    _check(add_zw_inputs(make_chunk(4096, 35, 3, seed=22), 22), opts, "35 levels")
    _check(add_zw_inputs(make_chunk(1024, 72, 10, seed=23, ptop=100.0, inversion_frac=0.5, inversion_dt=1.5), 23),
           opts, "72 levels, 10 constituents")


def test_zw_large_launch_with_deferred_second_passes():
    # at least 2 x 1024 x 64 columns: the occupancy-2 kernel and the deferred
    # second-pass launch
    ch = add_zw_inputs(make_chunk(140000, 35, 3, seed=24, inversion_frac=0.6, inversion_dt=1.5), 24)
    _check(ch, OPTIONS[-1], "large launch")


def test_zw_with_options_off_equals_the_uw_kernel():
    from uwshcu_amd import _capi, device
    ch = add_zw_inputs(make_chunk(4096, 35, 3, seed=25, inversion_frac=1.0, inversion_dt=1.0), 25)
    zw = _gpu_zw(ch, dict(lf=0))
    par, _ = _capi.zw_default_params()
    ctx = device.default_context()
    uw = _to_host(ctx.compute_uwshcu_inv(device.chunk_to_device(ch), dt=ch.dt, params=par, diag=True), ch.ncnst)
    for k in uw:
        np.testing.assert_array_equal(zw[k], uw[k], err_msg=k)


def test_zw_zero_columns_and_a_ragged_chunk():
    import torch
    from uwshcu_amd import _capi, device
    ch = add_zw_inputs(make_chunk(100, 30, 3, seed=26, inversion_frac=1.0, inversion_dt=1.0), 26)
    ctx = device.default_context()
    inp = device.chunk_to_device(ch)
    par, z = _capi.zw_default_params(ent=1, pertt=1)
    out = ctx.alloc_outputs(100, 30, 3, diag=True)
    for v in out.values():
        v.fill_(-7)
    cush = inp["cush"].clone()
    ctx.zw_forward_into(inp, out, cush, z, ncol=0, dt=ch.dt, params=par)
    ctx.check()
    assert all(bool((v == -7).all()) for v in out.values())
    # ncol = 77 of pcols = 100: the first 77 columns equal a 77-column run, the rest untouched
    ctx.zw_forward_into(inp, out, cush, z, ncol=77, dt=ch.dt, params=par)
    ctx.check()
    got = _to_host(out, 3)
    small = ch.copy()
    small.ncol = 77
    small.fields = {k: np.ascontiguousarray(v[..., :77]) for k, v in ch.fields.items()}
    ref, rdiag, rc, _ = oracle.run_zw(small, oracle.zw_defaults(ent=1, pertt=1),
                                      oracle.default_consts(rpen=15.0, rkm=10.0), portable=True)
    assert rc == 0
    for k in ref:
        if k == "cush":
            continue
        np.testing.assert_array_equal(got[k][..., :77], ref[k], err_msg=k)
        assert np.all(got[k][..., 77:] == -7), k
    np.testing.assert_array_equal(cush.cpu().numpy()[:77], ref["cush"])
    np.testing.assert_array_equal(got["diag"][:77], rdiag)
    torch.cuda.synchronize()


def test_zw_host_array_entry_point():
    """uwshcu_zw_forward_host: the entry point a Fortran driver binds in place of
    compute_zwshcu_inv, on host arrays with pcols > ncol."""
    import ctypes as C
    from uwshcu_amd import _capi, device
    full = add_zw_inputs(make_chunk(300, 30, 3, seed=27, inversion_frac=1.0, inversion_dt=1.0), 27)
    ncol = 257
    opts = dict(ent=1, pertt=1, pertw=1, cloud_sh=1)
    small = full.copy()
    small.ncol = ncol
    small.fields = {k: np.ascontiguousarray(v[..., :ncol]) for k, v in full.fields.items()}
    ref, rdiag, rc, _ = oracle.run_zw(small, oracle.zw_defaults(**opts), oracle.default_consts(rpen=15.0, rkm=10.0),
                                      portable=True)
    assert rc == 0
    f = full.fields
    n, Lv, nc = full.ncol, full.pver, full.ncnst
    outs = {k: np.full((Lv + 1, n), -7.0) for k in ("umf", "slflx", "qtflx", "flxprc1", "flxsnow1", "wu")}
    outs.update({k: np.full((Lv, n), -7.0) for k in ("qvten", "qlten", "qiten", "sten", "uten", "vten", "qrten",
                                                      "qsten", "evapc", "cufrc", "qcu", "qlu", "qiu", "qc")})
    outs["trten"] = np.full((nc, Lv, n), -7.0)
    outs.update({k: np.full(n, -7.0) for k in ("precip", "snow", "cbmf", "rliq", "cnt", "cnb")})
    cush = f["cush"].copy()
    diag = np.zeros((_capi.NDIAG, n), dtype=np.int32)
    pd = lambda a: a.ctypes.data_as(C.c_void_p)
    par, z = _capi.zw_default_params(**opts)
    ctx = device.default_context()
    arrays = {**f, **outs, "cush": cush}
    args = [ctx._h, C.byref(par), C.byref(z), n, ncol, Lv, nc, C.c_double(full.dt)]
    args += [pd(np.ascontiguousarray(arrays[k])) if k in f else pd(arrays[k]) for k in _capi.ZW_ARRAY_ARGS]
    bad = C.c_int(0)
    rc = _capi.lib().uwshcu_zw_forward_host(*args, pd(diag), C.byref(bad))
    assert rc == 0 and bad.value == -1
    outs["cush"] = cush
    for k in ref:
        np.testing.assert_array_equal(outs[k][..., :ncol], ref[k], err_msg=k)
        assert np.all(outs[k][..., ncol:] == (f["cush"][ncol:] if k == "cush" else -7.0)), k
    np.testing.assert_array_equal(np.ascontiguousarray(diag[:, :ncol].T), rdiag)
