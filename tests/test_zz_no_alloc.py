"""uwshcu_forward_device_noalloc: the FP64 forward with its workspace in the library's
static device scratch, called through the C ABI.

Two calls in a row on one context (the second with a smaller chunk, so the scratch
still holds the first call's contents) must equal uwshcu_forward_device bit for bit on
every output and diagnostic; the calls must leave the device's free memory unchanged;
a chunk larger than the scratch is refused with UWSHCU_ERR_SCRATCH before anything is
launched."""
import ctypes as C

import numpy as np
import pytest

from uwshcu_amd import _capi
from uwshcu_amd.generator import make_chunk

pytestmark = pytest.mark.gpu

SCRATCH_ERR = -15


def _args(ctx_h, ch, inp, out, cush, stream, noalloc):
    import torch  # noqa: F401
    pver, pcols = ch.pver, ch.ncol
    ptrs = {k: inp[k].data_ptr() for k in _capi.INPUT_NAMES + ("dpdry0",)}
    ptrs["cush"] = cush.data_ptr()
    for k in _capi.OUTPUT_NAMES + ("wu",):
        ptrs[k] = out[k].data_ptr()
    head = [ctx_h] + ([] if noalloc else [_capi.FP64]) + [C.byref(_capi.default_params()), pcols, pcols, pver,
                                                          ch.ncnst, C.c_double(ch.dt)]
    return head + [C.c_void_p(ptrs[k]) for k in _capi.ARRAY_ARGS] + [C.c_void_p(out["diag"].data_ptr()),
                                                                      C.c_void_p(stream)]


def _run(ctx, ch, noalloc):
    import torch
    from uwshcu_amd import device
    inp = device.chunk_to_device(ch, "cuda:0")
    out = ctx.alloc_outputs(ch.ncol, ch.pver, ch.ncnst, diag=True)
    for v in out.values():
        v.fill_(-7)                       # no output may keep a fill value
    cush = inp["cush"].clone()
    stream = torch.cuda.current_stream(0).cuda_stream
    L = _capi.lib()
    fn = L.uwshcu_forward_device_noalloc if noalloc else L.uwshcu_forward_device
    return fn, _args(ctx._h, ch, inp, out, cush, stream, noalloc), out, cush, inp


def _host(out, cush):
    r = {k: v.cpu().numpy() for k, v in out.items()}
    r["cush"] = cush.cpu().numpy()
    return r


def test_noalloc_equals_forward_device_on_two_calls_in_a_row():
    import torch
    from uwshcu_amd import device
    chunks = [make_chunk(4096, 35, 3, seed=71), make_chunk(700, 30, 3, seed=72, recipe="baseline")]
    ctx = device.Context(0)
    # the allocating entry point first: its results, and every kernel loaded
    want = []
    for ch in chunks:
        fn, args, out, cush, _ = _run(ctx, ch, noalloc=False)
        assert fn(*args) == 0
        torch.cuda.synchronize()
        want.append(_host(out, cush))
    calls = [_run(ctx, ch, noalloc=True) for ch in chunks]   # inputs and outputs allocated up front
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info(0)[0]
    for i, (fn, args, _, _, _) in enumerate(calls):
        print(f"[noalloc] call {i}: {chunks[i].ncol}x{chunks[i].pver}x{chunks[i].ncnst}", flush=True)
        assert fn(*args) == 0
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info(0)[0]
    rc, bad = ctx.status()
    assert rc == 0, (rc, bad)
    assert free1 == free0, f"device free memory moved by {free0 - free1} bytes"
    for i, (_, _, out, cush, _) in enumerate(calls):
        got = _host(out, cush)
        for k, v in want[i].items():
            np.testing.assert_array_equal(got[k], v, err_msg=f"call {i}: {k}")
    ctx.close()


def test_noalloc_refuses_a_chunk_larger_than_its_scratch():
    import torch
    from uwshcu_amd import device
    per_col = _capi.workspace_bytes_per_column(35, 3)
    fits = _capi.NOALLOC_SCRATCH_BYTES // per_col // 64 * 64
    ch = make_chunk(fits + 64, 35, 3, seed=73)
    ctx = device.Context(0)
    fn, args, out, _, _ = _run(ctx, ch, noalloc=True)
    torch.cuda.synchronize()
    assert fn(*args) == SCRATCH_ERR
    torch.cuda.synchronize()
    assert (out["sten"] == -7).all()          # nothing was launched
    ctx.close()
