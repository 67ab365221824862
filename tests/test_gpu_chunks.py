"""One launch over many chunks (uwshcu_forward_chunks_device): against the portable-math
oracle on the same columns and against the same columns run as one chunk, bit for bit,
with untouched padding."""
import numpy as np
import pytest

import oracle
from uwshcu_amd.generator import make_chunk

pytestmark = pytest.mark.gpu

SENTINEL = 12345.0


def _split(inp, sizes, pad):
    """Copy consecutive column ranges of inp into separate chunks with pcols = ncol + pad."""
    import torch
    chunks, c0 = [], 0
    for n in sizes:
        pc = n + pad
        sub = {}
        for k, v in inp.items():
            t = torch.full(v.shape[:-1] + (pc,), SENTINEL, dtype=v.dtype, device=v.device)
            t[..., :n] = v[..., c0:c0 + n]
            sub[k] = t
        chunks.append((sub, n, c0))
        c0 += n
    return chunks


def _run_chunks(ctx, inp, sizes, pad, dt):
    import torch
    pver, _ = inp["p0"].shape
    ncnst = inp["tr0"].shape[0]
    parts = _split(inp, sizes, pad)
    calls = []
    for sub, n, _ in parts:
        out = ctx.alloc_outputs(n + pad, pver, ncnst)
        for v in out.values():
            v.fill_(SENTINEL)
        calls.append((sub, out, sub["cush"], n))
    ctx.forward_chunks(calls, dt=dt)
    ctx.check()
    torch.cuda.synchronize()
    return parts, calls


@pytest.mark.parametrize("sizes,pad", [([1000, 1, 999, 0, 1000], 3), ([64] * 9 + [7], 0)])
def test_chunks_equal_one_chunk(sizes, pad):
    import torch
    from uwshcu_amd import device
    ch = make_chunk(sum(sizes), 30, 3, seed=41)
    ctx = device.default_context()
    inp = device.chunk_to_device(ch)
    ref = ctx.compute_uwshcu_inv(inp, dt=ch.dt)
    torch.cuda.synchronize()
    parts, calls = _run_chunks(ctx, inp, sizes, pad, ch.dt)
    orc, _, orc_rc, _ = oracle.run(ch, nthreads=8, portable=True)
    assert orc_rc == 0
    for (sub, n, c0), (_, out, cush, _) in zip(parts, calls):
        got = dict(out, cush=cush)
        for k, r in ref.items():
            g = got[k].cpu().numpy()
            np.testing.assert_array_equal(g[..., :n], orc[k][..., c0:c0 + n], err_msg=k)   # the oracle
            np.testing.assert_array_equal(g[..., :n], r[..., c0:c0 + n].cpu().numpy(), err_msg=k)
            assert np.all(g[..., n:] == SENTINEL), k


def test_chunks_large_enough_for_the_deferring_pipeline():
    """40 chunks of 4096 columns: the packed range (163840 columns) runs the
    two-wavefront column kernel with deferred second passes.  Every chunk is a copy of
    the same 4096 columns, so every chunk must equal the oracle on them."""
    import torch
    from uwshcu_amd import device
    base = make_chunk(4096, 35, 3, seed=43)
    ctx = device.default_context()
    one = device.chunk_to_device(base)
    inp = {k: v.repeat(*([1] * (v.dim() - 1)), 40).contiguous() for k, v in one.items()}
    ref = ctx.compute_uwshcu_inv(inp, dt=base.dt)
    torch.cuda.synchronize()
    parts, calls = _run_chunks(ctx, inp, [4096] * 40, 0, base.dt)
    orc, _, orc_rc, _ = oracle.run(base, nthreads=8, portable=True)
    assert orc_rc == 0
    for (sub, n, c0), (_, out, cush, _) in zip(parts, calls):
        got = dict(out, cush=cush)
        for k, r in ref.items():
            assert torch.equal(got[k], r[..., c0:c0 + n]), k
            np.testing.assert_array_equal(got[k].cpu().numpy(), orc[k], err_msg=k)
