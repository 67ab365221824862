"""The forward entry point under HIP graph capture: once the context's workspace
exists, uwshcu_forward_device issues only kernels (no allocation, no host
synchronisation, no memset node), so a caller can capture it and replay it.  Each
replay, with new inputs written in place between replays, must equal an eager call
on the same inputs bit for bit."""
import pytest

from uwshcu_amd.generator import make_chunk

pytestmark = pytest.mark.gpu


def test_captured_forward_replays_equal_eager_calls():
    import torch
    from uwshcu_amd import device
    ch = make_chunk(4096, 35, 3, seed=51)
    ctx = device.Context(torch.cuda.current_device())
    inp = device.chunk_to_device(ch)
    t0_base = inp["t0"].clone()
    cush0 = inp["cush"].clone()
    out = ctx.alloc_outputs(ch.ncol, ch.pver, ch.ncnst)
    cush = cush0.clone()

    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(2):                     # sizes the workspace outside the capture
            cush.copy_(cush0)
            ctx.forward_into(inp, out, cush, dt=ch.dt)
    torch.cuda.current_stream().wait_stream(s)
    ctx.check()

    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        cush.copy_(cush0)
        ctx.forward_into(inp, out, cush, dt=ch.dt)
    torch.cuda.synchronize()

    rng = torch.Generator(device="cpu").manual_seed(3)
    for rep in range(3):
        # new temperatures, written into the captured input array
        pert = (torch.rand(t0_base.shape, generator=rng, dtype=torch.float64) - 0.5).to(t0_base.device)
        inp["t0"].copy_(t0_base + 0.5 * pert)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        got = {k: v.clone() for k, v in out.items()}
        got_cush = cush.clone()
        ref = ctx.compute_uwshcu_inv(inp, dt=ch.dt)   # eager, same inputs
        torch.cuda.synchronize()
        for k, v in got.items():
            assert torch.equal(v, ref[k]), f"replay {rep}: {k}"
        assert torch.equal(got_cush, ref["cush"]), f"replay {rep}: cush"
        assert not torch.equal(got["sten"], torch.zeros_like(got["sten"]))
    ctx.close()
