"""torch.autograd binding of compute_uwshcu_inv and its adjoint.

``uwshcu(inputs, params)`` runs the FP64 HIP forward and returns the outputs as
tensors that carry gradients back to t0, qv0, u0, v0, tke, pblh and the tunables
``params = tensor([rpen, rle, rkm])``.  The backward pass calls the library's adjoint
(``uwshcu_adjoint_device``): it re-runs each column's FP64 forward while recording it on
a tape, so the branch pattern is the forward's own, and sweeps the tape back from the
output cotangents.  The forward keeps no tape; the backward recomputes it in column
batches that fit the context's workspace.

Gradients are those of the scheme as computed: piecewise-smooth, exact between the
scheme's branch points (saturation-table segments, min/max limiters, level choices).
pblh acts only through the integer inversion level, so its gradient is zero.
"""
from __future__ import annotations

import torch

from . import _capi
from .device import Context, default_context

DIFF_INPUTS = ("t0", "qv0", "u0", "v0", "tke", "pblh")
OUTPUTS = ("umf", "slflx", "qtflx", "flxprc1", "flxsnow1", "wu", "qvten", "qlten", "qiten", "sten", "uten", "vten",
           "qrten", "qsten", "evapc", "cufrc", "qcu", "qlu", "qiu", "qc", "trten", "precip", "snow", "cbmf", "rliq",
           "cnt", "cnb", "cush")


class UwshcuFunction(torch.autograd.Function):
    @staticmethod
    def forward(fctx, t0, qv0, u0, v0, tke, pblh, params, static, context, dt):
        inp = dict(static)
        inp.update(t0=t0, qv0=qv0, u0=u0, v0=v0, tke=tke, pblh=pblh)
        pv = [float(x) for x in params.detach().cpu().tolist()]
        par = _capi.Params(*pv)
        pver, pcols = inp["p0"].shape
        ncnst = inp["tr0"].shape[0] if inp["tr0"].dim() == 3 else 0
        out = context.alloc_outputs(pcols, pver, ncnst)
        cush = inp["cush"].clone()
        context.forward_into(inp, out, cush, dt=dt, params=par)
        # Every input tensor goes through save_for_backward, so an in-place change
        # between forward and backward raises autograd's version error instead of
        # re-recording the tape on changed values.
        names = tuple(inp.keys())
        fctx.save_for_backward(*(inp[k] for k in names))
        fctx.names = names
        fctx.par = pv
        fctx.context = context
        fctx.dt = dt
        fctx.params_device = params.device
        fctx.params_dtype = params.dtype
        res = [out[k] for k in OUTPUTS[:-1]] + [cush]
        for k in ("cnt", "cnb"):
            fctx.mark_non_differentiable(out[k])
        return tuple(res)

    @staticmethod
    def backward(fctx, *grad_out):
        context = fctx.context
        inp = dict(zip(fctx.names, fctx.saved_tensors))
        pver, pcols = inp["p0"].shape
        ncnst = inp["tr0"].shape[0] if inp["tr0"].dim() == 3 else 0
        cot = {k: g for k, g in zip(OUTPUTS, grad_out) if g is not None}
        dev = inp["p0"].device
        grads = {k: torch.empty_like(inp[k]) for k in DIFF_INPUTS}
        grads["par_sum"] = torch.empty(3, dtype=torch.float64, device=dev)
        scratch = context.alloc_outputs(pcols, pver, ncnst)
        cush = inp["cush"].clone()
        context.adjoint_into(inp, scratch, cush, cot, grads, dt=fctx.dt, params=_capi.Params(*fctx.par))
        context.check()
        gpar = grads["par_sum"].to(device=fctx.params_device, dtype=fctx.params_dtype)
        return tuple(grads[k] for k in DIFF_INPUTS) + (gpar, None, None, None)


def uwshcu(inputs: dict, params: torch.Tensor, *, dt: float = 1800.0, context: Context | None = None) -> dict:
    """Differentiable compute_uwshcu_inv on device tensors (layout as in ``device``).

    ``params`` is a float64 tensor [rpen, rle, rkm].  Returns a dict of outputs.
    """
    context = context or default_context()
    static = {k: v for k, v in inputs.items() if k not in DIFF_INPUTS}
    res = UwshcuFunction.apply(*(inputs[k] for k in DIFF_INPUTS), params, static, context, dt)
    return dict(zip(OUTPUTS, res))
