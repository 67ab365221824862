"""Host-side mirror of compute_uwshcu_inv over the C ABI, on device-resident torch tensors.

``compute_uwshcu_inv`` takes the same chunk fields as the reference routine
(uwshcu.F90:251-265) and returns the same outputs.  Field layout: a layer field is a
C-contiguous float64 tensor ``[pver, pcols]`` (= Fortran ``(pcols, pver)``), an
interface field ``[pver+1, pcols]``, constituents ``[ncnst, pver, pcols]``, column
scalars ``[pcols]``; level 0 is the model top.  Tensors live on the GPU; the library
writes the outputs in place, stream-ordered on torch's current stream.

PyTorch only supplies device memory and streams here; every computation runs in
the HIP kernels of libuwshcu_hip.so.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _capi
from ._capi import FP64, MIXED, MIXED_UNCHECKED, UwshcuError  # noqa: F401

LAYER_OUT = ("qvten", "qlten", "qiten", "sten", "uten", "vten", "qrten", "qsten", "evapc", "cufrc", "qcu", "qlu",
             "qiu", "qc")
IFACE_OUT = ("umf", "slflx", "qtflx", "flxprc1", "flxsnow1", "wu")
COLUMN_OUT = ("precip", "snow", "cbmf", "rliq", "cnt", "cnb")


class Context:
    """A library context on one GPU: device constants, saturation table, workspace."""

    def __init__(self, device: int | torch.device | None = None, consts: _capi.Consts | None = None,
                 estbl=None, workspace_limit: int | None = None):
        L = _capi.lib()
        if device is None:
            device = torch.cuda.current_device()
        if isinstance(device, torch.device):
            device = device.index if device.index is not None else torch.cuda.current_device()
        self.device = int(device)
        self.consts = consts or _capi.default_consts()
        tab = None
        if estbl is not None:
            tab = (C.c_double * _capi.ESTBL_LEN)(*estbl)
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            rc = L.uwshcu_ctx_create(C.byref(self.consts), tab, self.device, C.byref(h))
        if rc:
            raise UwshcuError(rc, what="uwshcu_ctx_create")
        self._h = h
        if workspace_limit:
            L.uwshcu_ctx_set_workspace_limit(self._h, C.c_size_t(int(workspace_limit)))

    def close(self):
        if getattr(self, "_h", None):
            _capi.lib().uwshcu_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------
    def alloc_outputs(self, pcols: int, pver: int, ncnst: int, diag: bool = False) -> dict:
        dev = torch.device("cuda", self.device)
        o = {k: torch.empty((pver + 1, pcols), dtype=torch.float64, device=dev) for k in IFACE_OUT}
        o.update({k: torch.empty((pver, pcols), dtype=torch.float64, device=dev) for k in LAYER_OUT})
        o["trten"] = torch.empty((max(ncnst, 1), pver, pcols), dtype=torch.float64, device=dev)
        o.update({k: torch.empty((pcols,), dtype=torch.float64, device=dev) for k in COLUMN_OUT})
        if diag:
            o["diag"] = torch.empty((_capi.NDIAG, pcols), dtype=torch.int32, device=dev)
        return o

    def forward_into(self, inp: dict, out: dict, cush: torch.Tensor, *, ncol: int | None = None, dt: float = 1800.0,
                     precision: int = FP64, params: _capi.Params | None = None, stream=None) -> None:
        """Launch the forward physics; writes ``out`` and ``cush`` in place (asynchronous)."""
        pver, pcols = inp["p0"].shape
        ncnst = inp["tr0"].shape[0] if inp["tr0"].dim() == 3 else 0
        ncol = pcols if ncol is None else ncol
        par = params or _capi.default_params()
        ptrs = {}
        for k in _capi.INPUT_NAMES + ("dpdry0",):
            t = inp[k]
            if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
                raise ValueError(f"{k}: expected a contiguous float64 CUDA tensor")
            ptrs[k] = t.data_ptr()
        ptrs["cush"] = cush.data_ptr()
        for k in _capi.OUTPUT_NAMES + ("wu",):
            ptrs[k] = out[k].data_ptr()
        diag = out.get("diag")
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        sp = stream.cuda_stream if hasattr(stream, "cuda_stream") else stream
        args = [self._h, precision, C.byref(par), pcols, ncol, pver, ncnst, C.c_double(dt)]
        args += [C.c_void_p(ptrs[k]) for k in _capi.ARRAY_ARGS]
        args += [C.c_void_p(diag.data_ptr() if diag is not None else 0), C.c_void_p(sp)]
        rc = _capi.lib().uwshcu_forward_device(*args)
        if rc:
            raise UwshcuError(rc, what="uwshcu_forward_device")

    def zw_forward_into(self, inp: dict, out: dict, cush: torch.Tensor, zw: _capi.ZwParams, *,
                        ncol: int | None = None, dt: float = 1800.0, params: _capi.Params | None = None,
                        stream=None) -> None:
        """The modified scheme compute_zwshcu_inv (uwshcu_zw_forward_device, FP64):
        ``inp`` also holds ``omega`` [pver, pcols], ``tpert`` and ``qpert`` [pcols].
        Writes ``out`` and ``cush`` in place (asynchronous)."""
        pver, pcols = inp["p0"].shape
        ncnst = inp["tr0"].shape[0] if inp["tr0"].dim() == 3 else 0
        ncol = pcols if ncol is None else ncol
        par = params or _capi.zw_default_params()[0]
        ptrs = {}
        for k in _capi.INPUT_NAMES + _capi.ZW_INPUTS + ("dpdry0",):
            t = inp[k]
            if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
                raise ValueError(f"{k}: expected a contiguous float64 CUDA tensor")
            ptrs[k] = t.data_ptr()
        if inp["omega"].shape != inp["p0"].shape or inp["tpert"].shape != (pcols,) or inp["qpert"].shape != (pcols,):
            raise ValueError("omega must be [pver, pcols]; tpert and qpert [pcols]")
        ptrs["cush"] = cush.data_ptr()
        for k in _capi.OUTPUT_NAMES + ("wu",):
            ptrs[k] = out[k].data_ptr()
        diag = out.get("diag")
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        sp = stream.cuda_stream if hasattr(stream, "cuda_stream") else stream
        args = [self._h, C.byref(par), C.byref(zw), pcols, ncol, pver, ncnst, C.c_double(dt)]
        args += [C.c_void_p(ptrs[k]) for k in _capi.ZW_ARRAY_ARGS]
        args += [C.c_void_p(diag.data_ptr() if diag is not None else 0), C.c_void_p(sp)]
        rc = _capi.lib().uwshcu_zw_forward_device(*args)
        if rc:
            raise UwshcuError(rc, what="uwshcu_zw_forward_device")

    def compute_zwshcu_inv(self, inp: dict, zw: _capi.ZwParams | None = None, *, dt: float = 1800.0,
                           params: _capi.Params | None = None, diag: bool = False,
                           ncol: int | None = None) -> dict:
        """Run the modified scheme on the chunk and return the outputs (synchronous)."""
        pver, pcols = inp["p0"].shape
        ncnst = inp["tr0"].shape[0] if inp["tr0"].dim() == 3 else 0
        out = self.alloc_outputs(pcols, pver, ncnst, diag)
        cush = inp["cush"].clone()
        self.zw_forward_into(inp, out, cush, zw if zw is not None else _capi.zw_default_params()[1], ncol=ncol,
                             dt=dt, params=params)
        self.check()
        out["cush"] = cush
        return out

    @staticmethod
    def chunk_table(chunks: list):
        """The descriptor table of uwshcu_forward_chunks_device for ``chunks``, a list of
        ``(inp, out, cush, ncol)`` (``ncol`` None = all of the chunk's pcols), each laid
        out as for ``forward_into``.  Build it once and pass it to ``forward_chunks`` for
        every step that reuses the same arrays."""
        if not chunks:
            raise ValueError("no chunks")
        pver = chunks[0][0]["p0"].shape[0]
        tr = chunks[0][0]["tr0"]
        ncnst = tr.shape[0] if tr.dim() == 3 else 0
        desc = (_capi.ChunkDesc * len(chunks))()
        for j, (inp, out, cush, ncol) in enumerate(chunks):
            pv, pcols = inp["p0"].shape
            if pv != pver:
                raise ValueError("all chunks must have the same number of levels")
            d = desc[j]
            d.pcols, d.ncol = pcols, pcols if ncol is None else ncol
            for k in _capi.CHUNK_ARRAYS:
                t = cush if k == "cush" else (inp[k] if k in inp else out[k])
                if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
                    raise ValueError(f"{k}: expected a contiguous float64 CUDA tensor")
                setattr(d, k, t.data_ptr())
        return (desc, len(chunks), pver, ncnst)

    def forward_chunks(self, chunks, *, dt: float = 1800.0, precision: int = FP64,
                       params: _capi.Params | None = None, stream=None) -> None:
        """One launch over several chunks (a list as for ``chunk_table``, or its result).
        Writes every chunk's ``out`` and ``cush`` in place (asynchronous)."""
        desc, n, pver, ncnst = chunks if isinstance(chunks, tuple) else self.chunk_table(chunks)
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        sp = stream.cuda_stream if hasattr(stream, "cuda_stream") else stream
        par = params or _capi.default_params()
        rc = _capi.lib().uwshcu_forward_chunks_device(self._h, precision, C.byref(par), n, desc, pver, ncnst,
                                                       C.c_double(dt), C.c_void_p(sp))
        if rc:
            raise UwshcuError(rc, what="uwshcu_forward_chunks_device")

    def status(self, stream=None) -> tuple[int, int]:
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        sp = stream.cuda_stream if hasattr(stream, "cuda_stream") else stream
        bad = C.c_int(-1)
        rc = _capi.lib().uwshcu_status(self._h, C.c_void_p(sp), C.byref(bad))
        return rc, bad.value

    def mixed_recomputed(self, stream=None) -> int:
        """Columns the last MIXED call recomputed in FP64 (synchronises the stream)."""
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        sp = stream.cuda_stream if hasattr(stream, "cuda_stream") else stream
        n = C.c_int(0)
        rc = _capi.lib().uwshcu_mixed_recomputed(self._h, C.c_void_p(sp), C.byref(n))
        if rc:
            raise UwshcuError(rc, what="uwshcu_mixed_recomputed")
        return n.value

    def set_mixed_tolerance(self, tol: float) -> None:
        rc = _capi.lib().uwshcu_ctx_set_mixed_tolerance(self._h, C.c_double(tol))
        if rc:
            raise UwshcuError(rc, what="uwshcu_ctx_set_mixed_tolerance")

    def check(self, stream=None):
        rc, bad = self.status(stream)
        if rc:
            raise UwshcuError(rc, bad)

    def adjoint_into(self, inp: dict, out: dict, cush: torch.Tensor, cot: dict, grads: dict, *,
                     ncol: int | None = None, dt: float = 1800.0, params: _capi.Params | None = None,
                     stream=None) -> None:
        """Launch the adjoint (asynchronous): recompute the FP64 forward on tapes,
        write ``out``/``cush`` like :meth:`forward_into`, and write the gradients named
        in ``grads`` (t0, qv0, u0, v0, tke, pblh, par_col [3, pcols], par_sum [3])
        from the output cotangents ``cot`` (missing = zero)."""
        pver, pcols = inp["p0"].shape
        ncnst = inp["tr0"].shape[0] if inp["tr0"].dim() == 3 else 0
        ncol = pcols if ncol is None else ncol
        par = params or _capi.default_params()
        for k in _capi.INPUT_NAMES + ("dpdry0",):
            t = inp[k]
            if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
                raise ValueError(f"{k}: expected a contiguous float64 CUDA tensor")
        ptrs = {k: inp[k].data_ptr() for k in _capi.INPUT_NAMES + ("dpdry0",)}
        ptrs["cush"] = cush.data_ptr()
        for k in _capi.OUTPUT_NAMES + ("wu",):
            ptrs[k] = out[k].data_ptr()
        keep = []

        def dptr(t):
            if t is None:
                return None
            t = t.to(dtype=torch.float64).contiguous()
            keep.append(t)
            return t.data_ptr()
        ct = _capi.Cotangents(*[dptr(cot.get(k)) for k in _capi.COT_FIELDS])
        gr = _capi.Gradients(*[(grads[k].data_ptr() if grads.get(k) is not None else None)
                               for k in _capi.GRAD_FIELDS])
        diag = out.get("diag")
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        sp = stream.cuda_stream if hasattr(stream, "cuda_stream") else stream
        args = [self._h, C.byref(par), pcols, ncol, pver, ncnst, C.c_double(dt)]
        args += [C.c_void_p(ptrs[k]) for k in _capi.ARRAY_ARGS]
        args += [C.c_void_p(diag.data_ptr() if diag is not None else 0), C.byref(ct), C.byref(gr), C.c_void_p(sp)]
        rc = _capi.lib().uwshcu_adjoint_device(*args)
        if rc:
            raise UwshcuError(rc, what="uwshcu_adjoint_device")
        self._keep = keep   # cotangent copies stay alive until the next call

    def set_tape_capacity(self, nodes_per_column: int) -> None:
        rc = _capi.lib().uwshcu_ctx_set_tape_capacity(self._h, C.c_longlong(int(nodes_per_column)))
        if rc:
            raise UwshcuError(rc, what="uwshcu_ctx_set_tape_capacity")

    def compute_uwshcu_inv(self, inp: dict, *, dt: float = 1800.0, precision: int = FP64,
                           params: _capi.Params | None = None, diag: bool = False, ncol: int | None = None) -> dict:
        """Run the chunk and return the outputs (synchronous; raises on a fatal condition)."""
        pver, pcols = inp["p0"].shape
        ncnst = inp["tr0"].shape[0] if inp["tr0"].dim() == 3 else 0
        out = self.alloc_outputs(pcols, pver, ncnst, diag)
        cush = inp["cush"].clone()
        self.forward_into(inp, out, cush, ncol=ncol, dt=dt, precision=precision, params=params)
        self.check()
        out["cush"] = cush
        return out


def merge_shallow(fields: dict, *, ncol: int | None = None, latvap: float | None = None, stream=None) -> None:
    """convect_shallow_tend's merge after the scheme (uwshcu_merge_shallow_device) on
    device tensors laid out [levels, pcols]; updates the in/out fields and writes the
    outputs in place (asynchronous).  ``latvap`` defaults to the library's xlv."""
    pver, pcols = fields["pmid"].shape
    ncol = pcols if ncol is None else ncol
    latvap = _capi.default_consts().xlv if latvap is None else latvap
    f = _capi.MergeFields()
    for k in _capi.MERGE_FIELDS:
        t = fields[k]
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
            raise ValueError(f"{k}: expected a contiguous float64 CUDA tensor")
        setattr(f, k, t.data_ptr())
    if stream is None:
        stream = torch.cuda.current_stream(fields["pmid"].device)
    sp = stream.cuda_stream if hasattr(stream, "cuda_stream") else stream
    rc = _capi.lib().uwshcu_merge_shallow_device(pcols, ncol, pver, C.c_double(latvap), C.byref(f), C.c_void_p(sp))
    if rc:
        raise UwshcuError(rc, what="uwshcu_merge_shallow_device")


def chunk_to_device(chunk, device="cuda") -> dict:
    """Copy a generator chunk (numpy fields) to device tensors."""
    return {k: torch.from_numpy(v).to(device) for k, v in chunk.fields.items()}


_default_ctx = {}


def default_context(device=None) -> Context:
    d = torch.cuda.current_device() if device is None else int(device)
    if d not in _default_ctx or _default_ctx[d]._h is None:
        _default_ctx[d] = Context(d)
    return _default_ctx[d]


def compute_uwshcu_inv(inp: dict, *, dt: float = 1800.0, precision: int = FP64, params=None, diag=False) -> dict:
    """Module-level form of :meth:`Context.compute_uwshcu_inv` on the current device."""
    return default_context().compute_uwshcu_inv(inp, dt=dt, precision=precision, params=params, diag=diag)
