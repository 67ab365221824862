"""ctypes binding of the C ABI in ``include/uwshcu.h`` (libuwshcu_hip.so).

The library is built in-tree by ``uwshcu_amd.build.build()`` (hipcc for gfx950).
Loading fails loudly when it is missing: there is no CPU path in the product.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libuwshcu_hip.so")

NCNST_MAX = 32
ESTBL_LEN = 250
NDIAG = 12

FP64 = 0
MIXED = 1                  # fast FP64 pass (converged saturation adjustment) + check + FP64 recompute
MIXED_UNCHECKED = 2        # experimental FP32-plume pass alone (tendencies off by 1-2 %)
MIXED_FAST_UNCHECKED = 3   # the fast FP64 pass alone

ERRORS = {
    0: "ok",
    -1: "findsp not converging",
    -2: "findsp enthalpy not conserved",
    -3: "mu > mumin2 in the cloud-base closure",
    -10: "invalid argument",
    -11: "workspace allocation failed",
    -12: "HIP runtime error",
    -13: "no usable GPU",
    -14: "adjoint tape capacity exceeded",
    -15: "chunk needs more than the static scratch of the allocation-free entry point",
    -16: "drop-in called before uwshcu_set_dropin_params (uwshcu_rpen unset)",
}

NOALLOC_SCRATCH_BYTES = 256 << 20   # UWSHCU_NOALLOC_SCRATCH_BYTES


class Consts(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("xlv", "xlf", "xls", "cp", "zvir", "r", "g", "ep2", "p00", "rovcp",
                                          "tmin", "tmax", "ttrice", "epsqs", "rgasv", "hlatf", "hlatv", "cpws",
                                          "tmelt")] + [
        ("icephs", C.c_int), ("ixnumliq", C.c_int), ("ixnumice", C.c_int), ("pad_", C.c_int),
        ("pcf", C.c_double * 5),
        ("qmin", C.c_double * (NCNST_MAX + 1)),
        ("wet", C.c_int * (NCNST_MAX + 1)),
    ]


class Params(C.Structure):
    _fields_ = [("rpen", C.c_double), ("rle", C.c_double), ("rkm", C.c_double)]


# Output fields in the order of uwshcu_cotangents (uwshcu_types.h).
COT_FIELDS = ("umf", "slflx", "qtflx", "flxprc1", "flxsnow1", "wu", "qvten", "qlten", "qiten", "sten", "uten", "vten",
              "qrten", "qsten", "evapc", "cufrc", "qcu", "qlu", "qiu", "qc", "trten", "precip", "snow", "cbmf", "rliq",
              "cnt", "cnb", "cush")
GRAD_FIELDS = ("t0", "qv0", "u0", "v0", "tke", "pblh", "par_col", "par_sum")


class Cotangents(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in COT_FIELDS]


class Gradients(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in GRAD_FIELDS]


# uwshcu_chunk (uwshcu_types.h): one chunk of uwshcu_forward_chunks_device
CHUNK_ARRAYS = ("ps0", "zs0", "p0", "z0", "dp0", "u0", "v0", "qv0", "ql0", "qi0", "t0", "s0", "tr0", "tke", "pblh",
                "dpdry0", "cush", "umf", "slflx", "qtflx", "flxprc1", "flxsnow1", "wu", "qvten", "qlten", "qiten",
                "sten", "uten", "vten", "trten", "qrten", "qsten", "evapc", "cufrc", "qcu", "qlu", "qiu", "qc",
                "precip", "snow", "cbmf", "rliq", "cnt", "cnb")


class ChunkDesc(C.Structure):
    _fields_ = [("pcols", C.c_int), ("ncol", C.c_int)] + [(n, C.c_void_p) for n in CHUNK_ARRAYS]


# uwshcu_merge_fields (uwshcu_types.h), in declaration order
MERGE_FIELDS = ("umf", "slflx", "qtflx", "qrten", "qsten", "qcu", "qc2", "rliq2", "cnt2", "cnb2", "cbmf", "pmid",
                "rprddp", "cmfmc", "cnt", "cnb", "rliq", "qc", "icwmr", "rprdsh", "rprdtot", "cmfsl", "cmflq",
                "freqsh", "pcnt", "pcnb", "massflx_sh")


class MergeFields(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in MERGE_FIELDS]


class ZwParams(C.Structure):
    """Options of the modified scheme compute_zwshcu_inv (uwshcu_zw_params)."""
    _fields_ = [(n, C.c_int) for n in ("lf", "ent", "cloud_sh", "pertt", "pertw", "pad_")] + [
        (n, C.c_double) for n in ("fwcrit", "ent0", "ufrcd", "ufrc0", "ufrcz", "tpertmax", "qpertmax")]


# Argument names of the forward entry points, after (ctx, precision, par, pcols, ncol,
# pver, ncnst, dt), in the order of compute_uwshcu_inv (uwshcu.F90:251-265).
INPUT_NAMES = ("ps0", "zs0", "p0", "z0", "dp0", "u0", "v0", "qv0", "ql0", "qi0", "t0", "s0", "tr0", "tke",
               "cldfrc", "concldfrc", "pblh")
OUTPUT_NAMES = ("umf", "slflx", "qtflx", "flxprc1", "flxsnow1", "qvten", "qlten", "qiten", "sten", "uten",
                "vten", "trten", "qrten", "qsten", "precip", "snow", "evapc", "cufrc", "qcu", "qlu", "qiu",
                "cbmf", "qc", "rliq", "cnt", "cnb")
# full ordered list of the array arguments
ARRAY_ARGS = INPUT_NAMES + ("cush",) + OUTPUT_NAMES + ("dpdry0", "wu")
# uwshcu_zw_forward_device: after (ctx, par, zw, pcols, ncol, pver, ncnst, dt), in the
# order of compute_zwshcu_inv (zwshcu.F90:336-350) without its unused inputs
ZW_INPUTS = ("omega", "tpert", "qpert")
ZW_ARRAY_ARGS = INPUT_NAMES + ("cush",) + ZW_INPUTS + OUTPUT_NAMES + ("dpdry0", "wu")

SYMBOLS = ("uwshcu_default_consts", "uwshcu_build_estbl", "uwshcu_default_params", "uwshcu_ctx_create",
           "uwshcu_ctx_destroy", "uwshcu_ctx_set_workspace_limit", "uwshcu_workspace_bytes_per_column",
           "uwshcu_forward_device", "uwshcu_status", "uwshcu_forward_host", "bxcompute_uwshcu_inv_",
           "uwshcu_set_dropin_params", "uwshcu_dropin_last_status", "uwshcu_ctx_set_mixed_tolerance",
           "uwshcu_mixed_recomputed", "uwshcu_adjoint_device", "uwshcu_ctx_set_tape_capacity",
           "uwshcu_forward_chunks_device", "uwshcu_merge_shallow_device", "uwshcu_zw_default_params",
           "uwshcu_zw_forward_device", "uwshcu_zw_forward_host", "uwshcu_forward_device_noalloc")

_lib = None


class UwshcuError(RuntimeError):
    def __init__(self, code, col=-1, what="uwshcu"):
        self.code, self.col = code, col
        super().__init__(f"{what}: {ERRORS.get(code, 'error')} (code {code}, column {col})")


def lib():
    """Load libuwshcu_hip.so; raise if it has not been built.

    torch is imported first: torch ships its own libamdhip64.so.7 with the same
    soname as the ROCm installation's, and the first one loaded serves the whole
    process.  Loading torch's first keeps torch and this library on one HIP
    runtime (the library is ABI-compatible with either).
    """
    global _lib
    if _lib is not None:
        return _lib
    import torch  # noqa: F401  (see above)
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with uwshcu_amd.build.build() (hipcc, gfx950)")
    L = C.CDLL(LIB_PATH)
    vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.c_double
    L.uwshcu_default_consts.argtypes = [C.POINTER(Consts)]
    L.uwshcu_build_estbl.argtypes = [C.POINTER(Consts), C.POINTER(C.c_double)]
    L.uwshcu_default_params.argtypes = [C.POINTER(Params)]
    L.uwshcu_ctx_create.argtypes = [C.POINTER(Consts), C.POINTER(C.c_double), C.c_int, C.POINTER(vp)]
    L.uwshcu_ctx_destroy.argtypes = [vp]
    L.uwshcu_ctx_set_workspace_limit.argtypes = [vp, C.c_size_t]
    L.uwshcu_workspace_bytes_per_column.argtypes = [C.c_int, C.c_int, C.c_int]
    L.uwshcu_workspace_bytes_per_column.restype = C.c_size_t
    head = [vp, C.c_int, C.POINTER(Params), C.c_int, C.c_int, C.c_int, C.c_int, dp]
    L.uwshcu_forward_device.argtypes = head + [vp] * len(ARRAY_ARGS) + [vp, vp]
    L.uwshcu_forward_device_noalloc.argtypes = [vp, C.POINTER(Params), C.c_int, C.c_int, C.c_int, C.c_int, dp] + \
        [vp] * len(ARRAY_ARGS) + [vp, vp]
    L.uwshcu_zw_default_params.argtypes = [C.POINTER(Params), C.POINTER(ZwParams)]
    L.uwshcu_zw_forward_device.argtypes = [vp, C.POINTER(Params), C.POINTER(ZwParams), C.c_int, C.c_int, C.c_int,
                                           C.c_int, dp] + [vp] * len(ZW_ARRAY_ARGS) + [vp, vp]
    L.uwshcu_zw_forward_host.argtypes = [vp, C.POINTER(Params), C.POINTER(ZwParams), C.c_int, C.c_int, C.c_int,
                                         C.c_int, dp] + [vp] * len(ZW_ARRAY_ARGS) + [vp, ip]
    L.uwshcu_forward_host.argtypes = head + [vp] * len(ARRAY_ARGS) + [vp, ip]
    L.uwshcu_status.argtypes = [vp, vp, ip]
    L.uwshcu_set_dropin_params.argtypes = [C.POINTER(Params)]
    L.uwshcu_dropin_last_status.argtypes = [ip]
    L.uwshcu_ctx_set_mixed_tolerance.argtypes = [vp, dp]
    L.uwshcu_mixed_recomputed.argtypes = [vp, vp, ip]
    L.uwshcu_adjoint_device.argtypes = [vp, C.POINTER(Params), C.c_int, C.c_int, C.c_int, C.c_int, dp] + \
        [vp] * len(ARRAY_ARGS) + [vp, C.POINTER(Cotangents), C.POINTER(Gradients), vp]
    L.uwshcu_ctx_set_tape_capacity.argtypes = [vp, C.c_longlong]
    L.uwshcu_forward_chunks_device.argtypes = [vp, C.c_int, C.POINTER(Params), C.c_int, C.POINTER(ChunkDesc), C.c_int,
                                               C.c_int, dp, vp]
    L.uwshcu_merge_shallow_device.argtypes = [C.c_int, C.c_int, C.c_int, dp, C.POINTER(MergeFields), vp]
    _lib = L
    return L


def default_consts() -> Consts:
    c = Consts()
    lib().uwshcu_default_consts(C.byref(c))
    return c


def build_estbl(c: Consts | None = None):
    c = c or default_consts()
    t = (C.c_double * ESTBL_LEN)()
    rc = lib().uwshcu_build_estbl(C.byref(c), t)
    if rc:
        raise UwshcuError(rc, what="uwshcu_build_estbl")
    return list(t)


def default_params(rpen=10.0, rle=0.1, rkm=14.0) -> Params:
    return Params(rpen, rle, rkm)


def zw_default_params(**kw) -> tuple:
    """(Params, ZwParams) with the modified scheme's defaults (uwshcu_zw_default_params);
    keyword arguments override ZwParams fields."""
    p, z = Params(), ZwParams()
    rc = lib().uwshcu_zw_default_params(C.byref(p), C.byref(z))
    if rc:
        raise UwshcuError(rc, what="uwshcu_zw_default_params")
    for k, v in kw.items():
        setattr(z, k, v)
    return p, z


def workspace_bytes_per_column(pver, ncnst, precision=FP64) -> int:
    return int(lib().uwshcu_workspace_bytes_per_column(pver, ncnst, precision))
