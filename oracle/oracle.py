"""ctypes access to the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only the test suite, the smoke check and the CPU-baseline leg of the benchmark
import this module.  The oracle restates the reference Fortran
(uwshcu.F90:251-5454, wv_saturation.F90, gffgch.F90) in C; see
``uwshcu_oracle.c`` for the per-function citations.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")          # host libm
LIB_PATH_PM = os.path.join(HERE, "liboracle_pm.so")    # portable exp/log/pow/erfc
NCNST_MAX = 32
NDIAG = 12

OUTPUT_IFACE = ("umf", "slflx", "qtflx", "flxprc1", "flxsnow1", "wu")
OUTPUT_LAYER = ("qvten", "qlten", "qiten", "sten", "uten", "vten", "qrten", "qsten", "evapc",
                "cufrc", "qcu", "qlu", "qiu", "qc")
OUTPUT_COLUMN = ("precip", "snow", "cbmf", "rliq", "cnt", "cnb", "cush")


class Consts(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("xlv", "xlf", "xls", "cp", "zvir", "r", "g", "ep2", "p00", "rovcp",
                                          "rpen", "rle", "rkm", "tmin", "tmax", "ttrice", "epsqs", "rgasv",
                                          "hlatf", "hlatv", "cpws", "tmelt")] + [
        ("icephs", C.c_int),
        ("pcf", C.c_double * 5),
        ("estbl", C.c_double * 250),
        ("qmin", C.c_double * (NCNST_MAX + 1)),
        ("wet", C.c_int * (NCNST_MAX + 1)),
        ("ixnumliq", C.c_int),
        ("ixnumice", C.c_int),
    ]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_libs = {}


def lib(portable: bool = False):
    """The oracle library; ``portable`` selects the portable-math flavour."""
    if portable not in _libs:
        path = LIB_PATH_PM if portable else LIB_PATH
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        d = C.c_double
        pc = C.POINTER(Consts)
        L.orc_default_consts.argtypes = [pc]
        L.orc_gffgch.argtypes = [pc, d, dp, C.c_int]
        L.orc_estblf.argtypes = [pc, d]
        L.orc_estblf.restype = d
        L.orc_qsat.argtypes = [pc, d, d, dp, dp, dp]
        L.orc_conden.argtypes = [pc, d, d, d, dp, dp, dp, dp, dp, C.POINTER(C.c_int)]
        L.orc_roots.argtypes = [d, d, d, dp, dp, C.POINTER(C.c_int)]
        L.orc_slope.argtypes = [C.c_int, dp, dp, dp]
        L.orc_qsinvert.argtypes = [pc, d, d, d]
        L.orc_qsinvert.restype = d
        L.orc_compute_alpha.argtypes = [d, d]
        L.orc_compute_alpha.restype = d
        L.orc_compute_mumin2.argtypes = [d, d, d]
        L.orc_compute_mumin2.restype = d
        L.orc_compute_ppen.argtypes = [d, d, d, d, d, d]
        L.orc_compute_ppen.restype = d
        L.orc_fluxbelowinv.argtypes = [pc, d, dp, C.c_int, C.c_int, d, d, d, d, d, dp]
        L.orc_positive_moisture_single.argtypes = [d, d, C.c_int, d, d, d, d] + [dp] * 9
        L.orc_single_cin.argtypes = [pc, d, d, d, d, d, d]
        L.orc_single_cin.restype = d
        L.orc_getbuoy.argtypes = [pc, d, d, d, d, d, d, dp, dp]
        L.orc_findsp_point.argtypes = [pc, d, d, d, dp, dp, dp]
        L.orc_findsp_point.restype = C.c_int
        L.orc_compute_uwshcu_inv_omp.restype = C.c_int
        _libs[portable] = L
    return _libs[portable]


def default_consts(rpen=10.0, rle=0.1, rkm=14.0) -> Consts:
    c = Consts()
    lib().orc_default_consts(C.byref(c))
    c.rpen, c.rle, c.rkm = rpen, rle, rkm
    return c


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def run(chunk, consts: Consts | None = None, nthreads: int = 0, portable: bool = False):
    """Run the oracle on a generator chunk; returns (outputs dict, diag [ncol, NDIAG], rc, bad_col).

    portable=False uses the host libm (the faithful reading of the reference);
    portable=True the portable FP64 functions of orc_math.h.
    """
    c = consts or default_consts()
    f = chunk.fields
    n, L, ncnst = chunk.ncol, chunk.pver, chunk.ncnst
    out = {k: np.zeros((L + 1, n)) for k in OUTPUT_IFACE}
    out.update({k: np.zeros((L, n)) for k in OUTPUT_LAYER})
    out["trten"] = np.zeros((ncnst, L, n))
    for k in ("precip", "snow", "cbmf", "rliq", "cnt", "cnb"):
        out[k] = np.zeros(n)
    cush = f["cush"].copy()
    diag = np.zeros((n, NDIAG), dtype=np.int32)
    bad = C.c_int(-1)
    ins = [f[k] for k in ("ps0", "zs0", "p0", "z0", "dp0", "u0", "v0", "qv0", "ql0", "qi0", "t0", "s0",
                          "tr0", "tke", "cldfrc", "concldfrc", "pblh")]
    args = [C.byref(c), C.c_int(nthreads), C.c_int(n), C.c_int(L), C.c_int(n), C.c_int(ncnst),
            C.c_double(chunk.dt)] + [_p(np.ascontiguousarray(a)) for a in ins] + [_p(cush)]
    args += [_p(out[k]) for k in ("umf", "slflx", "qtflx", "flxprc1", "flxsnow1", "qvten", "qlten", "qiten",
                                  "sten", "uten", "vten", "trten", "qrten", "qsten", "precip", "snow", "evapc",
                                  "cufrc", "qcu", "qlu", "qiu", "cbmf", "qc", "rliq", "cnt", "cnb")]
    args += [_p(np.ascontiguousarray(f["dpdry0"])), _p(out["wu"]),
             diag.ctypes.data_as(C.POINTER(C.c_int)), C.byref(bad)]
    rc = lib(portable).orc_compute_uwshcu_inv_omp(*args)
    out["cush"] = cush
    return out, diag, rc, bad.value
