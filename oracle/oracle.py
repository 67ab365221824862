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


class ZW(C.Structure):
    """Options of the modified scheme (orc_zw in uwshcu_oracle.h)."""
    _fields_ = [(n, C.c_int) for n in ("lf", "ent", "cloud_sh", "pertt", "pertw", "pad_")] + [
        (n, C.c_double) for n in ("fwcrit", "ent0", "ufrcd", "ufrc0", "ufrcz", "tpertmax", "qpertmax")]


def zw_defaults(**kw) -> ZW:
    """The reference defaults (zwshcu.F90:30-57): lf on, every other option off.
    zmh_ent0 = 0.015 and zmh_shallow_qpertmax = 1.0e-3 are single-precision
    literals there, so their FP64 values are those of the nearest floats."""
    z = ZW(lf=1, ent=0, cloud_sh=0, pertt=0, pertw=0, pad_=0, fwcrit=1.0,
           ent0=float(np.float32(0.015)), ufrcd=2.0, ufrc0=2.0, ufrcz=2.0, tpertmax=2.0,
           qpertmax=float(np.float32(1.0e-3)))
    for k, v in kw.items():
        setattr(z, k, v)
    return z


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
        L.orc_compute_zwshcu_inv.restype = C.c_int
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


def run_zw(chunk, zw: ZW, consts: Consts | None = None, portable: bool = False, nthreads: int = 1):
    """The modified scheme (orc_compute_zwshcu_inv) on a chunk that carries the
    ``omega``, ``tpert`` and ``qpert`` inputs (generator.add_zw_inputs).  Consts
    supply rpen and rkm (zwshcu's uwshcu_rpen and zmh_rkm).  ``nthreads`` > 1 runs
    blocks of columns on a thread pool (columns are independent).  Returns like
    ``run``."""
    c = consts or default_consts(rpen=15.0, rkm=10.0)
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
                          "tr0", "tke", "pblh")]
    arrays = [np.ascontiguousarray(a) for a in ins] + [cush]
    arrays += [np.ascontiguousarray(f[k]) for k in ("omega", "tpert", "qpert")]
    arrays += [out[k] for k in ("umf", "slflx", "qtflx", "flxprc1", "flxsnow1", "qvten", "qlten", "qiten",
                                "sten", "uten", "vten", "trten", "qrten", "qsten", "precip", "snow", "evapc",
                                "cufrc", "qcu", "qlu", "qiu", "cbmf", "qc", "rliq", "cnt", "cnb")]
    arrays += [np.ascontiguousarray(f["dpdry0"]), out["wu"]]
    olib = lib(portable)

    def block(c0, c1):
        # columns [c0, c1): base pointers offset by c0, leading dimension n kept
        b = C.c_int(-1)
        c0, c1 = int(c0), int(c1)
        ptrs = [C.cast(a.ctypes.data + 8 * c0, C.POINTER(C.c_double)) for a in arrays]
        dptr = C.cast(diag.ctypes.data + 4 * NDIAG * c0, C.POINTER(C.c_int))
        r = olib.orc_compute_zwshcu_inv(C.byref(c), C.byref(zw), C.c_int(n), C.c_int(L), C.c_int(c1 - c0),
                                     C.c_int(ncnst), C.c_double(chunk.dt), *ptrs, dptr, C.byref(b))
        return r, (c0 + b.value if r else -1)

    nb = max(1, int(nthreads))
    edges = np.linspace(0, n, nb + 1).astype(int) if n else np.zeros(2, dtype=int)
    if nb == 1:
        res = [block(0, n)]
    else:
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(nb) as ex:
            res = list(ex.map(lambda j: block(edges[j], edges[j + 1]), range(nb)))
    rc, badc = 0, -1
    for r, b in res:
        if r and not rc:
            rc, badc = r, b
    out["cush"] = cush
    return out, diag, rc, badc
