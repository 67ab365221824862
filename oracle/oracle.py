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
        for fn in ("orc_xexp", "orc_xlog", "orc_xerfc"):
            getattr(L, fn).argtypes = [d]
            getattr(L, fn).restype = d
        L.orc_xpow.argtypes = [d, d]
        L.orc_xpow.restype = d
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
        L.orc_set_probe.argtypes = [dp]
        L.orc_probe_layout.argtypes = [C.POINTER(C.c_int)] * 3
        L.orc_probe_layout.restype = C.c_int
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


# Stage-probe record (uwshcu_oracle.h, ORC_P_* and ORC_PL_*), in enum order.
PROBE_SCALARS = ("closure", "wcrit", "sigmaw", "mu", "rho0inv", "cbmf", "winv", "wlcl", "ufrclcl", "cinlcl",
                 "rbuoy", "rmaxfrac", "cbmflimit", "mumin2_applied",
                 "ppen_set", "wtwb", "drage", "bogbot", "bogtop", "rho0j", "dpen", "ppen",
                 "krel", "kbup", "kpen", "rpen", "PGFc", "g", "criqc", "emf_set", "limit_emf")
PROBE_LEVELS = ("set", "case", "xc", "wue", "cridis", "thvj", "thv0j", "xsat", "thvxsat", "thlue", "qtue",
                "excessu", "excess0", "thle", "qte", "ue", "ve", "pe", "dpe", "fer", "fdr",
                "thlu0", "qtu0", "uu0", "vu0", "thlu1", "qtu1", "uu1", "vu1", "ssthl", "ssqt", "ssu", "ssv",
                "umf0", "umf1", "rei", "wtwb_k", "wtw_k", "drage_k", "bogbot_k", "bogtop_k", "rho0j_k",
                "emf", "thlu_emf", "qtu_emf", "thl0", "qt0", "ps0", "p0", "ssthl0", "ssqt0", "umf", "reif", "dp0")


def column_chunk(chunk, col: int):
    """The one-column chunk holding column ``col`` of ``chunk``."""
    from copy import copy
    one = copy(chunk)
    one.fields = {k: np.ascontiguousarray(v[..., col:col + 1]) for k, v in chunk.fields.items()}
    one.ncol = 1
    return one


def probe(chunk, col: int, consts: Consts | None = None, portable: bool = False):
    """Run the oracle on column ``col`` alone with the stage probe on.  Returns (outputs,
    diag row, record) where record maps PROBE_SCALARS names to floats and PROBE_LEVELS
    names to arrays over UW levels 0..ORC_PRB_KMAX-1."""
    L = lib(portable)
    ns, km, nf = C.c_int(), C.c_int(), C.c_int()
    total = L.orc_probe_layout(C.byref(ns), C.byref(km), C.byref(nf))
    assert (ns.value, nf.value) == (len(PROBE_SCALARS), len(PROBE_LEVELS)), "probe layout differs"
    buf = np.zeros(total)
    L.orc_set_probe(buf.ctypes.data_as(C.POINTER(C.c_double)))
    try:
        out, diag, rc, _ = run(column_chunk(chunk, col), consts=consts, nthreads=1, portable=portable)
    finally:
        L.orc_set_probe(None)
    rec = {k: float(buf[i]) for i, k in enumerate(PROBE_SCALARS)}
    base = ns.value
    for f, k in enumerate(PROBE_LEVELS):
        rec[k] = buf[base + f * km.value: base + (f + 1) * km.value].copy()
    assert rc == 0
    return out, diag[0], rec


LIB_PATH_COUNT = os.path.join(HERE, "liboracle_count.so")
LIB_PATH_DUAL = os.path.join(HERE, "liboracle_dual.so")
OP_KINDS = ("add", "mul", "div", "sqrt", "cmp", "other")
_IN_ORDER = ("ps0", "zs0", "p0", "z0", "dp0", "u0", "v0", "qv0", "ql0", "qi0", "t0", "s0", "tr0", "tke", "cldfrc",
             "concldfrc", "pblh", "cush", "dpdry0")
_OUT_ORDER = ("umf", "slflx", "qtflx", "flxprc1", "flxsnow1", "qvten", "qlten", "qiten", "sten", "uten", "vten",
              "trten", "qrten", "qsten", "precip", "snow", "evapc", "cufrc", "qcu", "qlu", "qiu", "cbmf", "qc", "rliq",
              "cnt", "cnb", "wu")


def count_ops(chunk, cols=None, consts: Consts | None = None):
    """FP64 operations per column of the portable-math oracle (liboracle_count.so):
    returns (ops [len(cols), 6] in OP_KINDS order, outputs, diag).  The outputs equal
    run(portable=True)'s bit for bit: the counting number computes with doubles."""
    if "count" not in _libs:
        if not os.path.exists(LIB_PATH_COUNT):
            subprocess.run(["make", "-s", "-C", HERE, "liboracle_count.so"], check=True)
        Lc = C.CDLL(LIB_PATH_COUNT)
        Lc.orc_default_consts.argtypes = [C.POINTER(Consts)]
        Lc.orc_count_column.restype = C.c_int
        _libs["count"] = Lc
    Lc = _libs["count"]
    c = consts or default_consts()
    f = chunk.fields
    n, L, ncnst = chunk.ncol, chunk.pver, chunk.ncnst
    cols = range(n) if cols is None else cols
    ins = {k: np.ascontiguousarray(f[k], dtype=np.float64) for k in _IN_ORDER}
    ins["cush"] = ins["cush"].copy()
    out = {k: np.zeros((L + 1, n)) for k in OUTPUT_IFACE}
    out.update({k: np.zeros((L, n)) for k in OUTPUT_LAYER})
    out["trten"] = np.zeros((ncnst, L, n))
    for k in ("precip", "snow", "cbmf", "rliq", "cnt", "cnb"):
        out[k] = np.zeros(n)
    pin = (C.POINTER(C.c_double) * len(_IN_ORDER))(*[_p(ins[k]) for k in _IN_ORDER])
    pout = (C.POINTER(C.c_double) * len(_OUT_ORDER))(*[_p(out[k]) for k in _OUT_ORDER])
    diag = np.zeros((n, NDIAG), dtype=np.int32)
    ops = np.zeros((len(cols), len(OP_KINDS)), dtype=np.uint64)
    one = (C.c_ulonglong * len(OP_KINDS))()
    for j, col in enumerate(cols):
        rc = Lc.orc_count_column(C.byref(c), C.c_int(n), C.c_int(L), C.c_int(ncnst), C.c_double(chunk.dt),
                                 C.c_int(int(col)), pin, pout,
                                 C.cast(diag.ctypes.data + 4 * NDIAG * int(col), C.POINTER(C.c_int)), one)
        assert rc == 0, (col, rc)
        ops[j] = list(one)
    out["cush"] = ins["cush"]
    return ops, out, diag


def jvp(chunk, tangents: dict, dpar=(0.0, 0.0, 0.0), par=(10.0, 0.1, 14.0), cols=None):
    """Forward-mode derivative of the portable-math oracle (liboracle_dual.so, dual
    numbers): the outputs and their tangents along the input direction ``tangents``
    (arrays shaped like the chunk's fields, missing = zero) and the tunables' direction
    ``dpar`` (d rpen, d rle, d rkm).  Returns (outputs, output tangents, diag) for the
    columns ``cols`` (default all; other columns stay zero).  The values equal
    run(portable=True)'s bit for bit."""
    if "dual" not in _libs:
        if not os.path.exists(LIB_PATH_DUAL):
            subprocess.run(["make", "-s", "-C", HERE, "liboracle_dual.so"], check=True)
        Ld = C.CDLL(LIB_PATH_DUAL)
        Ld.orc_dual_jvp.restype = C.c_int
        _libs["dual"] = Ld
    Ld = _libs["dual"]
    f = chunk.fields
    n, L, ncnst = chunk.ncol, chunk.pver, chunk.ncnst
    cols = range(n) if cols is None else cols
    ins = {k: np.ascontiguousarray(f[k], dtype=np.float64) for k in _IN_ORDER}
    tin = {k: np.ascontiguousarray(tangents[k], dtype=np.float64) for k in _IN_ORDER if k in tangents}
    for k, v in tin.items():
        assert v.shape == ins[k].shape, k

    def zeros():
        o = {k: np.zeros((L + 1, n)) for k in OUTPUT_IFACE}
        o.update({k: np.zeros((L, n)) for k in OUTPUT_LAYER})
        o["trten"] = np.zeros((ncnst, L, n))
        for k in ("precip", "snow", "cbmf", "rliq", "cnt", "cnb"):
            o[k] = np.zeros(n)
        return o
    out, tout = zeros(), zeros()
    P = C.POINTER(C.c_double)
    pin = (P * len(_IN_ORDER))(*[_p(ins[k]) for k in _IN_ORDER])
    ptin = (P * len(_IN_ORDER))(*[(_p(tin[k]) if k in tin else None) for k in _IN_ORDER])
    pout = (P * len(_OUT_ORDER))(*[_p(out[k]) for k in _OUT_ORDER])
    ptout = (P * len(_OUT_ORDER))(*[_p(tout[k]) for k in _OUT_ORDER])
    par_a = (C.c_double * 3)(*par)
    dpar_a = (C.c_double * 3)(*dpar)
    diag = np.zeros((n, NDIAG), dtype=np.int32)
    for col in cols:
        rc = Ld.orc_dual_jvp(C.c_int(n), C.c_int(L), C.c_int(ncnst), C.c_double(chunk.dt), C.c_int(int(col)), par_a,
                             dpar_a, pin, ptin, pout, ptout,
                             C.cast(diag.ctypes.data + 4 * NDIAG * int(col), C.POINTER(C.c_int)))
        assert rc == 0, (col, rc)
    return out, tout, diag


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
