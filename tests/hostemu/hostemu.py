"""Test harness: the product's column code compiled for the CPU (see hostemu.cpp).

Used by the CPU test tier to check the kernel's logic against the oracle without a
GPU.  Never imported by the product.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
PKG = os.path.join(ROOT, "hpc-and-climate-model_amd")
LIB_PATH = os.path.join(HERE, "libhostemu.so")
LIB_PATH_CHECKED = os.path.join(HERE, "libhostemu_bc.so")   # every workspace/array index asserted
SRC = os.path.join(HERE, "hostemu.cpp")
HDRS = [os.path.join(PKG, "csrc", f) for f in ("uwshcu_column.h", "uwshcu_thermo.h", "uwshcu_math.h")] + \
       [os.path.join(ROOT, "include", "uwshcu_types.h")]

if PKG not in sys.path:
    sys.path.insert(0, PKG)
from uwshcu_amd import _capi  # noqa: E402

NDIAG = 12


def build(force=False, checked=False):
    path = LIB_PATH_CHECKED if checked else LIB_PATH
    newest = max(os.path.getmtime(p) for p in [SRC] + HDRS)
    if not force and os.path.exists(path) and os.path.getmtime(path) >= newest:
        return
    extra = ["-O1", "-DUWSHCU_BOUNDS_CHECK"] if checked else ["-O2"]
    cmd = ["/opt/rocm/bin/hipcc", "-x", "hip", "--offload-host-only"] + extra + [
           "-ffp-contract=off", "-fopenmp", "-std=c++17", "-shared", "-fPIC", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(PKG, "csrc"), SRC, "-o", path]
    subprocess.run(cmd, check=True, capture_output=True)


_libs = {}


def lib(checked=False):
    if checked not in _libs:
        build(checked=checked)
        L = C.CDLL(LIB_PATH_CHECKED if checked else LIB_PATH)
        L.hostemu_forward.restype = C.c_int
        _libs[checked] = L
    return _libs[checked]


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def run(chunk, precision=_capi.FP64, params=None, consts=None, estbl=None, checked=False, stats=None):
    """Run the host-compiled column code on a generator chunk.

    Returns (outputs dict, diag [ncol, NDIAG], rc, bad_col) like ``oracle.run``.
    The constants and table come from the product library's host functions, which
    need no GPU.
    """
    c = consts or _capi.default_consts()
    t = estbl or _capi.build_estbl(c)
    tab = (C.c_double * 250)(*t)
    par = params or _capi.default_params()
    f = chunk.fields
    n, L, ncnst = chunk.ncol, chunk.pver, chunk.ncnst
    out = {k: np.zeros((L + 1, n)) for k in ("umf", "slflx", "qtflx", "flxprc1", "flxsnow1", "wu")}
    for k in ("qvten", "qlten", "qiten", "sten", "uten", "vten", "qrten", "qsten", "evapc", "cufrc", "qcu",
              "qlu", "qiu", "qc"):
        out[k] = np.zeros((L, n))
    out["trten"] = np.zeros((ncnst, L, n))
    for k in ("precip", "snow", "cbmf", "rliq", "cnt", "cnb"):
        out[k] = np.zeros(n)
    cush = f["cush"].copy()
    diag = np.zeros((NDIAG, n), dtype=np.int32)
    bad = C.c_int(-1)
    ins = [np.ascontiguousarray(f[k]) for k in ("ps0", "zs0", "p0", "z0", "dp0", "u0", "v0", "qv0", "ql0", "qi0",
                                                 "t0", "s0", "tr0", "tke", "pblh")]
    args = [C.byref(c), tab, C.c_int(precision), C.byref(par), C.c_int(n), C.c_int(n), C.c_int(L), C.c_int(ncnst),
            C.c_double(chunk.dt)] + [_p(a) for a in ins] + [_p(cush)]
    args += [_p(out[k]) for k in ("umf", "slflx", "qtflx", "flxprc1", "flxsnow1", "qvten", "qlten", "qiten", "sten",
                                  "uten", "vten", "trten", "qrten", "qsten", "precip", "snow", "evapc", "cufrc",
                                  "qcu", "qlu", "qiu", "cbmf", "qc", "rliq", "cnt", "cnb")]
    nfix = C.c_int(0)
    args += [_p(np.ascontiguousarray(f["dpdry0"])), _p(out["wu"]), _p(diag), C.byref(bad), C.byref(nfix)]
    rc = lib(checked).hostemu_forward(*args)
    if stats is not None:
        stats["recomputed"] = nfix.value
    out["cush"] = cush
    return out, np.ascontiguousarray(diag.T), rc, bad.value
