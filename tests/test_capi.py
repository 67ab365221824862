"""The C ABI library (CPU only: loads and host-side functions, no GPU calls)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle
from uwshcu_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names += re.findall(r"^\s*(?:int|void|size_t)\s+\**(\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = _capi.lib()
    declared = _declared_functions()
    assert "uwshcu_forward_device" in declared and "bxcompute_uwshcu_inv_" in declared
    for name in declared:
        assert hasattr(lib, name), name
    assert set(_capi.SYMBOLS) <= set(declared)


def test_struct_layout_matches_header():
    # offsets the HIP side relies on: the ctypes mirror must match the C struct
    assert C.sizeof(_capi.Consts) == 8 * 19 + 4 * 4 + 8 * 5 + 8 * 33 + 4 * 33 + 4
    assert C.sizeof(_capi.Params) == 24
    # uwshcu_chunk: two ints, then 44 array pointers
    assert C.sizeof(_capi.ChunkDesc) == 8 + 8 * 44
    assert _capi.ChunkDesc.ps0.offset == 8 and _capi.ChunkDesc.cnb.offset == 8 + 8 * 43
    # uwshcu_zw_params: six ints, then seven doubles (same layout as the oracle's orc_zw)
    assert C.sizeof(_capi.ZwParams) == 4 * 6 + 8 * 7 == C.sizeof(oracle.ZW)
    assert _capi.ZwParams.fwcrit.offset == 24 and _capi.ZwParams.qpertmax.offset == 72


def test_zw_defaults_equal_the_oracle_defaults():
    par, z = _capi.zw_default_params()
    assert (par.rpen, par.rle, par.rkm) == (15.0, 0.1, 10.0)
    o = oracle.zw_defaults()
    for name, _ in _capi.ZwParams._fields_:
        assert getattr(z, name) == getattr(o, name), name


def test_default_constants_equal_the_oracle():
    c = _capi.default_consts()
    o = oracle.default_consts()
    for f, _ in _capi.Consts._fields_:
        if f in ("pad_", "pcf", "qmin", "wet", "ixnumliq", "ixnumice"):
            continue
        if hasattr(o, f):
            assert getattr(c, f) == getattr(o, f), f
    assert list(c.pcf) == list(o.pcf)
    assert list(c.qmin) == list(o.qmin) and list(c.wet) == list(o.wet)
    assert (c.ixnumliq, c.ixnumice) == (o.ixnumliq, o.ixnumice) == (4, 5)


def test_table_builder_equals_the_oracle_table():
    t = np.array(_capi.build_estbl())
    np.testing.assert_array_equal(t, np.array(oracle.default_consts().estbl))


def test_workspace_size():
    # an FP64 environment block (23 arrays) and a plume block of 29 + 4*ncnst arrays,
    # each of pver+1 levels; the plume block is FP32 in the FP32-plume pass only
    fp64 = _capi.workspace_bytes_per_column(35, 3, _capi.FP64)
    assert fp64 == 36 * 8 * (23 + 29 + 12)
    assert _capi.workspace_bytes_per_column(35, 3, _capi.MIXED) == fp64
    assert _capi.workspace_bytes_per_column(35, 3, _capi.MIXED_FAST_UNCHECKED) == fp64
    assert _capi.workspace_bytes_per_column(35, 3, _capi.MIXED_UNCHECKED) == 36 * (8 * 23 + 4 * (29 + 12))


def test_device_entry_points_reject_bad_arguments_without_a_gpu():
    lib = _capi.lib()
    par = _capi.default_params()
    nul = [C.c_void_p(0)] * len(_capi.ARRAY_ARGS)
    rc = lib.uwshcu_forward_device(None, 0, C.byref(par), 10, 10, 30, 3, C.c_double(1800.0), *nul, None, None)
    assert rc == -10
    rc = lib.uwshcu_forward_host(None, 0, C.byref(par), 10, 10, 30, 3, C.c_double(1800.0), *nul, None, None)
    assert rc == -10
    _, z = _capi.zw_default_params()
    nulz = [C.c_void_p(0)] * len(_capi.ZW_ARRAY_ARGS)
    rc = lib.uwshcu_zw_forward_device(None, C.byref(par), C.byref(z), 10, 10, 30, 3, C.c_double(1800.0), *nulz,
                                      None, None)
    assert rc == -10
    rc = lib.uwshcu_zw_forward_host(None, C.byref(par), C.byref(z), 10, 10, 30, 3, C.c_double(1800.0), *nulz,
                                    None, None)
    assert rc == -10
    desc = (_capi.ChunkDesc * 2)()
    rc = lib.uwshcu_forward_chunks_device(None, 0, C.byref(par), 2, desc, 30, 3, C.c_double(1800.0), None)
    assert rc == -10


def test_product_fails_loudly_without_the_library(monkeypatch, tmp_path):
    monkeypatch.setattr(_capi, "_lib", None)
    monkeypatch.setattr(_capi, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(RuntimeError, match="missing"):
        _capi.lib()
