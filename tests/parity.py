"""Comparison helpers shared by the parity tests."""
from __future__ import annotations

import numpy as np

COLUMN_FIELDS = ("precip", "snow", "cbmf", "rliq", "cnt", "cnb", "cush")


def column_rel_err(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Per-column relative error of ``a`` against the reference ``b``.

    Layer/interface fields are ``[levels, ncol]`` (constituents ``[m, levels, ncol]``):
    the error of a column is max_k |a-b| / max_k |b|, with the denominator floored
    at 1e-12 of the field's global maximum so that an all-zero reference column
    must be matched to that absolute level.  Column scalars use |a-b| / max(|b|, floor).
    """
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.ndim == 3:
        a = a.reshape(-1, a.shape[-1])
        b = b.reshape(-1, b.shape[-1])
    gmax = np.abs(b).max() if b.size else 0.0
    floor = max(gmax * 1e-12, 1e-300)
    if a.ndim == 1:
        return np.abs(a - b) / np.maximum(np.abs(b), floor)
    if a.shape[0] == 0:                     # no constituents: no error per column
        return np.zeros(a.shape[-1])
    num = np.abs(a - b).max(axis=0)
    den = np.maximum(np.abs(b).max(axis=0), floor)
    return num / den


def compare(out: dict, ref: dict, fields=None) -> dict:
    """Max per-column relative error for every field present in both dicts."""
    res = {}
    for k in (fields or ref.keys()):
        if k in out and k in ref:
            res[k] = float(column_rel_err(out[k], ref[k]).max()) if np.asarray(ref[k]).size else 0.0
    return res


def reference_rms_errors(out: dict, ref: dict, cpair: float = 1.00464e3) -> dict:
    """The GPU-vs-CPU error figures printed by the reference harness
    (convect_shallow.F90.error:662-743), evaluated over all columns:
    normalised RMS of the updraft mass flux, qv tendency and s/cp tendency."""
    def rms(x, y):
        den = float((y ** 2).sum())
        return float(np.sqrt(((x - y) ** 2).sum() / den / y.size)) if den > 0 else 0.0
    return {"CMFMCrelativeerror": rms(out["umf"], ref["umf"]),
            "CMFDQrelativeerror": rms(out["qvten"], ref["qvten"]),
            "CMFDTrelativeerror": rms(out["sten"] / cpair, ref["sten"] / cpair)}
