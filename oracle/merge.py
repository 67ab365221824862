"""numpy restatement of convect_shallow_tend's merge after the shallow scheme
(TEST INFRASTRUCTURE ONLY: imported by the test suite, never by the product).

Arrays are the driver's Fortran (pcols, levels) layout held in numpy as
[levels, pcols] (column fastest), CAM level order.  Every statement cites the line of
convect_shallow.F90 it follows.
"""
from __future__ import annotations

import numpy as np

FIELDS_IN = ("umf", "slflx", "qtflx", "qrten", "qsten", "qcu", "qc2", "rliq2", "cnt2", "cnb2", "cbmf", "pmid",
             "rprddp")
FIELDS_IO = ("cmfmc", "cnt", "cnb", "rliq", "qc")
FIELDS_OUT = ("icwmr", "rprdsh", "rprdtot", "cmfsl", "cmflq", "freqsh", "pcnt", "pcnb", "massflx_sh")


def merge(f: dict, ncol: int, pver: int, latvap: float) -> dict:
    """Returns the in/out and out fields after the merge (copies; ``f`` is not changed).
    Columns >= ncol keep their input values (out fields: their ``f`` values), except
    freqsh, which the driver zeroes over all pcols first, and massflx_sh, copied over
    all pcols."""
    o = {k: np.array(f[k], copy=True) for k in FIELDS_IO + FIELDS_OUT}
    n = slice(0, ncol)
    o["icwmr"][:, n] = f["qcu"][:, n]                                   # :616  icwmr = iccmr_UW
    o["rprdsh"][:, n] = f["qrten"][:, n] + f["qsten"][:, n]             # :617  rprdsh = rprdsh + cmfdqs
    o["massflx_sh"][:] = f["cbmf"][:]                                   # :623  massflx_sh(1:pcols) = cbmf(1:pcols)
    o["cmfsl"][:, n] = f["slflx"][:, n]                                 # :653  cmfsl = slflx
    o["cmflq"][:, n] = f["qtflx"][:, n] * latvap                        # :654  cmflq = qtflx * latvap
    o["freqsh"][:] = 0.0                                                # :790  freqsh(:) = 0
    mx = f["umf"][:pver, n].max(axis=0)                                 # :791-795 freqsh
    o["freqsh"][n] = np.where(mx <= 0.0, 1.0, 0.0)
    o["cmfmc"][:pver, n] = f["cmfmc"][:pver, n] + f["umf"][:pver, n]    # :806
    cnt = np.where(f["cnt2"][n] < f["cnt"][n], f["cnt2"][n], f["cnt"][n])   # :816
    cnb = np.where(f["cnb2"][n] > f["cnb"][n], f["cnb2"][n], f["cnb"][n])   # :817
    o["cnt"][n], o["cnb"][n] = cnt, cnb
    cols = np.arange(ncol)
    o["pcnt"][n] = f["pmid"][np.clip(cnt.astype(int), 1, pver) - 1, cols]   # :818 pmid(i, int(cnt(i)))
    o["pcnb"][n] = f["pmid"][np.clip(cnb.astype(int), 1, pver) - 1, cols]   # :819
    o["rprdtot"][:, n] = o["rprdsh"][:, n] + f["rprddp"][:, n]          # :822
    o["qc"][:, n] = f["qc"][:, n] + f["qc2"][:, n]                      # :829
    o["rliq"][n] = f["rliq"][n] + f["rliq2"][n]                         # :830
    return o


def random_fields(pcols: int, pver: int, seed: int = 0) -> dict:
    """Random driver fields of the right shapes and ranges (level indices 1..pver)."""
    rng = np.random.default_rng(seed)
    KS = pver + 1
    f = {}
    for k in ("umf", "slflx", "qtflx", "cmfmc", "cmfsl", "cmflq"):
        f[k] = rng.normal(0.0, 0.05, (KS, pcols))
    for k in ("qrten", "qsten", "qcu", "qc2", "rprddp", "qc", "icwmr", "rprdsh", "rprdtot"):
        f[k] = rng.normal(0.0, 1e-7, (pver, pcols))
    f["pmid"] = np.sort(rng.uniform(200.0, 101000.0, (pver, pcols)), axis=0)
    for k in ("rliq2", "cbmf", "rliq", "freqsh", "pcnt", "pcnb", "massflx_sh"):
        f[k] = rng.uniform(0.0, 1e-3, pcols)
    for k in ("cnt2", "cnb2", "cnt", "cnb"):
        f[k] = rng.integers(1, pver + 1, pcols).astype(np.float64)
    f["umf"][:, rng.random(pcols) < 0.3] = 0.0                           # columns without shallow convection
    return f
