"""Seeded synthetic physics chunks for the UW shallow-cumulus column physics.

A chunk holds ``ncol`` columns of ``pver`` layers in the layout the physics
driver passes to ``compute_uwshcu_inv`` (uwshcu.F90:251-265, called from
convect_shallow.F90:585-600): Fortran ``(pcols, pver)`` arrays with the column
index fastest and CAM level order (index 0 = model top).  In numpy terms a
layer field is a C-contiguous ``[pver, ncol]`` array, an interface field is
``[pver+1, ncol]`` and the constituents are ``[ncnst, pver, ncol]``.

Column recipe (all random draws from one ``numpy.random.Generator(PCG64(seed))``):

* hybrid sigma-pressure interfaces from 2 hPa to the surface, ``p = A*1e5 + B*ps``
  with ``ps ~ U(98000, 102000)`` Pa;
* surface temperature ``~U(295, 302)`` K; a well-mixed PBL (dry-adiabatic lapse rate
  g/cp below ``pblh``), then 6.5 K/km to 16 km, isothermal above;
* heights from the hypsometric equation bottom-up with the virtual temperature
  (as geopotential.F90:103-110 with zero surface geopotential), ``s = cp*T + g*z``;
* relative humidity 0.85 at the surface falling linearly in pressure to 0.2 at
  500 hPa (0.2 above), plus ``N(0, 0.05)`` noise clipped to ``[0.05, 1]``;
  ``qv = RH * qs(T, p)`` with the table saturation humidity above the PBL, and the
  surface value ``0.85 * qs(Tsfc, ps)`` (capped at 98 % RH) inside the well-mixed PBL;
* ``ql, qi ~ U(0, 1e-5)``, ``u, v ~ N(0, 5)``;
* ``tke ~ U(0.1, 1)`` at interfaces below the PBL top and 0.01 above,
  ``pblh ~ U(300, 1500)`` m, ``cldfrc ~ U(0, 0.3)``, ``cush = 1000`` m;
* constituents ``~U(0, 1e-6)``;
* a fraction ``inversion_frac`` of the columns (default 0.25) is ``inversion_dt``
  (default 7 K) warmer above the PBL top, which caps convection there.  With the
  well-mixed, near-saturated PBL a 3 K cap is overcome by the deeper PBLs (about a
  third of capped columns still convect); 7 K caps all of them, so the convecting
  fraction is about ``0.96 * (1 - inversion_frac)`` and is set through the capped
  fraction.
* free-troposphere RH falls from 0.90 at the PBL top (not the surface) to 0.2 at
  500 hPa; with the surface value the lower free troposphere is dry enough that
  uncapped shallow PBLs rarely convect.

``recipe="baseline"`` is the stated column recipe of the benchmark configurations
word for word: 6.5 K/km from the surface (no well-mixed PBL), RH 0.85 at the surface
falling linearly in pressure to 0.2 at 500 hPa plus the clipped noise at every
layer, ``qv = RH * qs(T, p)`` throughout, and a +3 K inversion (``inversion_dt``
default 3) above the PBL on ``inversion_frac`` (0.25) of the columns.  The random
draws are the same as the default recipe's, in the same order.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# CESM shr_const values used by the reference physics.
_BOLTZ = 1.38065e-23
_AVOGAD = 6.02214e26
_RGAS = _AVOGAD * _BOLTZ
_MWDAIR = 28.966
_MWWV = 18.016
RAIR = _RGAS / _MWDAIR
RH2O = _RGAS / _MWWV
ZVIR = RH2O / RAIR - 1.0
GRAV = 9.80616
CPAIR = 1.00464e3
EPSILO = _MWWV / _MWDAIR
TMELT = 273.15
RH_FT_BASE = 0.90
RH_FT_BASE_CAPPED = 0.90

LAYER_FIELDS = ("p0", "z0", "dp0", "dpdry0", "u0", "v0", "qv0", "ql0", "qi0", "t0", "s0",
                "cldfrc", "concldfrc")
IFACE_FIELDS = ("ps0", "zs0", "tke")
COLUMN_FIELDS = ("pblh", "cush")


def _goff_gratch_table(tmin=173.16, tmax=375.16, trice=20.0):
    """Saturation vapour pressure table at 1 K steps (input synthesis only)."""
    n = int(tmax - tmin + 2.000001)
    t = tmin + np.arange(n, dtype=np.float64)
    ts = 373.16
    e1 = 11.344 * (1.0 - t / ts)
    e2 = -3.49149 * (ts / t - 1.0)
    f = (-7.90298 * (ts / t - 1.0) + 5.02808 * np.log10(ts / t)
         - 1.3816 * (10.0 ** e1 - 1.0) / 1e7 + 8.1328 * (10.0 ** e2 - 1.0) / 1e3
         + np.log10(1013.246))
    esw = 10.0 ** f * 100.0
    r = TMELT / t
    esi = 575.185606e10 * np.exp(-(2.01889049 / r + 3.56654 * np.log(r) + 20.947031 * r))
    w = np.clip((TMELT - t) / trice, 0.0, 1.0)
    es = np.where(t >= TMELT, esw, np.where(t < TMELT - trice, esi, w * esi + (1 - w) * esw))
    return tmin, tmax, es


_TABLE = _goff_gratch_table()


def qsat_table(t, p):
    """Saturation specific humidity from the interpolated table (vqsatd semantics)."""
    tmin, tmax, es_tab = _TABLE
    e = np.clip(t, tmin, tmax)
    i = np.floor(e - tmin).astype(np.int64)
    i = np.minimum(i, len(es_tab) - 2)
    frac = e - (tmin + i)
    es = (1.0 - frac) * es_tab[i] + frac * es_tab[i + 1]
    qs = EPSILO * es / (p - (1.0 - EPSILO) * es)
    return np.where(qs < 0.0, 1.0, np.minimum(qs, 1.0))


def hybrid_coefficients(pver: int, ptop: float = 200.0, p0ref: float = 1.0e5):
    """Interface hybrid coefficients ``A, B`` (top-down, length pver+1)."""
    s = np.linspace(0.0, 1.0, pver + 1)
    eta_top = ptop / p0ref
    shape = 0.6 * s + 0.4 * (0.5 - 0.5 * np.cos(np.pi * s))
    eta = eta_top + (1.0 - eta_top) * shape
    b = np.clip((eta - 0.1) / 0.9, 0.0, None) ** 1.2
    b[-1] = 1.0
    a = eta - b
    a[-1] = 0.0
    return a, b


@dataclass
class Chunk:
    ncol: int
    pver: int
    ncnst: int
    dt: float
    fields: dict = field(default_factory=dict)

    def __getitem__(self, name):
        return self.fields[name]

    def copy(self):
        return Chunk(self.ncol, self.pver, self.ncnst, self.dt, {k: v.copy() for k, v in self.fields.items()})


def make_chunk(ncol: int, pver: int = 30, ncnst: int = 3, seed: int = 0, dt: float = 1800.0,
               inversion_frac: float = 0.25, ptop: float = 200.0, inversion_dt: float | None = None,
               recipe: str = "pbl") -> Chunk:
    if recipe not in ("pbl", "baseline"):
        raise ValueError(f"unknown recipe {recipe!r}")
    well_mixed = recipe == "pbl"
    if inversion_dt is None:
        inversion_dt = 7.0 if well_mixed else 3.0
    rng = np.random.default_rng(np.random.PCG64(seed))
    nl, ni = pver, pver + 1
    a, b = hybrid_coefficients(pver, ptop)
    ps = rng.uniform(98000.0, 102000.0, ncol)
    pint = a[:, None] * 1.0e5 + b[:, None] * ps[None, :]          # [ni, ncol]
    pmid = 0.5 * (pint[:-1] + pint[1:])
    dp = pint[1:] - pint[:-1]
    tsfc = rng.uniform(295.0, 302.0, ncol)
    pblh = rng.uniform(300.0, 1500.0, ncol)
    inv = rng.uniform(0.0, 1.0, ncol) < inversion_frac
    rh_noise = rng.normal(0.0, 0.05, (nl, ncol))

    gamma_d = GRAV / CPAIR

    def temp_of_z(z):
        zb = pblh[None, :]
        if not well_mixed:
            t = tsfc[None, :] - 6.5e-3 * np.minimum(z, 16000.0)
            return t + np.where(inv[None, :] & (z > zb), inversion_dt, 0.0)
        t_pbl = tsfc[None, :] - gamma_d * np.minimum(z, zb)
        t = t_pbl - 6.5e-3 * (np.minimum(z, 16000.0) - np.minimum(np.minimum(z, 16000.0), zb))
        return t + np.where(inv[None, :] & (z > zb), inversion_dt, 0.0)

    def rh_of_p(p, p_bot, rh_bot):
        frac = np.clip((p_bot[None, :] - p) / (p_bot[None, :] - 50000.0), 0.0, 1.0)
        return rh_bot[None, :] + (0.2 - rh_bot[None, :]) * frac

    # free troposphere: RH falls linearly in pressure from RH_FT_BASE at the PBL top
    # (hypsometric estimate of its pressure) to 0.2 at 500 hPa
    p_pbl = ps * np.exp(-pblh / ((RAIR / GRAV) * tsfc))
    rh_base = np.where(inv, RH_FT_BASE_CAPPED, RH_FT_BASE)
    if well_mixed:
        rh = np.clip(rh_of_p(pmid, p_pbl, rh_base) + rh_noise, 0.05, 1.0)
    else:
        rh = np.clip(rh_of_p(pmid, ps, np.full(ncol, 0.85)) + rh_noise, 0.05, 1.0)
    qv_sfc = 0.85 * qsat_table(tsfc, ps)
    zi = np.zeros((ni, ncol))
    zm = np.zeros((nl, ncol))
    t = np.zeros((nl, ncol))
    qv = np.zeros((nl, ncol))
    rog = RAIR / GRAV
    for k in range(nl - 1, -1, -1):                 # bottom-up
        zb = zi[k + 1]
        zmid = zb + rog * tsfc * np.log(pint[k + 1] / pmid[k])
        for _ in range(3):
            tk = temp_of_z(zmid[None, :])[0]
            qs_k = qsat_table(tk, pmid[k])
            # well-mixed PBL: the surface specific humidity, capped at 98 % RH
            if well_mixed:
                qk = np.where(zmid < pblh, np.minimum(qv_sfc, 0.98 * qs_k), rh[k] * qs_k)
            else:
                qk = rh[k] * qs_k
            tv = tk * (1.0 + ZVIR * qk)
            zmid = zb + rog * tv * np.log(pint[k + 1] / pmid[k])
        t[k], qv[k], zm[k] = tk, qk, zmid
        zi[k] = zb + rog * tv * np.log(pint[k + 1] / pint[k])
    ql = rng.uniform(0.0, 1.0e-5, (nl, ncol))
    qi = rng.uniform(0.0, 1.0e-5, (nl, ncol))
    u = rng.normal(0.0, 5.0, (nl, ncol))
    v = rng.normal(0.0, 5.0, (nl, ncol))
    tke = np.where(zi < pblh[None, :], rng.uniform(0.1, 1.0, (ni, ncol)), 0.01)
    cld = rng.uniform(0.0, 0.3, (nl, ncol))
    concld = rng.uniform(0.0, 0.3, (nl, ncol))
    tr = rng.uniform(0.0, 1.0e-6, (ncnst, nl, ncol))
    s = CPAIR * t + GRAV * zm
    f = {
        "ps0": pint, "zs0": zi, "p0": pmid, "z0": zm, "dp0": dp, "dpdry0": dp * (1.0 - qv - ql - qi),
        "u0": u, "v0": v, "qv0": qv, "ql0": ql, "qi0": qi, "t0": t, "s0": s, "tr0": tr, "tke": tke,
        "cldfrc": cld, "concldfrc": concld, "pblh": pblh, "cush": np.full(ncol, 1000.0),
    }
    f = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in f.items()}
    return Chunk(ncol, pver, ncnst, dt, f)


def add_zw_inputs(ch: Chunk, seed: int = 1) -> Chunk:
    """Add the extra inputs of the modified scheme (compute_zwshcu_inv): layer
    pressure velocity ``omega`` [Pa/s] (CAM level order), and the PBL perturbations
    ``tpert`` [K] and ``qpert`` [kg/kg].

    omega is a per-column large-scale value U(-1.5, 0.5) (negative = ascent) plus
    N(0, 0.3) noise per layer; tpert ~ U(-0.5, 3) and qpert ~ U(-2e-4, 2e-3), so that
    some columns exceed the scheme's default caps (2 K and 1e-3)."""
    rng = np.random.default_rng(np.random.PCG64(seed + 7919))
    n, nl = ch.ncol, ch.pver
    mean = rng.uniform(-1.5, 0.5, n)
    ch.fields["omega"] = np.ascontiguousarray(mean[None, :] + rng.normal(0.0, 0.3, (nl, n)))
    ch.fields["tpert"] = rng.uniform(-0.5, 3.0, n)
    ch.fields["qpert"] = rng.uniform(-2.0e-4, 2.0e-3, n)
    return ch


def algorithmic_bytes_per_column(pver: int, ncnst: int) -> int:
    """FP64 bytes one column must move: inputs the physics reads plus the full output set.

    Inputs: 3 interface fields (ps0, zs0, tke), 10 layer fields (p0, z0, dp0, u0, v0,
    qv0, ql0, qi0, t0, s0), the constituents and 2 column scalars (pblh, cush).
    Outputs: 6 interface fields (umf, slflx, qtflx, flxprc1, flxsnow1, wu), 14 layer
    fields, the constituent tendencies and 7 column scalars (precip, snow, cbmf, rliq,
    cnt, cnb, cush).
    """
    return 8 * (9 * (pver + 1) + 24 * pver + 2 * ncnst * pver + 9)


def chunk_summary(ch: Chunk) -> str:
    return (f"ncol={ch.ncol} pver={ch.pver} ncnst={ch.ncnst} dt={ch.dt} "
            f"ptop={ch['ps0'][0].mean():.1f}Pa psfc={ch['ps0'][-1].mean():.0f}Pa "
            f"ztop={ch['zs0'][0].mean()/1000:.1f}km")
