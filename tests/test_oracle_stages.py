"""Stage pins for the oracle's compute_uwshcu (CPU only).

The oracle records the internal quantities of its stages for one column at a time
(``oracle.probe``).  Each stage is checked here against what the mathematics fixes,
computed a different way from the restatement's closed forms:

* cloud-base closure (uwshcu.F90:2092-2184): cbmf, the updraft fraction and winv are
  the mass flux, area and mean velocity of the Gaussian vertical-velocity tail above
  the critical velocity, by numerical quadrature; the constraints on mu hold;
* buoyancy sorting (:2529-2547): xc is the first mixing fraction at which a mixture
  can no longer rise the critical distance, found by bisection on the stopping
  condition with the mixture's virtual temperature interpolated through the
  saturation point, and the saturation-point mixture matches conden;
* entrainment and detrainment rates (:2556-2571): rei times the integrals of a
  uniform mixing-fraction distribution below and above xc;
* plume update (:2592-2614): thl, qt, u, v at the layer top against RK4 integration
  of d(phi)/dp = eps (phi_e(p) - phi) (+ the pressure-gradient term for u, v) with
  the linear environment profile, and the mass flux against exp of (eps - delta);
* vertical velocity (:2700-2712) and penetration depth (compute_ppen :4981-5028,
  :2873-2899): the w^2 equation with drag and linear buoyancy integrated by RK4, and
  its first zero found by bisection;
* penetrative entrainment (:3055-3183): sign and bounds of emf, the entrained air's
  thl/qt as the mass-weighted mean of the layers it came from (a direct sum, not the
  recurrence), and exact proportionality of emf to rpen;
* precipitation evaporation (:3808-3934): never above the wet-bulb limit (:3851),
  never in saturated layers, never more than was produced.

A dropped term, a wrong sign or index, or a transposed operand in one of these
stages changes the recorded quantity by far more than the tolerances below.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from uwshcu_amd.generator import make_chunk

SQ2PI = np.sqrt(2.0 * np.pi)
CST = oracle.default_consts()


def _chunks():
    return [("pbl35", make_chunk(96, 35, 3, seed=11)),
            ("baseline30", make_chunk(96, 30, 3, seed=12, recipe="baseline", inversion_frac=0.0)),
            ("pbl72", make_chunk(48, 72, 10, seed=13, ptop=100.0))]


_CACHE = {}


def _records(portable=False):
    if portable not in _CACHE:
        recs = []
        for name, ch in _chunks():
            for c in range(ch.ncol):
                out, diag, rec = oracle.probe(ch, c, portable=portable)
                recs.append((name, ch, c, out, diag, rec))
        _CACHE[portable] = recs
    return _CACHE[portable]


def _with(key):
    rs = [r for r in _records() if r[5][key]]
    assert rs, key
    return rs


# ---------------------------------------------------------------------------
# cloud-base closure


def _gauss_tail(wc, sig):
    """P(w > wc), E[w; w > wc] for w ~ N(0, sig^2), by quadrature."""
    from scipy.integrate import quad
    pdf = lambda w: np.exp(-0.5 * (w / sig) ** 2) / (sig * SQ2PI)  # noqa: E731
    hi = wc + 40.0 * sig
    p = quad(pdf, wc, hi, epsabs=0, epsrel=1e-13, limit=200)[0]
    m = quad(lambda w: w * pdf(w), wc, hi, epsabs=0, epsrel=1e-13, limit=200)[0]
    return p, m


def test_closure_is_the_gaussian_updraft_tail():
    n = 0
    for name, ch, c, out, diag, r in _with("closure"):
        sig, mu, rho = r["sigmaw"], r["mu"], r["rho0inv"]
        wc = mu * 1.4142 * sig                      # the scheme's sqrt(2)
        p, m = _gauss_tail(wc, sig)
        tol = 2e-4 * (1.0 + mu * mu)                # 2.5066 and 1.4142 for sqrt(2 pi) and sqrt(2)
        assert r["cbmf"] == pytest.approx(rho * m, rel=tol), (name, c)
        ufrcinv = r["cbmf"] / r["winv"] / rho
        assert ufrcinv == pytest.approx(p, rel=tol), (name, c)
        assert r["winv"] == pytest.approx(m / p, rel=tol), (name, c)
        # constraints: mu only raised, at least mumin1 = 0.906, cbmf within the layer mass
        assert mu >= r["wcrit"] / sig / 1.4142 * (1 - 1e-15)
        assert mu >= 0.906
        assert r["cbmf"] <= r["cbmflimit"] * (1 + 1e-12)
        n += 1
    assert n > 100


def test_closure_keeps_the_lcl_updraft_fraction_within_rmaxfrac():
    seen = 0
    for name, ch, c, out, diag, r in _with("closure"):
        if not r["wlcl"]:
            continue
        # mass flux conserved from PBL top to LCL, energy lost to the CIN below the LCL
        assert r["ufrclcl"] * r["wlcl"] * r["rho0inv"] == pytest.approx(r["cbmf"], rel=1e-14)
        assert r["wlcl"] ** 2 == pytest.approx(r["winv"] ** 2 - 2 * r["cinlcl"] * r["rbuoy"], rel=1e-12, abs=1e-12)
        assert r["ufrclcl"] <= r["rmaxfrac"] * (1 + 5e-3), (name, c)
        if r["mumin2_applied"]:
            assert r["ufrclcl"] == pytest.approx(r["rmaxfrac"], rel=5e-3), (name, c)
            seen += 1
    assert seen >= 0


# ---------------------------------------------------------------------------
# buoyancy sorting: critical mixing fraction


def _first_stop(wue, K, thv_of_x, thv0):
    """Smallest x in [0, 1] at which a mixture of fraction x of environmental air, with
    velocity (1-x) wue and buoyancy K/2 * (thv(x)/thv0 - 1), can no longer rise the
    critical distance: (1-x)^2 wue^2 + K (thv(x) - thv0)/thv0 < 0.  Pure environment
    (x = 1) is neutral, so the condition is divided by (1 - x) before scanning."""
    def g(x):
        return (1.0 - x) * wue * wue + K * (thv_of_x(x) - thv0) / thv0 / (1.0 - x)
    xs = np.linspace(0.0, 1.0 - 1e-9, 20001)
    gv = g(xs)
    if gv[0] < 0.0:
        return 0.0
    neg = np.nonzero(gv < 0.0)[0]
    if not len(neg):
        return 1.0
    lo, hi = xs[neg[0] - 1], xs[neg[0]]
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if g(mid) >= 0.0 else (lo, mid)
    return 0.5 * (lo + hi)


def test_critical_mixing_fraction_by_bisection():
    n = {1: 0, 2: 0}
    for name, ch, c, out, diag, r in _records():
        for k in np.nonzero(r["case"])[0]:
            case = int(r["case"][k])
            wue, thvj, thv0 = r["wue"][k], r["thvj"][k], r["thv0j"][k]
            K = 2.0 * r["rbuoy"] * r["g"] * r["cridis"][k]
            if case == 1:
                thv = lambda x: (1 - x) * thvj + x * thv0  # noqa: E731
            else:
                xs, tvs = r["xsat"][k], r["thvxsat"][k]
                thv = lambda x: np.interp(x, [0.0, xs, 1.0], [thvj, tvs, thv0])  # noqa: E731
                # the scheme's quadratic on the cumulus segment, a x^2 + b x + c:
                # with both roots negative it takes xc = 0 (min of the clipped roots)
                a = wue * wue
                b = K * ((tvs - thvj) / xs) / thv0 - 2 * a
                cq = K * (thvj - thv0) / thv0 + a
                if b * b - 4 * a * cq >= 0 and b > 0 and cq > 0:
                    assert r["xc"][k] == 0.0
                    continue
            want = _first_stop(wue, K, thv, thv0)
            assert r["xc"][k] == pytest.approx(want, abs=2e-8), (name, c, k, case)
            n[case] += 1
    assert n[1] > 10 and n[2] > 500, n


def test_saturation_point_mixture_matches_conden():
    L = oracle.lib()
    D = C.c_double
    m = 0
    for name, ch, c, out, diag, r in _records():
        for k in np.nonzero(r["case"] == 2)[0]:
            x = r["xsat"][k]
            thl = r["thlue"][k] + x * (r["thle"][k] - r["thlue"][k])
            qt = r["qtue"][k] + x * (r["qte"][k] - r["qtue"][k])
            th, qv, ql, qi, rv = D(), D(), D(), D(), D()
            idc = C.c_int()
            L.orc_conden(C.byref(CST), r["pe"][k], thl, qt, C.byref(th), C.byref(qv), C.byref(ql), C.byref(qi),
                         C.byref(rv), C.byref(idc))
            thv = th.value * (1.0 + CST.zvir * qv.value - ql.value - qi.value)
            assert r["thvxsat"][k] == pytest.approx(thv, rel=1e-15)
            # xsat separates the saturated from the unsaturated mixtures
            assert 0.0 < x < 1.0 and r["excessu"][k] * r["excess0"][k] < 0.0
            m += 1
    assert m > 20


def test_entrainment_rates_integrate_a_uniform_mixing_distribution():
    from scipy.integrate import quad
    n = 0
    for name, ch, c, out, diag, r in _records():
        for k in np.nonzero(r["set"])[0]:
            xc, rei = r["xc"][k], r["rei"][k]
            ent = quad(lambda x: 2.0 * x, 0.0, xc)[0] if xc > 0 else 0.0           # mixtures kept in the updraft
            det = quad(lambda x: 2.0 * (1.0 - x), xc, 1.0)[0] if xc < 1 else 0.0   # mixtures detrained
            # (the scheme forms 1 - 2 xc + xc^2, which cancels to rounding near xc = 1)
            assert r["fer"][k] == pytest.approx(rei * ent, rel=1e-12, abs=4e-16 * rei)
            assert r["fdr"][k] == pytest.approx(rei * det, rel=1e-12, abs=4e-16 * rei)
            n += 1
    assert n > 200


# ---------------------------------------------------------------------------
# plume update across a layer


def _rk4(f, y0, s1, n=400, s0=0.0):
    """RK4 solution of y' = f(s, y) from s0 to s0 + s1 in n steps."""
    h = s1 / n
    y, s = np.array(y0, dtype=float), s0
    for _ in range(n):
        k1 = f(s, y)
        k2 = f(s + h / 2, y + h / 2 * k1)
        k3 = f(s + h / 2, y + h / 2 * k2)
        k4 = f(s + h, y + h * k3)
        y = y + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        s += h
    return y


def test_plume_update_is_the_entraining_plume_solution():
    n = 0
    for name, ch, c, out, diag, r in _records():
        for k in np.nonzero(r["set"])[0]:
            eps, dpe, pg = r["fer"][k], r["dpe"][k], r["PGFc"]
            # s = pressure distance above the layer's base; environment linear in s
            env = np.array([r["thle"][k], r["qte"][k], r["ue"][k], r["ve"][k]])
            slope = np.array([r["ssthl"][k], r["ssqt"][k], r["ssu"][k], r["ssv"][k]])
            pgf = np.array([0.0, 0.0, pg, pg])
            y0 = np.array([r["thlu0"][k], r["qtu0"][k], r["uu0"][k], r["vu0"][k]])
            got = np.array([r["thlu1"][k], r["qtu1"][k], r["uu1"][k], r["vu1"][k]])

            def rhs(s, y):
                phie = env + slope * (dpe / 2.0 - s)
                return eps * (phie - y) - pgf * slope

            want = _rk4(rhs, y0, dpe)
            scale = np.abs(y0 - env) + np.abs(slope) * dpe + np.abs(pgf * slope) * dpe
            # rounding: 400 RK4 steps (1e-12 of the value), the closed form's cancelling
            # slope/eps terms; truncation: (eps dpe)^2 / 2 of the scheme's first-order form
            rnd = 1e-12 * np.abs(want) + 1e-14 * (np.abs(env) + np.abs(y0) + np.abs(slope) / max(eps, 1e-300))
            tol = rnd + ((eps * dpe) * scale if eps * dpe < 1e-4 else 1e-9 * scale)
            assert np.all(np.abs(got - want) <= tol), (name, c, k, got - want, tol)
            # mass flux: dM/ds = (eps - delta) M
            m = _rk4(lambda s, y: (r["fer"][k] - r["fdr"][k]) * y, [r["umf0"][k]], dpe)[0]
            assert r["umf1"][k] == pytest.approx(m, rel=1e-9)
            n += 1
    assert n > 200


def _w2_rhs(D, bb, bt, dpen, rho):
    return lambda s, y: -2.0 * D * y + 2.0 * (bb + (bt - bb) * s / dpen) / rho


def test_vertical_velocity_equation_across_each_layer():
    n = 0
    for name, ch, c, out, diag, r in _records():
        for k in np.nonzero(r["set"])[0]:
            D, dpe, rho = r["drage_k"][k], r["dpe"][k], r["rho0j_k"][k]
            bb, bt, y0 = r["bogbot_k"][k], r["bogtop_k"][k], r["wtwb_k"][k]
            want = _rk4(_w2_rhs(D, bb, bt, dpe, rho), [y0], dpe)[0]
            scale = abs(y0) + dpe * (abs(bb) + abs(bt)) / rho
            tol = 1e-9 * scale if D * dpe > 1e-3 else 2.0 * D * dpe * scale + 1e-12 * scale
            assert abs(r["wtw_k"][k] - want) <= tol, (name, c, k, r["wtw_k"][k], want)
            n += 1
    assert n > 200


def test_updraft_drag_is_entrainment_plus_form_drag():
    """drage = fer (1 + rdrag) with rdrag = 1 (uwshcu.F90:1046, used at :2694): the
    vertical-velocity equation's drag is the lateral entrainment plus an equal form drag."""
    n = 0
    for name, ch, c, out, diag, r in _records():
        for k in np.nonzero(r["set"])[0]:
            assert r["drage_k"][k] == pytest.approx(2.0 * r["fer"][k], rel=1e-15, abs=0), (name, c, k)
            n += 1
    assert n > 200


def test_penetration_depth_is_the_first_zero_of_w2():
    n = 0
    for name, ch, c, out, diag, r in _with("ppen_set"):
        D, dpen, rho = r["drage"], r["dpen"], r["rho0j"]
        f = _w2_rhs(D, r["bogbot"], r["bogtop"], dpen, rho)
        steps = 2000
        h = dpen / steps
        y, s, zero = r["wtwb"], 0.0, None
        for _ in range(steps):
            k1 = f(s, y)
            k2 = f(s + h / 2, y + h / 2 * k1)
            k3 = f(s + h / 2, y + h / 2 * k2)
            k4 = f(s + h, y + h * k3)
            y1 = y + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
            if y1 <= 0.0:
                lo, hi, ylo = s, s + h, y      # bisection on RK4 solutions from the step's start
                for _ in range(60):
                    mid = 0.5 * (lo + hi)
                    ym = _rk4(f, [ylo], mid - lo, n=8, s0=lo)[0]
                    if ym > 0.0:
                        lo, ylo = mid, ym
                    else:
                        hi = mid
                zero = 0.5 * (lo + hi)
                break
            y, s = y1, s + h
        if zero is None:
            assert r["ppen"] == pytest.approx(-dpen, rel=1e-12), (name, c)
        else:
            # the reference's five fixed Newton steps leave up to 3.2e-5 of the layer depth
            # (the quadratic of the drag-free case is solved exactly)
            tol = 1e-4 * dpen if D > 0 else 1e-9 * dpen
            assert -r["ppen"] == pytest.approx(zero, rel=0, abs=tol), (name, c, r["ppen"], zero)
        n += 1
    assert n > 100


# ---------------------------------------------------------------------------
# penetrative entrainment


def test_penetrative_entrainment_is_a_mass_weighted_mixture():
    n = 0
    for name, ch, c, out, diag, r in _with("emf_set"):
        kbup, kpen = int(r["kbup"]), int(r["kpen"])
        emf = r["emf"]
        g, dt = r["g"], ch.dt
        for k in range(0, kpen + 1):
            if kbup <= k <= kpen - 1:
                assert emf[k] <= 0.0
                assert emf[k] >= -0.9 * r["dp0"][k + 1] / g / dt * (1 + 1e-12)
            else:
                assert emf[k] == 0.0
        if kpen - 1 < kbup:
            continue
        # entrained air at kpen-1: the kpen layer's value at its base interface
        top = r["thl0"][kpen] + r["ssthl0"][kpen] * (r["ps0"][kpen - 1] - r["p0"][kpen])
        assert r["thlu_emf"][kpen - 1] == pytest.approx(top, rel=1e-15)
        # below: where |emf| grows at every step the flux carries the mass-weighted
        # mean of everything it took in on the way down (direct sum, not the recurrence)
        ks = list(range(kpen - 2, kbup - 1, -1))
        if all(abs(emf[k]) > abs(emf[k + 1]) for k in ks):
            for k in ks:
                tot = top * emf[kpen - 1]
                totq = (r["qt0"][kpen] + r["ssqt0"][kpen] * (r["ps0"][kpen - 1] - r["p0"][kpen])) * emf[kpen - 1]
                for j in range(k + 1, kpen):
                    tot += r["thl0"][j] * (emf[j - 1] - emf[j])
                    totq += r["qt0"][j] * (emf[j - 1] - emf[j])
                assert r["thlu_emf"][k] == pytest.approx(tot / emf[k], rel=1e-12), (name, c, k)
                assert r["qtu_emf"][k] == pytest.approx(totq / emf[k], rel=1e-11), (name, c, k)
                n += 1
    assert n > 50


def test_penetrative_entrainment_accumulates_rpen_times_each_layers_detrainment():
    """Where no limit binds, the downward penetrative mass flux at interface k is rpen
    times the sum, over the layers j = k+1 .. kpen between k and the cloud top, of the
    mass detrained by buoyancy sorting in layer j: umf(j-1) rei(j) dp0(j), with the top
    layer's share taken over the penetration depth |ppen| (uwshcu.F90:3099, 3121).  A
    direct sum, not the recurrence."""
    n = 0
    for name, ch, c, out, diag, r in _with("emf_set"):
        if r["limit_emf"]:
            continue
        kbup, kpen, rpen = int(r["kbup"]), int(r["kpen"]), r["rpen"]
        for k in range(kpen - 1, kbup - 1, -1):
            s = r["umf"][kpen - 1] * (-r["ppen"]) * r["reif"][kpen]
            for j in range(k + 1, kpen):
                s += r["umf"][j - 1] * r["dp0"][j] * r["reif"][j]
            assert r["emf"][k] == pytest.approx(-rpen * s, rel=1e-12), (name, c, k)
            n += 1
    assert n > 50


def test_penetrative_entrainment_is_proportional_to_rpen():
    """Within one implicit-CIN pass rpen enters only the penetrative-entrainment mass
    flux, linearly: doubling it (an exact scaling in binary) doubles emf bit for bit
    wherever no limit binds, and leaves every earlier stage of the pass and the
    entrained air's properties unchanged.  Checked on columns whose result is their
    first pass (path 1: the second pass's CIN depends on the first pass's tendencies,
    and so on rpen)."""
    c2 = oracle.default_consts(rpen=20.0)
    n = 0
    for name, ch, c, out, diag, r in _with("emf_set"):
        if r["limit_emf"] or diag[7] != 1:
            continue
        _, d2, r2 = oracle.probe(ch, c, consts=c2)
        if r2["limit_emf"] or d2[7] != 1:
            continue
        for key in ("cbmf", "winv", "ppen", "kbup", "kpen"):
            assert r2[key] == r[key]
        for key in ("xc", "thlu1", "wtw_k", "umf"):
            np.testing.assert_array_equal(r2[key], r[key])
        np.testing.assert_array_equal(r2["emf"], 2.0 * r["emf"])
        np.testing.assert_array_equal(r2["thlu_emf"], r["thlu_emf"])
        n += 1
    assert n > 10


# ---------------------------------------------------------------------------
# precipitation evaporation


def test_evaporation_stays_within_the_wet_bulb_limit():
    L = oracle.lib()
    D = C.c_double
    checked = 0
    for name, ch in _chunks():
        out, diag, rc, _ = oracle.run(ch, nthreads=4)
        assert rc == 0
        mkx, dt, g = ch.pver, ch.dt, CST.g
        for i in range(ch.ncol):
            ev = out["evapc"][:, i]
            if not ev.any():
                continue
            for row in range(mkx):                  # CAM row (noevap_krelkpen is off, uwshcu.F90:1068)
                if ev[row] == 0.0:
                    continue
                assert ev[row] > 0.0
                qv, t, p = ch["qv0"][row, i], ch["t0"][row, i], ch["p0"][row, i]
                tsp, qsp, en = D(), D(), D()
                L.orc_findsp_point(C.byref(CST), qv, t, p, C.byref(tsp), C.byref(qsp), C.byref(en))
                assert ev[row] <= max(0.0, (qsp.value - qv) / dt) * (1 + 1e-12)
                es, qs, gam = D(), D(), D()
                L.orc_qsat(C.byref(CST), t, p, C.byref(es), C.byref(qs), C.byref(gam))
                assert qv < qs.value                                # only subsaturated layers evaporate
                checked += 1
            dp = ch["dp0"][:, i]
            produced = np.sum((out["qrten"][:, i] + out["qsten"][:, i]) * dp) / g
            assert np.sum(ev * dp) / g <= produced * (1 + 1e-12)
            assert out["precip"][i] >= 0.0 and 0.0 <= out["snow"][i] <= out["precip"][i] * (1 + 1e-15)
    assert checked > 50
