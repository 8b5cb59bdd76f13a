"""AGN X-ray bands of the spin-temperature step (Flag_IncludeAGNXray 1-3), on the CPU: the
library's host tables and the oracle's restatement of the grid work.

Pins:
  * prefactors const_zp_prefactor_AGN_soft / _hard = (L0 / L_band) / (h nu_th) c (1 + z')^(alpha + 3)
    with L0 / L_band from quadrature of (nu / nu_th)^-alpha over each band (ComputeTs.c:1045-1128),
    alpha = 1 included; a non-finite one is an error whatever the flag (the reference aborts);
  * the bands' frequency integrals: within 1% (the reference's QAG tolerance) of the ~1e-10
    Gauss-Legendre quadrature of oracle/ts_tables.py over each shell's band limits
    (NuXrayThreshold or NuXraySoftCut to NuXraySoftCut or NuXrayMax, scaled by (1+z')/(1+z''),
    floored at the tau = 1 frequency; ComputeTs.c:835-899);
  * the flag selects the bands: a band left out has zero tables, a band present has the same
    tables whatever the other band does;
  * oracle: zero AGN grids change nothing (bit for bit against Flag 0); each band's heating term
    does not depend on the other band; with uniform AGN grids and a uniform x_e the R_ct = 0
    emissivity is e 1e10 L_sun / V_pix and each band's heating average is the closed shell sum
    sum_r dt/dz'' dz'' e W_r 1e10 L_sun / V_pix (1 + z'')^-alpha K_b(r, x_e) A_b dt/dz 2/(3 k_B (1 + x_e)),
    W_r the history weight the plan's loads and accumulations give the shell (clearing and
    rescaling leave the AGN grids alone, load_reion_bh_grids reionization.c:2799-2870).
"""
import math
import os
import sys

import numpy as np
import pytest
from scipy import integrate

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "oracle"))

import mxr  # noqa: E402
import oracle_lib as ol  # noqa: E402
import ts_tables as O  # noqa: E402
from ts_cases import SOLAR_LUM, ZZ, agn_histories, inputs, params, state  # noqa: E402

PLANCK = 6.6262e-27
NU_over_EV = 1.60217646e-12 / PLANCK
BOLTZMANN = 1.3806e-16
BANDS = ("soft", "hard")


def _fcoll(snap):
    sf = np.zeros(len(ZZ))
    sf[: snap + 1] = 2e-3 * (1.0 + np.arange(snap + 1)) / (snap + 1)
    return sf


def _pd(p):
    d = {k: getattr(p, k) for k, _ in mxr.TsParams._fields_ if k not in ("ZZ", "TablesForXHeatingDir")}
    d["ZZ"] = ZZ
    return d


def _limits(p, t, band, r):
    """A band's integration limits at shell r, as the test reads ComputeTs.c:837-871."""
    zp = t.zp
    prev = zp if r == 0 else t.zpp_edge[r - 1]
    zpp = (t.zpp_edge[r] + prev) * 0.5
    sc = (1.0 + zp) / (1.0 + zpp)
    lo_ev, hi_ev = (p.NuXrayThreshold, p.NuXraySoftCut) if band == "soft" else (p.NuXraySoftCut, p.NuXrayMax)
    return max(t.nu_tau_one[r], lo_ev * NU_over_EV * sc), hi_ev * NU_over_EV * sc


@pytest.mark.parametrize("a_soft,a_hard", [(2.2, 1.7), (1.0, 1.0), (0.6, 2.5)])
@pytest.mark.parametrize("flag", [0, 1])
def test_agn_prefactors_closed_form(a_soft, a_hard, flag):
    snap = 20
    p = params(32, 40.0, snap, Flag_IncludeAGNXray=flag, SpecIndexXrayAGNSoft=a_soft, SpecIndexXrayAGNHard=a_hard)
    t = mxr.ts_build_tables(p, 2e-3, 2e-4, _fcoll(snap))
    nu_th = p.NuXrayThreshold * NU_over_EV
    for band, alpha, (lo, hi) in (("soft", a_soft, (p.NuXrayThreshold, p.NuXraySoftCut)),
                                  ("hard", a_hard, (p.NuXraySoftCut, p.NuXrayMax))):
        L = integrate.quad(lambda nu: (nu / nu_th) ** (-alpha), lo * NU_over_EV, hi * NU_over_EV, epsrel=1e-13)[0]
        want = (1.0 / L) / PLANCK / nu_th * 2.9979e10 * (1 + t.zp) ** (alpha + 3)
        assert getattr(t, f"const_zp_prefactor_AGN_{band}") == pytest.approx(want, rel=1e-9), band


def test_non_finite_agn_prefactor_is_an_error():
    """NuXraySoftCut = NuXrayMax empties the hard band: 1 / 0 in its conversion, an error even at flag 0."""
    p = params(32, 40.0, 20, Flag_IncludeAGNXray=0, NuXraySoftCut=10000.0)
    with pytest.raises(mxr.MxrError) as e:
        mxr.ts_build_tables(p, 2e-3, 2e-4, _fcoll(20))
    assert e.value.code == mxr.MXR_ERR_INVALID_ARG


@pytest.mark.parametrize("flag", [-1, 4])
def test_bad_flag(flag):
    with pytest.raises(mxr.MxrError) as e:
        mxr.ts_build_tables(params(32, 40.0, 20, Flag_IncludeAGNXray=flag), 2e-3, 2e-4, _fcoll(20))
    assert e.value.code == mxr.MXR_ERR_INVALID_ARG


@pytest.mark.parametrize("snap", [3, 30])
def test_agn_frequency_integrals(snap):
    p = params(64, 67.8, snap, Flag_IncludeAGNXray=1)
    pd = _pd(p)
    t = mxr.ts_build_tables(p, 2e-3, 3e-4, _fcoll(snap))
    cz = O.Cosmo(pd)
    xi = O.XInt(mxr.TABLES_DIR)
    errs = []
    for band, alpha in (("soft", p.SpecIndexXrayAGNSoft), ("hard", p.SpecIndexXrayAGNHard)):
        for r in range(0, 40, 7):
            lo, hi = _limits(p, t, band, r)
            for m in (0, 5, 10, 13):
                for flag, q in enumerate(("heat", "ion", "lya")):
                    ref = O.integrate_over_nu(xi, cz, pd, t.zp, float(O.X_INT_XHII[m]), lo, hi, flag,
                                              spec_index=alpha)
                    got = getattr(t, f"freq_int_{q}_AGN_{band}")[m][r]
                    assert ref > 0
                    errs.append(got / ref - 1)
    errs = np.abs(np.array(errs))
    assert errs.max() < 0.01, errs.max()
    assert np.median(errs) < 1e-3


def test_flag_selects_bands():
    snap = 20
    tabs = {f: mxr.ts_build_tables(params(32, 40.0, snap, Flag_IncludeAGNXray=f), 2e-3, 2e-4, _fcoll(snap))
            for f in range(4)}
    names = [f"freq_int_{q}_AGN_{b}" for q in ("heat", "ion", "lya") for b in BANDS]
    for n in names:
        assert not np.any(tabs[0].arr(n)), n
    for n in names:
        band = n.rsplit("_", 1)[1]
        single = tabs[3] if band == "soft" else tabs[2]
        other = tabs[2] if band == "soft" else tabs[3]
        assert np.all(tabs[1].arr(n) > 0), n
        np.testing.assert_array_equal(tabs[1].arr(n), single.arr(n))
        assert not np.any(other.arr(n)), n
    # the galaxy tables do not depend on the flag
    for f in range(1, 4):
        np.testing.assert_array_equal(tabs[f].arr("freq_int_heat"), tabs[0].arr("freq_int_heat"))


# ---- the oracle's grid work -------------------------------------------------------------------

def _run(p, t, d, depth, xe, Tk, soft=None, hard=None, **kw):
    return ol.ts_step(p, t, d["deltax"], d["stars"], d["sfr_histories"], d["xray_luminosity_histories"], depth, xe,
                      Tk, bh_soft_histories=soft, bh_hard_histories=hard, **kw)


def _case(N, box, snap, flag, **over):
    p = params(N, box, snap, Flag_IncludeAGNXray=flag, **over)
    depth = mxr.ts_history_depth(p, 40)
    d = inputs(N, box, snap, depth, scale=1e-2)
    xe, Tk = state(N, d["deltax"])
    t = mxr.ts_build_tables(p, 2e-3, 2e-4, _fcoll(snap))
    return p, depth, d, xe, Tk, t


def test_oracle_zero_agn_grids_change_nothing():
    N, box, snap = 16, 20.0, 14
    p0, depth, d, xe, Tk, t0 = _case(N, box, snap, 0)
    p1 = params(N, box, snap, Flag_IncludeAGNXray=1)
    t1 = mxr.ts_build_tables(p1, 2e-3, 2e-4, _fcoll(snap))
    z = np.zeros_like(d["sfr_histories"])
    r0 = _run(p0, t0, d, depth, xe, Tk)
    r1 = _run(p1, t1, d, depth, xe, Tk, z, z)
    for k in ("x_e_box_prev", "Tk_box", "TS_box"):
        np.testing.assert_array_equal(r1[k], r0[k], err_msg=k)
    assert r1["globals"]["volume_ave_Xheat_AGN_soft"] == 0.0
    assert r1["globals"]["XrayEmissivity_hard"] == 0.0


def test_oracle_bands_independent():
    N, box, snap = 16, 20.0, 20
    out = {}
    for flag in (1, 2, 3):
        p, depth, d, xe, Tk, t = _case(N, box, snap, flag)
        soft, hard = agn_histories(N, d, depth)
        out[flag] = _run(p, t, d, depth, xe, Tk, soft, hard)["globals"]
    g1, g2, g3 = out[1], out[2], out[3]
    # dT_K/dz terms: heating with z falling is negative (dt/dz < 0)
    assert g1["volume_ave_Xheat_AGN_soft"] < 0 and g1["volume_ave_Xheat_AGN_hard"] < 0
    assert g1["volume_ave_Xheat_AGN_soft"] == g3["volume_ave_Xheat_AGN_soft"]
    assert g1["volume_ave_Xheat_AGN_hard"] == g2["volume_ave_Xheat_AGN_hard"]
    assert g2["volume_ave_Xheat_AGN_soft"] == 0.0 and g3["volume_ave_Xheat_AGN_hard"] == 0.0
    assert g1["XrayEmissivity_soft"] == g3["XrayEmissivity_soft"] > 0
    assert g1["XrayEmissivity_hard"] == g2["XrayEmissivity_hard"] > 0


def _weights(t):
    """Each shell's AGN grid as a multiple of a uniform history: loads set it, accumulations add,
    clearing and rescaling leave it (it starts as the current snapshot's grid, weight 1)."""
    w, out = 1.0, []
    for ops in t.ops():
        for kind, _, wt in ops:
            wf = float(np.float32(wt))
            if kind == "load":
                w = wf
            elif kind == "add":
                w = w + wf
        out.append(w)
    return out


@pytest.mark.parametrize("snap", [14, 30])
def test_oracle_uniform_agn_closed_form(snap):
    N, box = 16, 20.0
    p, depth, d, xe, Tk, t = _case(N, box, snap, 1)
    assert any(k == "load" for o in t.ops() for k, _, _ in o), "the case must reload histories"
    e = {"soft": 0.3, "hard": 0.7}
    soft = np.full_like(d["sfr_histories"], e["soft"])
    hard = np.full_like(d["sfr_histories"], e["hard"])
    x0 = 3.0e-3
    xe = np.full_like(xe, np.float32(x0))
    r = _run(p, t, d, depth, xe, Tk, soft, hard)
    g = r["globals"]
    pv = (box / p.Hubble_h / N) ** 3
    unit = 1e10 * SOLAR_LUM / pv * p.UnitLength_in_cm ** -3
    pd = _pd(p)
    cz = O.Cosmo(pd)
    W = _weights(t)
    xs = O.X_INT_XHII.astype(np.float64)
    m = int(np.searchsorted(xs, np.float32(x0), side="right") - 1)
    zpf = float(np.float32(t.zp))
    for band, alpha in (("soft", p.SpecIndexXrayAGNSoft), ("hard", p.SpecIndexXrayAGNHard)):
        assert g[f"XrayEmissivity_{band}"] == pytest.approx(e[band] * unit, rel=1e-5)
        heat = getattr(t, f"freq_int_heat_AGN_{band}")
        total = 0.0
        for rr in range(t.n):
            zpp = (t.zpp_edge[rr] + (zpf if rr == 0 else t.zpp_edge[rr - 1])) * 0.5
            dzpp = (zpf if rr == 0 else t.zpp_edge[rr - 1]) - t.zpp_edge[rr]
            K = heat[m][rr] + (heat[m + 1][rr] - heat[m][rr]) / (xs[m + 1] - xs[m]) * (x0 - xs[m])
            total += cz.dtdz(zpp) * dzpp * e[band] * W[rr] * unit * (1 + zpp) ** -alpha * K
        A = getattr(t, f"const_zp_prefactor_AGN_{band}")
        want = total * A * t.dt_dzp * 2.0 / 3.0 / BOLTZMANN / (1.0 + x0)
        assert g[f"volume_ave_Xheat_AGN_{band}"] == pytest.approx(want, rel=1e-4), band
