"""Host tables of the spin-temperature step (mxr_ts_build_tables, no GPU) against
oracle/ts_tables.py, and the Gauss-Kronrod rule the library's adaptive quadrature uses.

Pins:
  * shell radii, zpp_edge, history plan, history depth, sum_lyn (with the kink correction),
    prefactors: exact agreement with an independent restatement (same double arithmetic);
  * the tau_X = 1 frequency: within 2.5% of the exact root (the reference's Brent bracket is
    2% relative, its QAG 0.5%);
  * the heat / ionisation / Lyman-alpha frequency integrals: within 1% (the reference's QAG
    relative tolerance) of ~1e-10-accurate piecewise Gauss-Legendre quadrature;
  * GK61: exact on x^k up to degree 91 (Kronrod) and 59 (Gauss subset).
"""
import os
import re
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "oracle"))

import mxr  # noqa: E402
import ts_tables as O  # noqa: E402
from ts_cases import ZZ, params  # noqa: E402


def _pd(p):
    d = {k: getattr(p, k) for k, _ in mxr.TsParams._fields_ if k not in ("ZZ", "TablesForXHeatingDir")}
    d["ZZ"] = ZZ
    return d


def _fcoll(snap):
    sf = np.zeros(len(ZZ))
    sf[: snap + 1] = 2e-3 * (1.0 + np.arange(snap + 1)) / (snap + 1)
    return sf


def _gk61():
    src = open(os.path.join(REPO, "meraxes-devs_qyxnew_amd", "csrc", "mxr_gk61.hpp")).read()

    def arr(name):
        body = re.search(name + r"\[\d+\] = \{(.*?)\};", src, re.S).group(1)
        return np.array([float(v) for v in body.replace("\n", " ").split(",") if v.strip()])

    return arr("GK61_XGK"), arr("GK61_WGK"), arr("GK61_WG")


def test_gk61_exactness():
    xgk, wgk, wg = _gk61()
    assert len(xgk) == 31 and len(wgk) == 31 and len(wg) == 15 and xgk[30] == 0.0
    xs = np.concatenate([xgk, -xgk[:30]])
    ws = np.concatenate([wgk, wgk[:30]])
    for k in range(0, 92, 2):  # odd powers vanish by symmetry
        assert abs(np.sum(ws * xs ** k) - 2.0 / (k + 1)) < 2e-15 * 2.0, k
    xg = np.concatenate([xgk[1::2], -xgk[1::2]])
    wgg = np.concatenate([wg, wg])
    for k in range(0, 60, 2):
        assert abs(np.sum(wgg * xg ** k) - 2.0 / (k + 1)) < 4e-15, k
    # the Gauss subset is the 30-point Gauss-Legendre rule
    ref_x, ref_w = np.polynomial.legendre.leggauss(30)
    np.testing.assert_allclose(np.sort(xg), np.sort(ref_x), atol=1e-15)
    # interpolatory rule on 61 nodes: a degree-91 polynomial is integrated exactly too
    assert abs(np.sum(ws * (1 + xs) ** 91) - 2.0 ** 92 / 92) / (2.0 ** 92 / 92) < 1e-13


@pytest.mark.parametrize("N,box", [(32, 40.0), (64, 67.8), (256, 200.0)])
@pytest.mark.parametrize("snap", [1, 5, 12, 25, 40])
def test_geometry_plan_lyn_prefactors(N, box, snap):
    p = params(N, box, snap)
    pd = _pd(p)
    t = mxr.ts_build_tables(p, 2e-3, 3e-4, _fcoll(snap))
    pl = O.plan(pd)
    assert t.homogeneous == 0 and t.n == 40
    assert np.array_equal(t.arr("R"), pl["R"])
    assert np.array_equal(t.arr("zpp_edge"), pl["zpp_edge"])
    assert np.array_equal(t.arr("zpp"), pl["zpp"])
    got = t.ops()
    assert len(got) == len(pl["ops"])
    for a, b in zip(got, pl["ops"]):
        assert [(k, h) for k, h, _ in a] == [(k, h) for k, h, _ in b]
        assert [w for _, _, w in a] == [w for _, _, w in b]
    assert t.hist_max == max([h for o in pl["ops"] for k, h, _ in o if k in ("load", "add")], default=-1)
    sl = O.sum_lyn(pd, pl, mxr.TABLES_DIR)
    np.testing.assert_array_equal(t.arr("sum_lyn"), sl)
    pf = O.prefactors(pd, pl["zp"])
    for k, v in pf.items():
        assert getattr(t, k) == v, k
    assert t.dzp == pl["dzp"]


def test_history_depth():
    for N, box in [(32, 40.0), (64, 67.8), (256, 200.0), (512, 400.0)]:
        p = params(N, box, 40)
        for last in (1, 10, 40):
            assert mxr.ts_history_depth(p, last) == O.history_depth(_pd(p), last)


def test_plan_edge_cases():
    """Early snapshots: shells beyond snapshot 0 clear the source (weight = -1 path); the first
    snapshots below ReionMaxHeatingRedshift have a reweighted R_ct = 0 source."""
    for snap in (1, 2, 3):
        t = mxr.ts_build_tables(params(32, 40.0, snap), 2e-3, 3e-4, _fcoll(snap))
        kinds = [k for o in t.ops() for k, _, _ in o]
        assert "clear" in kinds
    t = mxr.ts_build_tables(params(32, 40.0, 0), 0.0, 0.0, _fcoll(0))
    assert t.homogeneous == 1


@pytest.mark.parametrize("snap", [3, 12, 40])
def test_nu_tau_one(snap):
    p = params(64, 67.8, snap)
    pd = _pd(p)
    sf = _fcoll(snap)
    t = mxr.ts_build_tables(p, 2e-3, 3e-4, sf)
    cz = O.Cosmo(pd)
    ref = np.array([O.nu_tau_one(cz, pd, t.zp, t.zpp[r], 3e-4, sf) for r in range(0, 40, 3)])
    got = t.arr("nu_tau_one")[::3]
    assert np.max(np.abs(got / ref - 1)) < 0.025


@pytest.mark.parametrize("snap", [3, 40])
def test_frequency_integrals(snap):
    p = params(64, 67.8, snap)
    pd = _pd(p)
    t = mxr.ts_build_tables(p, 2e-3, 3e-4, _fcoll(snap))
    cz = O.Cosmo(pd)
    xi = O.XInt(mxr.TABLES_DIR)
    lo_thr = pd["NuXrayThreshold"] * O.NU_over_EV
    hi = pd["NuXrayMax"] * O.NU_over_EV
    errs = []
    for r in range(0, 40, 6):
        lo = max(t.nu_tau_one[r], lo_thr)
        for m in (0, 4, 9, 11, 13):
            for flag, name in enumerate(("freq_int_heat", "freq_int_ion", "freq_int_lya")):
                ref = O.integrate_over_nu(xi, cz, pd, t.zp, float(O.X_INT_XHII[m]), lo, hi, flag)
                errs.append(getattr(t, name)[m][r] / ref - 1)
    errs = np.abs(np.array(errs))
    assert errs.max() < 0.01, errs.max()
    assert np.median(errs) < 1e-3


def test_x_int_xhii_are_floats():
    t = mxr.ts_build_tables(params(32, 40.0, 12), 2e-3, 3e-4, _fcoll(12))
    np.testing.assert_array_equal(np.ctypeslib.as_array(t.x_int_XHII), O.X_INT_XHII)


def test_bad_parameters():
    p = params(32, 40.0, 12, TsNumFilterSteps=0)
    with pytest.raises(mxr.MxrError) as e:
        mxr.ts_build_tables(p, 2e-3, 3e-4, _fcoll(12))
    assert e.value.code == mxr.MXR_ERR_INVALID_ARG
    p = params(32, 40.0, 12, TablesForXHeatingDir=b"/nonexistent")
    with pytest.raises(mxr.MxrError) as e:
        mxr.ts_build_tables(p, 2e-3, 3e-4, _fcoll(12))
    assert e.value.code == mxr.MXR_ERR_TABLES
