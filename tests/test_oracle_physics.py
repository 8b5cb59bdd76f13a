"""Pins the oracle's host physics: natural cubic spline (GSL cspline) against scipy, the
RECFAST table at its nodes, units / critical density against textbook values, RtoM and the
radius schedule of find_HII_bubbles.c:256-285."""
import ctypes
import math

import numpy as np
import pytest
from scipy.interpolate import CubicSpline

import mxr
import oracle_lib as O


def spline(x, y, q):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    return O.lib().oracle_spline_natural(x.ctypes.data, y.ctypes.data, len(x), q)


def test_natural_spline_matches_scipy(rng):
    x = np.sort(rng.uniform(0, 10, 40))
    y = np.sin(x) + 0.1 * rng.standard_normal(40)
    cs = CubicSpline(x, y, bc_type="natural")
    for q in np.concatenate([rng.uniform(x[0], x[-1], 50), x]):
        assert spline(x, y, q) == pytest.approx(float(cs(q)), rel=1e-10, abs=1e-12)


def recfast_table():
    rows = np.loadtxt(mxr.TABLES_DIR + "/recfast_LCDM.dat")
    z = rows[::-1, 0].astype(np.float32).astype(np.float64)
    tk = rows[::-1, 3].astype(np.float32).astype(np.float64)
    return z, tk


def test_T_RECFAST_nodes_and_between():
    z, tk = recfast_table()
    f = lambda q: O.lib().oracle_T_RECFAST(mxr.TABLES_DIR.encode(), q)  # noqa: E731
    for i in (3, 8, 20, 100, 499):
        assert f(float(z[i])) == pytest.approx(tk[i], rel=1e-12)
    cs = CubicSpline(z, tk, bc_type="natural")
    for q in (7.3, 8.0, 12.25, 19.9):
        assert f(q) == pytest.approx(float(cs(np.float32(q))), rel=1e-10)
    assert f(501.0) < 0
    assert O.lib().oracle_T_RECFAST(b"/nonexistent", 8.0) < 0


def test_critical_density_and_hubble_time():
    p = mxr.hii_params(64, 67.8)
    # 3 H0^2 / (8 pi G) = 2.775e11 h^2 Msun / Mpc^3 = 27.75 (1e10 Msun/h) / (Mpc/h)^3
    assert O.lib().oracle_rho_crit(ctypes.byref(p)) == pytest.approx(27.75, rel=2e-3)
    # H0 = 100 h km/s/Mpc is 100 in internal units (Mpc/h, km/s) -> t_H = 1 / (100 E(z))
    Ez = math.sqrt(p.OmegaM * 9.0 ** 3 + p.OmegaLambda)
    assert O.lib().oracle_hubble_time(ctypes.byref(p)) == pytest.approx(1.0 / (100 * Ez), rel=1e-6)


def test_RtoM_is_top_hat_mass():
    p = mxr.hii_params(64, 67.8)
    rho = O.lib().oracle_rho_crit(ctypes.byref(p))
    for R in (0.7, 5.0, 40.0):
        assert O.lib().oracle_RtoM(ctypes.byref(p), R) == pytest.approx(4 / 3 * math.pi * R ** 3 * p.OmegaM * rho,
                                                                      rel=1e-14)
    p.ReionRtoMFilterType = 1
    assert O.lib().oracle_RtoM(ctypes.byref(p), 2.0) == pytest.approx((2 * math.pi) ** 1.5 * 8 * p.OmegaM * rho,
                                                                    rel=1e-14)


def test_radius_schedule():
    p = mxr.hii_params(512, 400.0)
    r = O.radii(p)
    assert r[0] == 50.0
    assert len(r) == 44
    ratios = np.array(r[1:-1]) / np.array(r[:-2])
    np.testing.assert_allclose(ratios, 1 / 1.1, rtol=1e-14)
    # cells smaller than 1 Mpc/h: the last radius is the cell size itself
    assert r[-1] == pytest.approx(400 / 512, rel=1e-15)
    # otherwise L_FACTOR times the cell size
    assert O.radii(mxr.hii_params(64, 67.8))[-1] == pytest.approx(0.620350491 * 67.8 / 64, rel=1e-15)
    # evolving R_max at z > 6 ignores ReionRBubbleMax
    p = mxr.hii_params(64, 67.8, Flag_EvolvingReionRBubbleMax=1)
    assert O.radii(p)[0] == 25.483241248322766
