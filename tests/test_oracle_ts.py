"""Pins of the oracle's spin-temperature step (oracle/ts.c) that do not retype its formulas.

  * homogeneous branch: x_e and T_K from scipy's natural CubicSpline of recfast_LCDM.dat
    (the GSL cspline the reference uses), T_K (1 + cT delta), and T_S between T_CMB and T_K
    (collisional coupling only);
  * shell smoothing: every smoothed shell equals a numpy float64 computation -- history
    operations, rfftn, the sharp-k window (the reference's float predicate), irfftn;
  * R_ct = 0 averages: <x_e> and <L_X> are plain means (the unfiltered round trip keeps the
    mean);
  * no sources (NO_LIGHT): adiabatic cooling dT/dz = 2 T / (1 + z) at delta = 0 and T = T_CMB
    (Compton term zero), x_e only recombines;
  * linearity: the X-ray heating and ionisation rates are linear in the X-ray source, the
    Lyman-alpha flux of the stars linear in the SFR.
"""
import numpy as np
import pytest
from scipy.interpolate import CubicSpline

import mxr
import oracle_lib as ol
from ts_cases import ZZ, inputs, params, state

TCMB = 2.728


def _tables(p, cf=2e-3, xe=2e-4):
    sf = np.zeros(len(ZZ))
    sf[: p.snapshot + 1] = cf
    return mxr.ts_build_tables(p, cf, xe, sf)


def test_homogeneous_vs_scipy_spline():
    N = 16
    p = params(N, 20.0, 0)
    d = inputs(N, 20.0, 0, 1)
    xe, Tk = state(N, d["deltax"])
    t = _tables(p)
    assert t.homogeneous == 1
    r = ol.ts_step(p, t, d["deltax"], d["stars"], d["sfr_histories"], d["xray_luminosity_histories"], 1, xe, Tk)
    tab = np.loadtxt(mxr.TABLES_DIR + "/recfast_LCDM.dat", dtype=np.float32).astype(np.float64)[::-1]
    zf = float(np.float32(ZZ[0]))
    xe_ref = np.float32(CubicSpline(tab[:, 0], tab[:, 1], bc_type="natural")(zf))
    T_ref = np.float32(CubicSpline(tab[:, 0], tab[:, 3], bc_type="natural")(zf))
    H = N // 2 + 1
    cells = r["x_e_box_prev"].reshape(N, N, 2 * H)[:, :, :N]
    assert np.allclose(cells, xe_ref, rtol=1e-6)
    dx = d["deltax"].reshape(N, N, 2 * H)[:, :, :N].reshape(-1)
    cT = np.float32(0.58 - 0.006 * (zf - 10.0))
    np.testing.assert_allclose(r["Tk_box"], T_ref * (1.0 + (cT * dx).astype(np.float64)), rtol=1e-6)
    Trad = TCMB * (1 + zf)
    lo = np.minimum(Trad, r["Tk_box"]) * (1 - 1e-6)
    hi = np.maximum(Trad, r["Tk_box"]) * (1 + 1e-6)
    assert np.all((r["TS_box"] >= lo) & (r["TS_box"] <= hi))
    assert r["globals"]["volume_ave_xe"] == pytest.approx(float(xe_ref), rel=1e-6)


def _numpy_smoothing(p, t, d, N):
    """float64 reference of the shell fields: ops in float32, rfftn, window, irfftn."""
    H = N // 2 + 1
    npad = N * N * 2 * H
    F = t.n
    hist = [d["sfr_histories"], d["xray_luminosity_histories"]]
    src = [h[:npad].copy() for h in hist]
    out = np.zeros((2, F, N ** 3))
    dk = np.float32(2.0 * np.pi / np.float32(p.BoxSize))
    n = np.arange(N)
    kx = (np.where(n > N // 2, n - N, n).astype(np.float32) * dk).astype(np.float32)
    kz = (np.arange(H).astype(np.float32) * dk).astype(np.float32)
    k2 = ((kx[:, None, None] * kx[:, None, None]).astype(np.float32) +
          (kx[None, :, None] * kx[None, :, None]).astype(np.float32)).astype(np.float32)
    k2 = (k2 + (kz[None, None, :] * kz[None, None, :]).astype(np.float32)).astype(np.float32)
    kmag = np.sqrt(k2).astype(np.float32)
    for r in range(F):
        for kind, hi, w in t.ops()[r]:
            for g in range(2):
                h = hist[g][hi * npad:(hi + 1) * npad]
                if kind == "clear":
                    src[g][:] = 0
                elif kind == "scale":
                    src[g] = (src[g].astype(np.float64) * w).astype(np.float32)
                elif kind == "load":
                    src[g] = (h * np.float32(w)).astype(np.float32)
                else:
                    src[g] = (src[g] + (h * np.float32(w)).astype(np.float32)).astype(np.float32)
        keep = np.ones_like(kmag, bool)
        if r > 0:
            kR = (kmag * np.float32(t.R[r])).astype(np.float32)
            keep = ~((kR.astype(np.float64) * 0.413566994).astype(np.float32) > 1)
        for g in range(2):
            f = src[g].reshape(N, N, 2 * H)[:, :, :N].astype(np.float64)
            s = np.fft.irfftn(np.fft.rfftn(f) * keep, s=(N, N, N), axes=(0, 1, 2))
            out[g, r] = np.maximum(s, 0.0).reshape(-1) if g == 0 else s.reshape(-1)
    return out


def test_smoothing_and_r0_means():
    N, box, snap = 16, 20.0, 14
    p = params(N, box, snap)
    depth = mxr.ts_history_depth(p, 40)
    d = inputs(N, box, snap, depth)
    xe, Tk = state(N, d["deltax"])
    t = _tables(p)
    assert t.hist_max >= 1, "the case must read older histories"
    r = ol.ts_step(p, t, d["deltax"], d["stars"], d["sfr_histories"], d["xray_luminosity_histories"], depth, xe, Tk,
                   want_smoothed=True)
    ref = _numpy_smoothing(p, t, d, N)
    sm = r["smoothed"].astype(np.float64)
    for g in range(2):
        scale = np.abs(ref[g]).max()
        assert np.abs(sm[g] - ref[g]).max() <= 2e-6 * scale
    H = N // 2 + 1
    cells = xe.reshape(N, N, 2 * H)[:, :, :N]
    assert r["red0"][1] == pytest.approx(float(cells.astype(np.float64).mean()), rel=1e-12)
    UL = p.UnitLength_in_cm
    pv = (box / p.Hubble_h / N) ** 3
    src0 = ref[1, 0].mean()
    assert r["red0"][2] == pytest.approx(src0 / pv * 1e40 * UL ** -3, rel=1e-5)


def test_no_light_adiabatic_cooling():
    N, box, snap = 16, 20.0, 14
    p = params(N, box, snap)
    depth = mxr.ts_history_depth(p, 40)
    d = inputs(N, box, snap, depth)
    for k in ("deltax", "stars", "sfr_histories", "xray_luminosity_histories"):
        d[k][:] = 0
    zf = np.float32(ZZ[snap])
    Trad = TCMB * (1.0 + float(zf))
    xe = np.zeros(mxr.padded_len(N), np.float32)
    H = N // 2 + 1
    xe.reshape(N, N, 2 * H)[:, :, :N] = 2e-4
    Tk = np.full(N ** 3, np.float32(Trad), np.float32)
    t = _tables(p, cf=0.0)
    assert t.NO_LIGHT == 1
    r = ol.ts_step(p, t, d["deltax"], d["stars"], d["sfr_histories"], d["xray_luminosity_histories"], depth, xe, Tk)
    dzp = ZZ[snap] - ZZ[snap - 1]
    expect = float(np.float32(Trad)) * (1.0 + 2.0 * dzp / (1.0 + float(zf)))
    np.testing.assert_allclose(r["Tk_box"], expect, rtol=2e-5)
    cells = r["x_e_box_prev"].reshape(N, N, 2 * H)[:, :, :N]
    assert np.all(cells < 2e-4) and np.all(cells > 1.9e-4)
    assert r["globals"]["volume_ave_J_alpha"] == 0.0 and r["globals"]["volume_ave_Xheat"] == 0.0


def test_linearity_in_sources():
    N, box, snap = 16, 20.0, 20
    p = params(N, box, snap)
    depth = mxr.ts_history_depth(p, 40)
    d = inputs(N, box, snap, depth, scale=1e-4)
    xe, Tk = state(N, d["deltax"])
    t = _tables(p)

    def run(sx, ss):
        return ol.ts_step(p, t, d["deltax"], d["stars"], d["sfr_histories"] * np.float32(ss),
                          d["xray_luminosity_histories"] * np.float32(sx), depth, xe, Tk)

    g0 = run(0.0, 0.0)["globals"]
    g1 = run(1.0, 1.0)["globals"]
    g2 = run(2.0, 1.0)["globals"]
    g3 = run(1.0, 2.0)["globals"]
    assert g1["volume_ave_Xheat"] != 0
    # X-ray heating / ionisation: linear in the X-ray source, independent of the SFR
    assert (g2["volume_ave_Xheat"] - g0["volume_ave_Xheat"]) == pytest.approx(
        2 * (g1["volume_ave_Xheat"] - g0["volume_ave_Xheat"]), rel=1e-5)
    assert (g2["volume_ave_Xion"] - g0["volume_ave_Xion"]) == pytest.approx(
        2 * (g1["volume_ave_Xion"] - g0["volume_ave_Xion"]), rel=1e-5)
    assert g3["volume_ave_Xheat"] == pytest.approx(g1["volume_ave_Xheat"], rel=1e-6)
    # J_alpha = X-ray part + stellar part: each linear in its own source
    Jx = g2["volume_ave_J_alpha"] - g1["volume_ave_J_alpha"]
    Js = g3["volume_ave_J_alpha"] - g1["volume_ave_J_alpha"]
    assert g1["volume_ave_J_alpha"] == pytest.approx(Jx + Js, rel=1e-5)
    assert Js > 0 and Jx > 0
