"""CPU pins of the oracle's construct_baryon_grids (reionization.c:2025-2591, one rank): NGP
cell assignment and its edges, float accumulation in galaxy order (two orders of the same
values round differently), clamps, the sfr conversion, the X-ray luminosity closed forms, the
black-hole mass limit and dead galaxies, the AGN grid selection and the history shift."""
import math

import numpy as np
import pytest

import mxr
import oracle_lib as O

SOLAR_MASS = 1.989e33
SEC_PER_YEAR = 3.155e7


def _gals(pos, **fields):
    n = len(pos)
    g = {k: np.zeros(n, np.float64) for k in mxr.GALAXY_FIELDS if k not in ("Pos", "Type")}
    g["Pos"] = np.ascontiguousarray(pos, np.float32).reshape(n, 3)
    g["Type"] = np.zeros(n, np.int32)
    for k, v in fields.items():
        g[k] = np.ascontiguousarray(v, np.int32 if k == "Type" else np.float64)
    return g


def _grids(N, nstore, fill=0.0):
    npad = N * N * 2 * (N // 2 + 1)
    g = {k: np.full(npad, fill, np.float32) for k in mxr.BARYON_GRIDS if "histories" not in k}
    for k in mxr.BARYON_GRIDS:
        if "histories" in k:
            g[k] = np.full(npad * nstore, fill, np.float32)
    return g


def _ip(N, ix, iy, iz):
    return iz + 2 * (N // 2 + 1) * (iy + N * ix)


def _real(x, N):
    return x[: N * N * 2 * (N // 2 + 1)].reshape(N, N, 2 * (N // 2 + 1))[:, :, :N]


def test_single_galaxy_lands_in_its_cell():
    N, box = 8, 16.0
    p = mxr.baryon_params(N, box)
    g = _gals([[3 * box / N, 5 * box / N, 7 * box / N]], FescWeightedGSM=[2.5], FescWeightedSfr=[0.75])
    out = _grids(N, 1, fill=9.0)
    O.construct_baryon_grids(p, g, out)
    s = _real(out["stars"], N)
    assert s[3, 5, 7] == np.float32(2.5) and np.count_nonzero(s) == 1
    assert out["weighted_sfr"][_ip(N, 3, 5, 7)] == np.float32(0.75)
    # stars and weighted_sfr are cleared over the padded length, padding included
    assert np.count_nonzero(out["stars"]) == 1
    # grids the flags do not build are untouched
    assert np.all(out["sfr"] == 9.0) and np.all(out["effective_bhm"] == 9.0)


@pytest.mark.parametrize("x,cell", [(0.0, 0), (1.49, 1), (1.5, 2), (2.5, 2), (3.5, 4), (7.49, 7), (7.5, 0),
                                    (7.999, 0)])
def test_pos_to_ngp_rounding_and_wrap(x, cell):
    # box = N: x / box * N = x; nearbyint rounds half to even, and N wraps to 0
    N = 8
    p = mxr.baryon_params(N, float(N))
    g = _gals([[x, 0.0, 0.0]], FescWeightedGSM=[1.0])
    out = _grids(N, 1)
    O.construct_baryon_grids(p, g, out)
    assert _real(out["stars"], N)[cell, 0, 0] == 1.0


def test_negative_position_is_an_error():
    N = 8
    p = mxr.baryon_params(N, 8.0)
    g = _gals([[-0.6, 0.0, 0.0]], FescWeightedGSM=[1.0])
    with pytest.raises(mxr.MxrError):
        O.construct_baryon_grids(p, g, _grids(N, 1))


def test_float_accumulation_follows_galaxy_order():
    # float buffer += double value, in galaxy order: [2^-24, 2^-24, 1] sums exactly to 1 + 2^-23,
    # while [1, 2^-24, 2^-24] rounds 1 + 2^-24 to even twice and stays 1
    N = 8
    p = mxr.baryon_params(N, 8.0)
    a, b = 2.0 ** -24, 1.0
    pos = [[1.0, 1.0, 1.0]] * 3
    up = _grids(N, 1)
    O.construct_baryon_grids(p, _gals(pos, FescWeightedGSM=[a, a, b]), up)
    down = _grids(N, 1)
    O.construct_baryon_grids(p, _gals(pos, FescWeightedGSM=[b, a, a]), down)
    ip = _ip(N, 1, 1, 1)
    assert up["stars"][ip] == np.float32(1 + 2.0 ** -23)
    assert down["stars"][ip] == np.float32(1.0)


def test_clamps_and_sfr_rate():
    N, t_sfr = 8, 0.37
    p = mxr.baryon_params(N, 8.0, spin_temp=True, instantaneous_sfr=False, sfr_timescale=t_sfr)
    g = _gals([[1, 1, 1], [2, 2, 2], [2, 2, 2]], FescWeightedGSM=[-1.0, 0.5, 0.25], FescWeightedSfr=[-2.0, 1, 1],
              GrossStellarMass=[-3.0, 0.3, 0.4], Sfr=[5, 5, 5])
    out = _grids(N, 2)
    O.construct_baryon_grids(p, g, out)
    assert out["stars"][_ip(N, 1, 1, 1)] == 0.0 and out["weighted_sfr"][_ip(N, 1, 1, 1)] == 0.0
    assert out["sfr"][_ip(N, 1, 1, 1)] == 0.0
    assert out["stars"][_ip(N, 2, 2, 2)] == np.float32(0.75)
    # sfr = GrossStellarMass sum / sfr_timescale (double), rounded to float
    acc = float(np.float32(float(np.float32(0.3)) + 0.4))   # float buffer += double, in double
    want = np.float32(acc / t_sfr)
    assert out["sfr"][_ip(N, 2, 2, 2)] == pytest.approx(float(want), rel=1e-7)
    assert out["sfr_histories"][_ip(N, 2, 2, 2)] == out["sfr"][_ip(N, 2, 2, 2)]


def _phys_sfr(p, sfr):
    return sfr * (p.UnitMass_in_g / SOLAR_MASS) * (SEC_PER_YEAR / p.UnitTime_in_s)


def test_xray_luminosity_closed_forms():
    N = 8
    sfr = 0.02
    # L_X = LXrayGal * SFR[Msun/yr] / 1e40 (no metallicity)
    p = mxr.baryon_params(N, 8.0, spin_temp=True)
    out = _grids(N, 1)
    O.construct_baryon_grids(p, _gals([[1, 1, 1]], Sfr=[sfr]), out)
    assert out["xray_luminosity"][_ip(N, 1, 1, 1)] == pytest.approx(3.16e40 * _phys_sfr(p, sfr) / 1e40, rel=1e-6)
    # Kyritsis et al.: log10(L_X/SFR) = B + A Z/Z_sun; alpha = 1 band fraction ln(2000/500)/ln 4 = 1
    pm = mxr.baryon_params(N, 8.0, spin_temp=True, xray_metallicity=True)
    Z = 0.004
    out = _grids(N, 1)
    O.construct_baryon_grids(pm, _gals([[1, 1, 1]], Sfr=[sfr], ColdGas=[1.0], MetalsColdGas=[Z]), out)
    lx = 10 ** (40.68 - 1.22 * Z / 0.0134)
    assert out["xray_luminosity"][_ip(N, 1, 1, 1)] == pytest.approx(lx * _phys_sfr(pm, sfr) / 1e40, rel=1e-6)
    # a flat spectrum (alpha = 0): band fraction (E_max - E_min) / (2000 - 500), which is 1/2 for
    # NuXraySoftCut = 1250 eV
    pb = mxr.baryon_params(N, 8.0, spin_temp=True, xray_metallicity=True, SpecIndexXrayGal=0.0,
                           NuXraySoftCut=1250.0)
    out = _grids(N, 1)
    O.construct_baryon_grids(pb, _gals([[1, 1, 1]], Sfr=[sfr], ColdGas=[1.0], MetalsColdGas=[Z]), out)
    assert out["xray_luminosity"][_ip(N, 1, 1, 1)] == pytest.approx(0.5 * lx * _phys_sfr(pb, sfr) / 1e40, rel=1e-6)
    # no luminosity for non-positive SFR or LXrayGal
    out = _grids(N, 1)
    O.construct_baryon_grids(mxr.baryon_params(N, 8.0, spin_temp=True, LXrayGal=0.0), _gals([[1, 1, 1]], Sfr=[1.0]),
                             out)
    assert out["xray_luminosity"][_ip(N, 1, 1, 1)] == 0.0


def test_black_hole_limit_dead_galaxies_and_agn_selection():
    N = 8
    p = mxr.baryon_params(N, 8.0, spin_temp=True, bh=True, agn_xray=mxr.MXR_AGN_XRAY_HARD,
                          BlackHoleMassLimitReion=1e-3)
    g = _gals([[1, 1, 1]] * 4, BlackHoleMass=[1e-2, 1e-4, 1e-2, 1e-2], EffectiveBHM=[1, 10, 100, 1000],
              EffectiveBHAR=[2, 20, 200, 2000], BHXrayEmissivity_hard=[3, 30, 300, 3000],
              BHXrayEmissivity_soft=[4, 40, 400, 4000], Type=[0, 1, 3, 2], FescWeightedGSM=[1, 1, 1, 1])
    out = _grids(N, 1, fill=7.0)
    below = O.construct_baryon_grids(p, g, out)
    ip = _ip(N, 1, 1, 1)
    assert below == 1                                   # galaxy 1; galaxy 2 is dead
    assert out["effective_bhm"][ip] == 1001 and out["effective_bhar"][ip] == 2002
    assert out["BHXrayEmissivity_hard"][ip] == 3003 and out["bh_xray_histories_hard"][ip] == 3003
    assert out["stars"][ip] == 3                        # dead galaxy skipped
    assert np.all(out["BHXrayEmissivity_soft"] == 7.0)  # hard only
    # the BH grids are written over real cells only: their padding keeps its contents
    pad = out["effective_bhm"].reshape(N, N, -1)[:, :, N:]
    assert np.all(pad == 7.0)
    assert np.count_nonzero(_real(out["effective_bhm"], N)) == 1


def test_history_shift():
    N, nstore = 8, 4
    npad = N * N * 2 * (N // 2 + 1)
    p = mxr.baryon_params(N, 8.0, spin_temp=True, nstore=nstore)
    out = _grids(N, nstore)
    for s in range(nstore):
        out["sfr_histories"][s * npad:(s + 1) * npad] = s + 1
        out["xray_luminosity_histories"][s * npad:(s + 1) * npad] = 10 * (s + 1)
    O.construct_baryon_grids(p, _gals([[1, 1, 1]], Sfr=[0.5]), out)
    h = out["sfr_histories"].reshape(nstore, npad)
    for s in range(1, nstore):
        assert np.all(h[s] == s)                        # slot s holds the old slot s - 1
    r0 = h[0].reshape(N, N, -1)
    assert np.all(r0[:, :, N:] == 1)                    # slot 0 padding: the old slot 0
    assert r0[1, 1, 1] == np.float32(0.5) and np.count_nonzero(r0[:, :, :N]) == 1
    hx = out["xray_luminosity_histories"].reshape(nstore, npad)
    assert np.all(hx[3] == 30)


def test_totals_conserved_on_a_random_catalogue():
    import synth
    N, box = 16, 20.0
    g = synth.galaxies(20000, box, seed=3)
    p = mxr.baryon_params(N, box, spin_temp=True, bh=True, agn_xray=mxr.MXR_AGN_XRAY_BOTH, nstore=2)
    out = synth.baryon_grids(N, 2, seed=1)
    O.construct_baryon_grids(p, g, out)
    live = g["Type"] <= 2
    assert _real(out["stars"], N).sum(dtype=np.float64) == pytest.approx(g["FescWeightedGSM"][live].sum(), rel=1e-5)
    assert _real(out["sfr"], N).sum(dtype=np.float64) == pytest.approx(g["Sfr"][live].sum(), rel=1e-5)
    assert _real(out["BHXrayEmissivity_soft"], N).sum(dtype=np.float64) == pytest.approx(
        g["BHXrayEmissivity_soft"][live].sum(), rel=1e-5)
