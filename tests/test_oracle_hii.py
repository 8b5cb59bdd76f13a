"""Pins the oracle's find_HII_bubbles restatement by invariants and closed forms that a
plausible mistake (a dropped filter, a wrong prefactor, a skipped last step, a reset z_in)
would break."""
import math

import numpy as np
import pytest

import mxr
import oracle_lib as O
from conftest import realisation


def uniform_inputs(N, stars_per_cell, delta=0.0):
    H = N // 2 + 1
    d = np.zeros((N, N, 2 * H), np.float32)
    d[:, :, :N] = delta
    s = np.zeros((N, N, 2 * H), np.float32)
    s[:, :, :N] = stars_per_cell
    return {"deltax": d.reshape(-1), "stars": s.reshape(-1), "weighted_sfr": (s / 10).reshape(-1)}


def matter_per_cell(p):
    import ctypes

    rho = O.lib().oracle_rho_crit(ctypes.byref(p))
    return p.OmegaM * rho * (p.BoxSize / p.ReionGridDim) ** 3


@pytest.mark.parametrize("ftype", [0, 1, 2])
def test_uniform_field_global_criterion(ftype):
    """Uniform fields: every filtered field is the constant itself, so the box ionises iff
    zeta * stars / (V_pix Omega_m rho_crit) > 1, and otherwise every cell gets the last step's
    partial xH = 1 - zeta * stars / (V_pix Omega_m rho_crit)."""
    N, box = 16, 20.0
    p = mxr.hii_params(N, box, filter_type=ftype, zeta=1.3)
    m = matter_per_cell(p)
    for frac, expect_ion in ((1.02, True), (0.7, False)):
        s = frac * m / p.ReionEfficiency
        st = mxr.HiiState.host(N)
        g = O.find_hii_bubbles(p, uniform_inputs(N, s), st)
        if expect_ion:
            assert np.all(st.xH == 0)
            assert np.all(st.z_at_ionization == np.float32(p.redshift))
        else:
            np.testing.assert_allclose(st.xH, 1 - frac, rtol=2e-5)
            assert np.all(st.z_at_ionization == -1)
        assert g.volume_weighted_global_xH == pytest.approx(float(st.xH.astype(np.float64).mean()), rel=1e-12)


def test_J21_at_ionisation_closed_form():
    """UVBFlag 1: J21 = stars_f * J_21_aux_constant at the first (largest) ionising radius,
    J_21_aux_constant = Gamma_R_prefactor h_P 1e21 / (4 pi) b_Gamma / t_sfr alpha_UV."""
    import ctypes

    N, box = 16, 20.0
    p = mxr.hii_params(N, box, filter_type=1)
    s = 1.5 * matter_per_cell(p)
    st = mxr.HiiState.host(N)
    O.find_hii_bubbles(p, uniform_inputs(N, s), st)
    R = O.radii(p)[0]
    UL, UM = p.UnitLength_in_cm, p.UnitMass_in_g
    UT = UL / p.UnitVelocity_in_cm_per_s
    vpix = (box / N) ** 3
    G = (1 + p.redshift) ** 2 * R * UL / vpix * UL ** -3 * (UM / UT) / 1.6726e-24 * p.ReionNionPhotPerBary
    t_sfr = p.ReionSfrTimescale * O.lib().oracle_hubble_time(ctypes.byref(p))
    J = np.float32(G * 6.6262e-27 * 1e21 / (4 * math.pi) * p.ReionGammaHaloBias / t_sfr * p.ReionAlphaUV)
    np.testing.assert_allclose(st.J_21_at_ionization, np.float32(s) * J, rtol=2e-6)


def test_state_ranges_and_monotone_in_efficiency():
    N, box = 32, 40.0
    inp = realisation(N, box, 1)
    prev_ion = None
    prev_x = 1.1
    for zeta in (0.5, 1.0, 2.0):
        st = mxr.HiiState.host(N)
        g = O.find_hii_bubbles(mxr.hii_params(N, box, zeta=zeta), inp, st)
        assert st.xH.min() >= 0 and st.xH.max() <= 1
        assert set(np.unique(st.z_at_ionization)) <= {-1.0, 8.0}
        assert g.volume_weighted_global_xH <= prev_x
        ion = st.xH == 0
        if prev_ion is not None:
            assert np.all(ion[prev_ion])
        prev_ion, prev_x = ion, g.volume_weighted_global_xH


def test_z_at_ionization_persists():
    N, box = 32, 40.0
    inp = realisation(N, box, 1)
    st = mxr.HiiState.host(N)
    O.find_hii_bubbles(mxr.hii_params(N, box, redshift=9.0, zeta=0.5), inp, st)
    first = st.z_at_ionization.copy()
    O.find_hii_bubbles(mxr.hii_params(N, box, redshift=8.0, zeta=1.0), inp, st)
    kept = first > 0
    assert kept.any()
    np.testing.assert_array_equal(st.z_at_ionization[kept], first[kept])
    assert (st.z_at_ionization == 8.0).any()


def test_point_source_bubble_matches_numpy_pipeline():
    """One bright cell in a uniform medium, sharp-k filter.  The ionised set equals the union
    over radii of {cells where the sharp-k-smoothed stars field (computed here with numpy's
    FFT in float64) times the prefactor exceeds 1}, except cells at the threshold."""
    N, box = 32, 40.0
    p = mxr.hii_params(N, box, filter_type=1)
    m = matter_per_cell(p)
    inp = uniform_inputs(N, 0.0)
    s3 = inp["stars"].reshape(N, N, 2 * (N // 2 + 1))
    s3[5, 7, 9] = 400 * m
    st = mxr.HiiState.host(N)
    _, margin = O.find_hii_bubbles(p, inp, st, margin=True)
    field = np.zeros((N, N, N))
    field[5, 7, 9] = np.float32(400 * m)
    F = np.fft.rfftn(field)
    n = np.arange(N)
    k1 = np.where(n > N // 2, n - N, n) * 2 * np.pi / box
    kz = np.arange(N // 2 + 1) * 2 * np.pi / box
    kmag = np.sqrt(k1[:, None, None] ** 2 + k1[None, :, None] ** 2 + kz[None, None, :] ** 2)
    ion = np.zeros((N, N, N), bool)
    radii = O.radii(p)
    for i, R in enumerate(radii):
        mask = (0.413566994 * kmag * R <= 1) if i < len(radii) - 1 else np.ones_like(kmag, bool)
        sm = np.maximum(np.fft.irfftn(F * mask, s=(N, N, N), axes=(0, 1, 2)), 0)
        pref = p.ReionEfficiency / (p.OmegaM * O.lib().oracle_rho_crit(__import__("ctypes").byref(p))
                                    * (box / N) ** 3)
        ion |= sm * pref > 1
    got = (st.xH == 0).reshape(N, N, N)
    diff = got != ion
    assert ion.sum() > 10
    assert np.all(np.abs(margin.reshape(N, N, N)[diff]) < 1e-4)


def test_fp32_restatement_against_fp64_oracle():
    N, box = 64, 67.8
    inp = realisation(N, box, 1)
    a, b = mxr.HiiState.host(N), mxr.HiiState.host(N)
    ga, margin = O.find_hii_bubbles(mxr.hii_params(N, box), inp, a, margin=True)
    gb = O.find_hii_bubbles(mxr.hii_params(N, box), inp, b, precision="f64")
    mism = (a.xH == 0) != (b.xH == 0)
    assert mism.sum() <= max(2, 5e-4 * N ** 3)
    assert np.all(np.abs(margin[mism]) < 1e-3)
    assert abs(ga.volume_weighted_global_xH - gb.volume_weighted_global_xH) < 1e-5 + mism.sum() / N ** 3


def test_recombinations_need_their_grids():
    N = 16
    p = mxr.hii_params(N, 20.0, Flag_IncludeRecombinations=1)
    with pytest.raises(mxr.MxrError) as e:
        O.find_hii_bubbles(p, uniform_inputs(N, 1.0), mxr.HiiState.host(N))
    assert e.value.code == mxr.MXR_ERR_INVALID_ARG


# ---- BH feedback (Flag_BHFeedback): effective_bhm and effective_bhar as filtered grids ----

def _bh_inputs(N, stars, bhm, bhar=0.0):
    d = uniform_inputs(N, stars)
    H = N // 2 + 1
    for k, v in (("effective_bhm", bhm), ("effective_bhar", bhar)):
        a = np.zeros((N, N, 2 * H), np.float32)
        a[:, :, :N] = v
        d[k] = a.reshape(-1)
    return d


@pytest.mark.parametrize("ftype", [0, 1])
def test_bh_uniform_field_criterion_and_partial_xH(ftype):
    """BH mass enters the criterion beside the stars: a uniform box ionises iff
    zeta (stars + bhm) / (V_pix Omega_m rho_crit) > 1, with BH sources alone as well, and the
    last step's partial xH is 1 - zeta (stars + bhm) / (V_pix Omega_m rho_crit)."""
    N, box = 16, 20.0
    p = mxr.hii_params(N, box, filter_type=ftype, zeta=1.3, bh=True)
    m = matter_per_cell(p) / p.ReionEfficiency
    for stars, bhm, expect_ion in ((0.0, 1.02 * m, True), (0.6 * m, 0.45 * m, True), (0.4 * m, 0.3 * m, False)):
        st = mxr.HiiState.host(N)
        O.find_hii_bubbles(p, _bh_inputs(N, stars, bhm), st)
        if expect_ion:
            assert np.all(st.xH == 0)
        else:
            np.testing.assert_allclose(st.xH, 1 - (stars + bhm) / m, rtol=3e-5)
            assert np.all(st.z_at_ionization == -1)


def test_bh_J21_adds_the_bh_term_with_its_own_index():
    """J21 at the first crossing = stars_f J_aux(alpha_UV) + bhm_f J_aux(alpha_UV,BH): doubling
    alpha_UV,BH doubles the BH part only."""
    N, box = 16, 20.0
    m = matter_per_cell(mxr.hii_params(N, box))
    s, b = 0.9 * m, 0.6 * m
    res = {}
    for a_bh in (2.0, 4.0):
        p = mxr.hii_params(N, box, filter_type=1, bh=True, ReionAlphaUVBH=a_bh)
        st = mxr.HiiState.host(N)
        O.find_hii_bubbles(p, _bh_inputs(N, s, b), st)
        assert np.all(st.xH == 0)
        res[a_bh] = float(st.J_21_at_ionization[0])
    p0 = mxr.hii_params(N, box, filter_type=1)
    st = mxr.HiiState.host(N)
    O.find_hii_bubbles(p0, uniform_inputs(N, s + b), st)  # alpha_UV = 2: the same per-mass constant
    j_all = float(st.J_21_at_ionization[0])
    assert res[2.0] == pytest.approx(j_all, rel=1e-6)
    j_star = j_all * s / (s + b)
    assert (res[4.0] - j_star) == pytest.approx(2 * (res[2.0] - j_star), rel=1e-5)


def test_bh_global_effective_bhar_and_split_equivalence():
    """volume-weighted effective_bhar is the unfiltered grid's mean in Msun/yr; moving part of
    the stellar mass into effective_bhm (same alpha_UV) leaves the flags unchanged up to the
    threshold cells and xH within rounding."""
    N, box = 32, 40.0
    inp = realisation(N, box, 1)
    p = mxr.hii_params(N, box, filter_type=1)
    pb = mxr.hii_params(N, box, filter_type=1, bh=True)
    a, b = mxr.HiiState.host(N), mxr.HiiState.host(N)
    ga, margin = O.find_hii_bubbles(p, inp, a, margin=True)
    split = dict(inp)
    split["stars"] = (inp["stars"] * np.float32(0.75)).astype(np.float32)
    split["effective_bhm"] = (inp["stars"] - split["stars"]).astype(np.float32)
    split["effective_bhar"] = (inp["weighted_sfr"] * np.float32(0.3)).astype(np.float32)
    gb = O.find_hii_bubbles(pb, split, b)
    mism = (a.xH == 0) != (b.xH == 0)
    assert mism.sum() <= max(2, 5e-4 * N ** 3)
    assert np.all(np.abs(margin[mism]) < 1e-3)
    ok = ~mism
    assert np.abs(a.xH[ok] - b.xH[ok]).max() < 1e-4
    UT = p.UnitLength_in_cm / p.UnitVelocity_in_cm_per_s
    H = N // 2 + 1
    bhar = split["effective_bhar"].reshape(N, N, 2 * H)[:, :, :N].astype(np.float64).mean()
    want = bhar * p.UnitMass_in_g / UT * 3.155e7 / 1.989e33
    assert gb.volume_weighted_global_effective_bhar == pytest.approx(want, rel=1e-12)
    assert ga.volume_weighted_global_effective_bhar == 0.0
