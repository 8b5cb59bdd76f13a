"""Pins of the oracle functions that the GPU is compared with, each against something the oracle
does not share: a documented closed form evaluated here, a textbook value, physical constants,
an independent root solve or numerical quadrature.  A plausible mistake in the restatement (an
exponent, a sign, a dropped term, a swapped argument) fails one of them.

  * fully and partially ionised gas temperatures (find_HII_bubbles.c:48-84) against the
    McQuinn & Upton Sanderbeck (2016) form of docs/formulas/igm.md (igm-ionized-temperature,
    igm-partial-temperature), and the post-loop T_K (1 + c_T delta) floor (:596-609) through the
    pipeline on uniform fields;
  * the brightness box without velocities (BrightnessTemperature.c:53-54, 94, 112-119) against
    T0 = 27 (Omega_b h^2 / 0.023) [0.15 / (Omega_m h^2) (1 + z) / 10]^(1/2) (igm-brightness);
  * get_Ts with Lyman-alpha coupling (XRayHeatingFunctions.c:1639-1684) against a bracketed root
    of the spin-temperature equation with Hirata's (2006) S_alpha, T_c,eff and tau_GP
    (igm-spin-temperature, igm-alpha-coupling, igm-alpha-correction), and its collisions-only
    closed form;
  * alpha_A against Osterbrock & Ferland's case-A coefficients, the Compton term against
    8 sigma_T a T_gamma^4 / (3 m_e c), dt/dz against -1 / ((1 + z) H(z));
  * the library's ComputeTs prefactors (ComputeTs.c:1013-1128): the X-ray luminosity conversion
    against quadrature of the power-law spectrum, dt/dz, and the growth factor and its derivative
    against the linear growth integral.
"""
import ctypes
import math

import numpy as np
import pytest
from scipy import integrate, optimize

import mxr
import oracle_lib as O
from ts_cases import ZZ, params

T_RE = 1.0e4
TCMB = 2.728           # K, the reference's constant (meraxes.h)
H100 = 3.2407789e-18   # 100 km/s/Mpc in 1/s
G_CGS = 6.672e-8
M_P = 1.6726e-24


# ---- ionised-gas temperatures ----------------------------------------------------------------

def t_ionised(z_re, z, d):
    """igm-ionized-temperature: McQuinn & Upton Sanderbeck (2016), in double."""
    if abs(z - z_re) < 1e-4:
        hist = T_RE ** 1.7
    else:
        d_r = d * (1 + z) / (1 + z_re)
        hist = (T_RE ** 1.7 * ((1 + d) / (1 + d_r)) ** 1.1333 * ((1 + z) / (1 + z_re)) ** 3.4 *
                math.exp(((1 + z) / 7.1) ** 2.5 - ((1 + z_re) / 7.1) ** 2.5))
    return (hist + (1e4 * (1 + z) / 4) ** 1.7 * (1 + d)) ** 0.5882


@pytest.mark.parametrize("z_re,z,d", [(8.0, 8.0, 0.3), (8.0, 8.0, -0.6), (8.00005, 8.0, 1.5), (10.0, 8.0, 0.5),
                                      (12.0, 6.5, -0.3), (9.0, 8.0, 2.0), (15.0, 8.0, 0.0)])
def test_fully_ionised_temperature_closed_form(z_re, z, d):
    got = O.lib().oracle_fully_ionized_T(z_re, z, d)
    want = t_ionised(np.float32(z_re), np.float32(z), np.float32(d))
    assert got == pytest.approx(want, rel=2e-6)
    assert 1e4 < got < 1e5


def test_partially_ionised_temperature():
    f = O.lib().oracle_partially_ionized_T
    assert f(50.0, 0.0) == T_RE and f(50.0, -0.1) == T_RE
    assert f(50.0, 1.0) == 50.0 and f(50.0, 1.5) == 50.0
    for T_HI, x in ((50.0, 0.25), (300.0, 0.9), (12.5, 0.5)):
        assert f(T_HI, x) == pytest.approx(x * T_HI + (1 - x) * T_RE, rel=1e-6)


def _uniform(N, stars, delta):
    H = N // 2 + 1
    d = np.zeros((N, N, 2 * H), np.float32)
    d[:, :, :N] = delta
    s = np.zeros((N, N, 2 * H), np.float32)
    s[:, :, :N] = stars
    return {"deltax": d.reshape(-1), "stars": s.reshape(-1), "weighted_sfr": (s / 10).reshape(-1)}


@pytest.mark.parametrize("TK,ionise", [(50.0, True), (1.0e6, True), (50.0, False)])
def test_post_loop_temperature_and_floor(TK, ionise):
    """Uniform fields through the pipeline: a box that ionises at this redshift gets
    max(T_ionised(z, z, delta), T_K (1 + c_T delta)); one that does not gets T_K (1 + c_T delta),
    with c_T = 0.58 - 0.006 (z - 10) (the linear adiabatic index of 2302.08506)."""
    N, box, z, delta = 16, 20.0, 8.0, 0.25
    p = mxr.hii_params(N, box, filter_type=1, zeta=1.0, redshift=z)
    p.TK_recfast = TK
    rho = O.lib().oracle_rho_crit(ctypes.byref(p))
    m = p.OmegaM * rho * (box / N) ** 3 * (1 + delta)
    inp = _uniform(N, (1.5 if ionise else 0.5) * m, delta)
    st = mxr.HiiState.host(N)
    O.find_hii_bubbles(p, inp, st)
    T = np.asarray(st.temp_kinetic_all_gas)
    floor = TK * (1 + (0.58 - 0.006 * (z - 10.0)) * delta)
    want = max(t_ionised(z, z, delta), floor) if ionise else floor
    assert np.all(np.asarray(st.xH) == (0.0 if ionise else np.asarray(st.xH)[0]))
    np.testing.assert_allclose(T, want, rtol=2e-6)


# ---- 21-cm brightness ---------------------------------------------------------------------

@pytest.mark.parametrize("spin", [False, True])
def test_brightness_base_closed_form(spin):
    N, z = 16, 8.5
    r = np.random.default_rng(4)
    H = N // 2 + 1
    xH = r.uniform(0, 1, N ** 3).astype(np.float32)
    dx = np.zeros((N, N, 2 * H), np.float32)
    dx[:, :, :N] = r.lognormal(0, 0.5, (N, N, N)) - 1
    TS = r.uniform(3, 300, N ** 3).astype(np.float32)
    p = mxr.tb_params(N, z, spin)
    dT, ave = O.brightness_temperature(p, xH, dx.reshape(-1), TS if spin else None)
    h, om = p.Hubble_h, p.OmegaM
    ob = om * p.BaryonFrac
    T0 = 27.0 * (ob * h * h / 0.023) * math.sqrt(0.15 / (om * h * h) * (1 + z) / 10)
    want = T0 * xH.astype(np.float64) * (1 + dx[:, :, :N].reshape(-1).astype(np.float64))
    if spin:
        want *= 1 - TCMB * (1 + z) / TS.astype(np.float64)
    np.testing.assert_allclose(dT, want, rtol=2e-6, atol=1e-6 * T0)
    assert ave == pytest.approx(want.mean(), rel=1e-5)


# ---- spin temperature helpers ---------------------------------------------------------------

def _cosmo(p):
    H0 = H100 * p.Hubble_h
    rho = 3 * H0 ** 2 / (8 * math.pi * G_CGS)
    nH0 = rho * p.OmegaM * p.BaryonFrac * (1 - p.Y_He) / M_P
    Hz = lambda z: H0 * math.sqrt(p.OmegaM * (1 + z) ** 3 + p.OmegaR * (1 + z) ** 4 + p.OmegaLambda)
    return Hz, nH0


def _ts_root(p, z, d, TK, xe, J, xc):
    """The spin-temperature equation solved for T_S by bracketing (igm-spin-temperature with
    Hirata's S_alpha, T_c,eff and the Gunn-Peterson depth, igm-alpha-correction)."""
    Hz, nH0 = _cosmo(p)
    Tcmb = TCMB * (1 + z)
    tau = 1.342881e-7 / Hz(z) * nH0 * (1 + z) ** 3 * (1 + d) * (1 - xe)
    xi = (1e-7 * tau / TK ** 2) ** (1 / 3)

    def xa(TS):
        S = (1 - 0.0631789 / TK + 0.115995 / TK ** 2 - 0.401403 / (TS * TK) + 0.336463 / (TS * TK ** 2)) / \
            (1 + 2.98394 * xi + 1.53583 * xi ** 2 + 3.85289 * xi ** 3)
        return 1.66e11 / (1 + z) * S * J

    def resid(TS):
        x = xa(TS)
        inv_tc = 1 / TK + 0.405535 / TK * (1 / TS - 1 / TK)
        return 1 / TS - (1 / Tcmb + x * inv_tc + xc / TK) / (1 + x + xc)

    lo, hi = 0.2 * min(Tcmb, TK), 5 * max(Tcmb, TK)
    TS = optimize.brentq(resid, lo, hi, xtol=1e-12 * hi, rtol=1e-14)
    return TS, xa(TS)


@pytest.mark.parametrize("z,d,TK,xe,J", [(10.0, 0.0, 10.0, 2e-4, 1e-11), (10.0, 0.5, 10.0, 2e-4, 1e-10),
                                         (15.0, -0.3, 30.0, 3e-4, 3e-11), (8.0, 1.0, 200.0, 1e-3, 1e-10),
                                         (20.0, 0.0, 8.0, 2e-4, 3e-11), (8.0, 0.2, 1500.0, 0.01, 5e-10)])
def test_get_Ts_solves_the_spin_temperature_equation(z, d, TK, xe, J):
    p = params(32, 40.0, 20)
    zf, df, TKf, xef, Jf = (float(np.float32(v)) for v in (z, d, TK, xe, J))
    TS, xa = O.ts_probe(p, 0, zf, df, TKf, xef, Jf)
    xc = O.ts_probe(p, 1, zf, TKf, df, xef)
    TS_root, xa_root = _ts_root(p, zf, df, TKf, xef, Jf, xc)
    assert 0.05 < xa_root  # the Lyman-alpha coupling matters in every case
    assert TS == pytest.approx(TS_root, rel=3e-3)
    assert xa == pytest.approx(xa_root, rel=3e-3)


@pytest.mark.parametrize("z,TK", [(10.0, 10.0), (20.0, 50.0), (8.0, 1000.0)])
def test_get_Ts_collisions_only(z, TK):
    p = params(32, 40.0, 20)
    TS, xa = O.ts_probe(p, 0, z, 0.3, TK, 2e-4, 0.0)
    xc = O.ts_probe(p, 1, z, TK, 0.3, 2e-4)
    Tcmb = TCMB * (1 + z)
    assert xa == 0.0
    assert TS == pytest.approx((1 + xc) / (1 / Tcmb + xc / TK), rel=1e-6)


def test_alpha_A_textbook_values():
    """Case-A hydrogen recombination coefficient (Osterbrock & Ferland 2006, table 2.1)."""
    p = params(32, 40.0, 20)
    for T, a in ((5000.0, 6.82e-13), (1e4, 4.18e-13), (2e4, 2.51e-13)):
        assert O.ts_probe(p, 2, T) == pytest.approx(a, rel=0.02)


@pytest.mark.parametrize("z,TK,xe", [(10.0, 20.0, 2e-4), (15.0, 5.0, 1e-3), (8.0, 100.0, 0.05)])
def test_compton_term_physical_constants(z, TK, xe):
    """dT/dz = -(8 sigma_T a T_gamma^4 / (3 m_e c)) x_e / (1 + f_He + x_e) (T_gamma - T_K) /
    ((1 + z) H(z)), to the 0.4% the reference's rounded constant 1.51e-4 allows."""
    p = params(32, 40.0, 20)
    Hz, _ = _cosmo(p)
    sigma_T, a_rad, m_e, c = 6.6524587e-25, 7.5657e-15, 9.1093837e-28, 2.99792458e10
    Y = p.Y_He
    f_He = (Y / 4) / ((1 - Y) + Y / 4)
    Tg = TCMB * (1 + z)
    want = -(8 * sigma_T * a_rad * Tg ** 4 / (3 * m_e * c)) * xe / (1 + f_He + xe) * (Tg - TK) / ((1 + z) * Hz(z))
    assert O.ts_probe(p, 3, z, TK, xe) == pytest.approx(want, rel=1e-2)


@pytest.mark.parametrize("z", [6.0, 8.0, 12.5, 20.0, 35.0])
def test_hubble_and_dtdz(z):
    p = params(32, 40.0, 20)
    Hz, _ = _cosmo(p)
    zf = float(np.float32(z))
    assert O.ts_probe(p, 4, zf) == pytest.approx(Hz(zf), rel=1e-9)
    assert O.ts_probe(p, 5, zf) == pytest.approx(-1 / ((1 + zf) * Hz(zf)), rel=1e-6)


# ---- the library's ComputeTs prefactors -----------------------------------------------------

def _tables(p, snap):
    sf = np.zeros(len(ZZ))
    sf[: snap + 1] = 2e-3
    return mxr.ts_build_tables(p, 2e-3, 2e-4, sf)


def _growth(p, z):
    """Linear growth D(z) / D(0) of a flat matter + Lambda universe: H(z) int_z^inf (1 + x) / H(x)^3."""
    E = lambda x: math.sqrt(p.OmegaM * (1 + x) ** 3 + p.OmegaLambda)
    D = lambda zz: E(zz) * integrate.quad(lambda x: (1 + x) / E(x) ** 3, zz, np.inf, epsabs=0, epsrel=1e-12)[0]
    return D(z) / D(0.0)


@pytest.mark.parametrize("alpha", [1.0, 0.8, 1.5])
@pytest.mark.parametrize("snap", [10, 40])
def test_ts_prefactors(alpha, snap):
    """const_zp_prefactor_GAL = (L0 / L_band) / (h nu_th) c (1 + z')^(alpha + 3), with L0 / L_band
    from quadrature of (nu / nu_th)^-alpha over [nu_th, nu_soft] (ComputeTs.c:1026-1108);
    dt/dz' = -1 / ((1 + z') H(z')); growth factor and derivative against the growth integral
    (the reference's fit is good to ~1%)."""
    p = params(32, 40.0, snap, SpecIndexXrayGal=alpha)
    t = _tables(p, snap)
    zp = t.zp
    NU_over_EV = 1.60217646e-12 / 6.62606896e-27
    PLANCK = 6.62606896e-27
    nu_th, nu_hi = p.NuXrayThreshold * NU_over_EV, p.NuXraySoftCut * NU_over_EV
    band = integrate.quad(lambda nu: (nu / nu_th) ** (-alpha), nu_th, nu_hi, epsrel=1e-13)[0]
    want = (1.0 / band) / PLANCK / nu_th * 2.99792458e10 * (1 + zp) ** (alpha + 3)
    assert t.const_zp_prefactor_GAL == pytest.approx(want, rel=1e-4)
    Hz, _ = _cosmo(p)
    assert t.dt_dzp == pytest.approx(-1 / ((1 + np.float32(zp)) * Hz(float(np.float32(zp)))), rel=1e-6)
    assert t.growth_factor_zp == pytest.approx(_growth(p, zp), rel=0.015)
    h = 1e-3
    dD = (_growth(p, zp + h) - _growth(p, zp - h)) / (2 * h)
    assert t.dgrowth_factor_dzp == pytest.approx(dD, rel=0.03)
