"""Pins of the oracle's recombination pieces (oracle/recomb.c) against what it does not share:
the ionisation-equilibrium equation that neutral_fraction solves, the published Rahmati et al.
(2013, eq. 14) self-shielding fit at its characteristic density, the MHR00 PDF's Gaussian in
Delta^(-2/3), scipy's natural cubic spline (CubicSpline, bc_type="natural") for C(z), beta(z) and
the per-cell table lookup, the hydrogen density from rho_crit = 3 H^2 / (8 pi G), and
dt/dz = -1 / ((1 + z) H(z)).  The integrands' products are pinned through ratios that isolate
each factor (n_H alpha_B, x_e, Delta, x_HI).

Reference: src/core/recombinations.c:35-82, 397-504, 561-766; src/core/reionization.h:15-31;
src/core/XRayHeatingFunctions.c:426-440.
"""
import math

import numpy as np
import pytest
from scipy.interpolate import CubicSpline

import mxr
import oracle_lib as O

H100 = 3.2407789e-18   # 100 km/s/Mpc in 1/s
G_CGS = 6.672e-8
M_P = 1.6726e-24
ALPHA_B_1E4 = 2.59e-13  # Osterbrock, case B at 1e4 K


@pytest.fixture(scope="module")
def p():
    return mxr.hii_params(32, 20.0, Flag_IncludeRecombinations=1, Y_He=0.24)


def corr_he(Y):
    return 1.0 / (4.0 / Y - 3)  # n_He / n_H


@pytest.mark.parametrize("n,T4,g12", [(1e-3, 1.0, 1.0), (1e-1, 0.5, 0.3), (3e-2, 2.0, 1e-3), (1.0, 1.0, 1e-4)])
def test_neutral_fraction_solves_ionisation_equilibrium(p, n, T4, g12):
    """(1 + n_He/n_H) n_H alpha_B(T) (1 - x)^2 = Gamma x: recombinations onto ionised hydrogen with
    the electrons of hydrogen and singly ionised helium balance photoionisations."""
    x = O.rec_probe(p, 0, n, T4, g12)
    assert 1e-5 < x < 1
    alpha = ALPHA_B_1E4 * T4 ** -0.75
    lhs = (1 + corr_he(0.24)) * n * alpha * (1 - x) ** 2
    rhs = g12 * 1e-12 * x
    assert lhs == pytest.approx(rhs, rel=1e-9)


def test_neutral_fraction_limits(p):
    # chi = (1 + c) n alpha / Gamma below 1e-5 is returned as is, and agrees with the root there
    n, T4 = 1e-4, 1.0
    alpha = ALPHA_B_1E4
    g12 = (1 + corr_he(0.24)) * n * alpha / 5e-6 / 1e-12  # chi = 5e-6
    x = O.rec_probe(p, 0, n, T4, g12)
    assert x == pytest.approx(5e-6, rel=1e-12)
    # the exact root of x^2 - (2 + 1/chi) x + 1 = 0 differs from chi by O(chi^2)
    b = 2 + 1 / 5e-6
    root = (b - math.sqrt(b * b - 4)) / 2
    assert x == pytest.approx(root, rel=2e-5)
    # below TINY = 1e-30: exactly zero
    assert O.rec_probe(p, 0, 1e-30, 1.0, 1e10) == 0.0


def test_gamma_ss_rahmati_fit(p):
    """Gamma / Gamma_bg = 0.98 (1 + (D/D_ss)^1.64)^-2.28 + 0.02 (1 + D/D_ss)^-0.84 with
    D_ss = 26.7 T4^0.17 ((1 + z)/10)^-3 Gamma_12^(2/3): at T4 = 1, z = 9, Gamma_12 = 1 the
    characteristic overdensity is 26.7."""
    at_ss = O.rec_probe(p, 1, 1.0, 26.7, 1.0, 9.0)
    assert at_ss == pytest.approx(0.98 * 2 ** -2.28 + 0.02 * 2 ** -0.84, rel=1e-12)
    assert O.rec_probe(p, 1, 0.7, 1e-9, 1.0, 9.0) == pytest.approx(0.7, rel=1e-9)  # no shielding at low D
    # D_ss scales as T4^0.17 (1 + z)^-3 Gamma^(2/3): the same ratio at the scaled density
    T4, z, g = 2.0, 6.0, 0.1
    d_ss = 26.7 * T4 ** 0.17 * ((1 + z) / 10) ** -3 * g ** (2 / 3)
    assert O.rec_probe(p, 1, g, d_ss, T4, z) / g == pytest.approx(at_ss, rel=1e-12)
    # deep in the shielded regime the 2% term dominates: ratio -> 0.02 (D/D_ss)^-0.84
    r = O.rec_probe(p, 1, 1.0, 26.7 * 1e4, 1.0, 9.0)
    assert r == pytest.approx(0.02 * (1 + 1e4) ** -0.84 + 0.98 * (1 + 1e4 ** 1.64) ** -2.28, rel=1e-12)


C_NODES = [0.558, 0.599, 0.611, 0.769, 0.868, 0.930, 0.964, 0.983, 0.993, 0.998, 0.999, 1.00]
BETA_NODES = [-2.23, -2.35, -2.48, -2.49, -2.50]


def test_mhr_parameters_natural_splines(p):
    cs = CubicSpline(np.arange(12) + 2.0, C_NODES, bc_type="natural")
    bs = CubicSpline(np.arange(5) + 2.0, BETA_NODES, bc_type="natural")
    for z in np.linspace(2.05, 12.95, 37):
        c, b = O.rec_probe(p, 2, z)
        assert c == pytest.approx(float(cs(z)), abs=1e-13)
        if z < 6.0:
            assert b == pytest.approx(float(bs(z)), abs=1e-13)
        else:
            assert b == -2.50
    for i, v in enumerate(C_NODES[1:-1]):
        assert O.rec_probe(p, 2, 3.0 + i)[0] == pytest.approx(v, abs=1e-14)
    assert O.rec_probe(p, 2, 1.0) == (0.558, -2.23)
    assert O.rec_probe(p, 2, 13.0)[0] == 1.0 and O.rec_probe(p, 2, 40.0) == (1.0, -2.50)


def test_mhr_pdf_gaussian_in_delta_power(p):
    """MHR00: P(Delta) = A exp(-(Delta^(-2/3) - C0)^2 / (2 (2 x 7.61 / (3 (1 + z)))^2)) Delta^beta."""
    z = 8.0
    c0, beta = O.rec_probe(p, 2, z)
    s = 2 * 7.61 / (3 * (1 + z))
    D0 = c0 ** -1.5                           # the Gaussian's centre
    assert O.rec_probe(p, 3, D0, z) == pytest.approx(D0 ** beta, rel=1e-12)
    D1 = (c0 + s) ** -1.5                     # one width away
    assert O.rec_probe(p, 3, D1, z) == pytest.approx(math.exp(-0.5) * D1 ** beta, rel=1e-12)
    # the integrand probe carries A linearly
    a1 = O.rec_probe(p, 4, math.log(2.0), z, 0.5, 1.0, 1.0)
    a3 = O.rec_probe(p, 4, math.log(2.0), z, 0.5, 1.0, 3.0)
    assert np.allclose(np.array(a3), 3 * np.array(a1), rtol=1e-14, atol=0)


def test_hydrogen_density_today(p):
    rho = 3 * (H100 * p.Hubble_h) ** 2 / (8 * math.pi * G_CGS)
    want = rho * p.BaryonFrac * p.OmegaM * (1 - p.Y_He) / M_P
    assert O.rec_probe(p, 5) == pytest.approx(want, rel=1e-12)
    assert O.rec_probe(p, 5) == pytest.approx(1.91e-7, rel=0.01)  # Genesis cosmology


@pytest.mark.parametrize("z", [6.0, 8.0, 12.5, 20.0])
def test_dtdz_closed_form(p, z):
    H = H100 * p.Hubble_h * math.sqrt(p.OmegaM * (1 + z) ** 3 + p.OmegaLambda)
    assert O.rec_probe(p, 6, z) == pytest.approx(-1 / ((1 + z) * H), rel=1e-10)
    assert O.lib().oracle_dtdz(p.OmegaM, p.OmegaLambda, p.Hubble_h, z) == O.rec_probe(p, 6, z)


@pytest.mark.parametrize("lnD,z,g12,T4", [(math.log(0.05), 7.0, 0.3, 1.0), (math.log(3.0), 9.0, 1.0, 0.8),
                                          (math.log(150.0), 12.0, 0.05, 2.0), (0.0, 6.0, 1e-3, 1.0)])
def test_integrand_factors(p, lnD, z, g12, T4):
    """MHR_rr = n_H P alpha_B x_e^2 Delta^2, cf numerator P x_e^2 Delta^3, denominator
    P x_e Delta^2, MHR_rnh = 1e4 P x_HI Delta (each with one extra Delta for d ln Delta)."""
    rr, num, den, rnh = O.rec_probe(p, 4, lnD, z, g12, T4, 1.0)
    D = math.exp(lnD)
    n_H = O.rec_probe(p, 5) * (1 + z) ** 3 * D
    x_hi = O.rec_probe(p, 0, n_H, T4, O.rec_probe(p, 1, g12, D, T4, z))
    x_e = 1 - x_hi
    alpha = ALPHA_B_1E4 * T4 ** -0.75
    assert rr / num == pytest.approx(O.rec_probe(p, 5) * (1 + z) ** 3 * alpha, rel=1e-12)
    assert num / den == pytest.approx(x_e * D, rel=1e-12)
    assert rnh / den == pytest.approx(1e4 * x_hi / (x_e * D), rel=1e-12)
    c0, beta = O.rec_probe(p, 2, z)
    pdf = O.rec_probe(p, 3, D, z)
    assert den == pytest.approx(pdf * x_e * D * D, rel=1e-12)


# ---- per-cell table lookup (splined_recombination) -----------------------------------------

def _nodes():
    g = np.arange(150, dtype=np.float32)
    return (np.float32(-10.0) + g * np.float32(0.1)).astype(np.float64)


@pytest.fixture(scope="module")
def tables():
    x = _nodes()
    zc = np.arange(200)[:, None, None]
    tc = np.arange(100)[None, :, None]
    xx = x[None, None, :]
    ln_rate = -35.0 + 0.01 * zc - 0.004 * tc + 0.8 * np.tanh(0.4 * xx) + 0.05 * np.sin(xx)
    clump = 1.0 + 0.02 * zc ** 0.5 + 0.3 / (1 + np.exp(-xx)) + 0.001 * tc
    ln_res = 5.0 - 0.7 * xx + 0.01 * np.cos(0.5 * xx) * (1 + 0.01 * tc) - 0.002 * zc
    O.rec_set_tables(x, ln_rate, clump, ln_res)
    return x, ln_rate, clump, ln_res


def _lookup_reference(tables, zc, tc, lnG):
    x, a, b, c = tables
    return (math.exp(CubicSpline(x, a[zc, tc], bc_type="natural")(lnG)),
            math.exp(CubicSpline(x, c[zc, tc], bc_type="natural")(lnG)),
            float(CubicSpline(x, b[zc, tc], bc_type="natural")(lnG)))


@pytest.mark.parametrize("z_eff,temp,g12", [(7.93, 1.2e4, 0.37), (0.04, 150.0, 5.0e-5), (39.9, 9.0e4, 100.0),
                                            (11.1, 1e4, 1.0), (5.0, 3.3e3, 2.2)])
def test_splined_recombination_against_scipy(tables, z_eff, temp, g12):
    # the float grid steps promoted to double, as in the reference's expressions
    zc = min(max(int(z_eff / float(np.float32(0.2)) + 0.5), 0), 199)
    tc = min(int((math.log10(temp) - 2.0) / float(np.float32(0.03)) + 0.5), 99)
    got = O.splined_recombination(z_eff, g12, temp)
    want = _lookup_reference(tables, zc, tc, math.log(g12))
    assert got == pytest.approx(want, rel=1e-12)


def test_splined_recombination_bounds(tables):
    # below the temperature or Gamma grid: no recombinations, residual 1e4, clumping 1
    assert O.splined_recombination(8.0, 1.0, 50.0) == (0.0, 1e4, 1.0)
    assert O.splined_recombination(8.0, math.exp(-10.5), 1e4) == (0.0, 1e4, 1.0)
    # redshift clamped to the first / last row
    assert O.splined_recombination(-0.5, 1.0, 1e4) == O.splined_recombination(0.0, 1.0, 1e4)
    assert O.splined_recombination(80.0, 1.0, 1e4) == O.splined_recombination(39.8, 1.0, 1e4)
    # temperature above the grid: the last row
    assert O.splined_recombination(8.0, 1.0, 1e7) == O.splined_recombination(8.0, 1.0, 10 ** 4.97)
    # Gamma_12 at or above the last node: evaluated just below it
    x = _nodes()
    top = O.splined_recombination(8.0, math.exp(x[-1] + 1.0), 1e4)
    zc, tc = 40, int((4 - 2.0) / float(np.float32(0.03)) + 0.5)
    assert top == pytest.approx(_lookup_reference(tables, zc, tc, x[-1] - 1e-7), rel=1e-12)
