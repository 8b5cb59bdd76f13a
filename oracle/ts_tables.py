"""ts_tables.py -- oracle restatement of the grid-independent part of the spin-temperature step.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's host tables
(mxr_ts_build_tables); never by the product.

Follows:
  src/core/ComputeTs.c:32-83       set_sfr_history (history depth)
  src/core/ComputeTs.c:339-478     shell radii, zpp_edge, history weighting plan
  src/core/ComputeTs.c:761-1011    nu_tau_one limits, sum_lyn and the kink correction
  src/core/ComputeTs.c:835-899     the AGN bands' frequency-integral limits
  src/core/ComputeTs.c:1013-1128   luminosity conversion, const_zp_prefactor_GAL
  src/core/XRayHeatingFunctions.c:262-424 (nu_tau_one, tauX), 426-453 (dtdz, hubble),
    458-516 (cross sections), 545-728 (zmax, nu_n, frecycle, spectral_emissivity),
    742-856 (integrate_over_nu), 861-1280 (x_int tables and their float interpolation),
    1595-1637 (dicke, ddicke_dz), 2010-2042 (interpolate_fcoll)

The reference evaluates tau_X and the frequency integrals with GSL's adaptive QAG
(Gauss-Kronrod 61, relative tolerance 0.005 and 0.01) and finds nu_tau_one with Brent's method
to a relative bracket of 0.02.  Here the same integrals are evaluated to ~1e-10 by
Gauss-Legendre quadrature on the subintervals between the integrand's kinks (table energies,
ionisation thresholds), and the root with scipy's brentq to 1e-12: the library's values must
agree within the reference's own tolerances.  Arithmetic the reference does in float (1 + z
with a float z, the secondary-electron interpolation, the stellar-spectrum normalisation) is
done in numpy float32.
"""
from __future__ import annotations

import math
import os

import numpy as np
from scipy.optimize import brentq

f32 = np.float32

SPEED_OF_LIGHT = 2.9979e10
PLANCK = 6.6262e-27
HUBBLE = 3.2407789e-18
GRAVITY = 6.672e-8
PROTONMASS = 1.6726e-24
MPC = 3.086e24
NU_over_EV = 1.60217646e-12 / PLANCK
NUIONIZATION = 13.60 * NU_over_EV
HeI_NUIONIZATION = 24.59 * NU_over_EV
HeII_NUIONIZATION = NUIONIZATION * 4
NU_LA = 10.2 * NU_over_EV
SIGMA_HI = 6.3e-18
TINY = 1e-30
L_FACTOR = 0.620350491
R_XLy_MAX = 500.0
NSPEC_MAX = 23
REL_TOL = float(f32(1e-5))
X_INT_XHII = np.array([1.0e-4, 2.318e-4, 4.677e-4, 1.0e-3, 2.318e-3, 4.677e-3, 1.0e-2, 2.318e-2, 4.677e-2,
                       1.0e-1, 0.5, 0.9, 0.99, 0.999], dtype=np.float32)
FRECYCLE = [1, 1, 1, 0, 0.2609, 0.3078, 0.3259, 0.3353, 0.3410, 0.3448, 0.3476, 0.3496, 0.3512, 0.3524, 0.3535,
            0.3543, 0.3550, 0.3556, 0.3561, 0.3565, 0.3569, 0.3572, 0.3575, 0.3578, 0.3580, 0.3582, 0.3584, 0.3586,
            0.3587, 0.3589, 0.3590]


class Cosmo:
    def __init__(self, p):
        self.OMm, self.OMl, self.OMr, self.OMk = p["OmegaM"], p["OmegaLambda"], p["OmegaR"], p["OmegaK"]
        self.wl, self.h = p["wLambda"], p["Hubble_h"]
        rho = 3.0 * HUBBLE * HUBBLE * self.h * self.h / (8.0 * math.pi * GRAVITY)
        OMb = p["BaryonFrac"] * p["OmegaM"]
        self.No = rho * OMb * (1 - p["Y_He"]) / PROTONMASS
        He_No = rho * OMb * p["Y_He"] / (4.0 * PROTONMASS)
        self.f_H = self.No / (self.No + He_No)
        self.f_He = He_No / (self.No + He_No)
        self.N_b0 = self.No + He_No

    def dtdz(self, z):
        opz = float(f32(1) + f32(z))
        x = math.sqrt(self.OMl / self.OMm) * opz ** (-1.5)
        dxdz = math.sqrt(self.OMl / self.OMm) * opz ** (-2.5) * (-1.5)
        const1 = 2 * math.sqrt(1 + self.OMm / self.OMl) / (3.0 * HUBBLE * self.h)
        numer = dxdz * (1 + x * (x ** 2 + 1) ** -0.5)
        denom = x + math.sqrt(x ** 2 + 1)
        return const1 * numer / denom

    def hubble(self, z):
        opz = float(f32(1) + f32(z))
        return HUBBLE * self.h * math.sqrt(self.OMm * opz ** 3 + self.OMr * opz ** 4 + self.OMl)

    def drdz(self, z):
        return (1.0 + float(f32(z))) * SPEED_OF_LIGHT * self.dtdz(z)

    def dicke(self, z):
        OMm, OMl, OMr = self.OMm, self.OMl, self.OMr
        if abs(OMm - 1.0) < 1e-4:
            return 1.0 / (1.0 + z)
        om_z = OMm * (1 + z) ** 3 / (OMl + OMm * (1 + z) ** 3 + OMr * (1 + z) ** 4)
        d_z = 2.5 * om_z / (1.0 / 70.0 + om_z * (209 - om_z) / 140.0 + om_z ** (4.0 / 7.0))
        d_0 = 2.5 * OMm / (1.0 / 70.0 + OMm * (209 - OMm) / 140.0 + OMm ** (4.0 / 7.0))
        return d_z / (d_0 * (1.0 + z))


def r_factor(R0, F):
    return (R_XLy_MAX / R0) ** float(f32(1) / f32(F))


def history_depth(p, last_snapshot):
    cz = Cosmo(p)
    ZZ = p["ZZ"]
    box = p["BoxSize"] / p["Hubble_h"]
    F = p["TsNumFilterSteps"]
    best = 0
    for snap in range(1, last_snapshot + 1):
        R = L_FACTOR * box / float(f32(p["ReionGridDim"]))
        Rf = r_factor(R, F)
        scb, prev_R, prev_zpp = 1, 0.0, ZZ[snap]
        for _ in range(F):
            edge = prev_zpp - (R - prev_R) * MPC / cz.drdz(prev_zpp)
            while edge > ZZ[snap - scb]:
                if snap - scb == 0:
                    break
                scb += 1
            prev_R, prev_zpp = R, edge
            R *= Rf
        best = max(best, scb)
    return best


def plan(p):
    """Shell radii, zpp_edge and the per-shell source operations (ComputeTs.c:339-478)."""
    cz = Cosmo(p)
    ZZ, snap = p["ZZ"], p["snapshot"]
    F = p["TsNumFilterSteps"]
    zp = ZZ[snap]
    prev_z = ZZ[snap - 1] if snap > 0 else ZZ[snap]
    dzp = zp - prev_z
    dt_dzp = cz.dtdz(zp)
    box = p["BoxSize"] / p["Hubble_h"]
    R = L_FACTOR * box / float(f32(p["ReionGridDim"]))
    Rf = r_factor(R, F)
    Rs, edges, zpps, ops = [], [], [], []
    scb = [1] * F
    weight = 1.0
    for r in range(F):
        Rs.append(R)
        if r == 0:
            prev_zpp, prev_R = zp, 0.0
        else:
            prev_zpp, prev_R = edges[r - 1], Rs[r - 1]
            scb[r] = scb[r - 1]
        edge = prev_zpp - (R - prev_R) * MPC / cz.drdz(prev_zpp)
        edges.append(edge)
        zpp = (edge + prev_zpp) * 0.5
        zpps.append(zpp)
        total_weight = (prev_zpp - edge) * cz.dtdz(zpp)
        o = []
        if snap - scb[r] >= 0:
            if snap - scb[r] == 0 and weight < 0:
                o.append(("clear", 0, 0.0))
                scb[r] += 1
            elif edge > ZZ[snap - scb[r]]:
                if r == 0:
                    weight = dt_dzp * dzp / total_weight
                    o.append(("scale", 0, weight))
                elif ZZ[snap - scb[r]] > edges[r - 1]:
                    weight = -cz.dtdz(0.5 * (ZZ[snap - scb[r]] + edges[r - 1]))
                    weight *= ZZ[snap - scb[r]] - edges[r - 1]
                    weight /= total_weight
                    o.append(("load", scb[r] - 1, weight))
                else:
                    o.append(("clear", 0, 0.0))
                while edge > ZZ[snap - scb[r]]:
                    if snap - scb[r] == 0:
                        weight = -1.0
                        break
                    if edge < ZZ[snap - scb[r] - 1]:
                        zedge = prev_zpp if prev_zpp > ZZ[snap - scb[r]] else ZZ[snap - scb[r]]
                        weight = edge - zedge
                        weight *= -cz.dtdz(0.5 * (edge + zedge))
                    else:
                        weight = ZZ[snap - scb[r] - 1] - ZZ[snap - scb[r]]
                        weight *= -cz.dtdz(0.5 * (ZZ[snap - scb[r] - 1] + ZZ[snap - scb[r]]))
                    weight /= total_weight
                    o.append(("add", scb[r], weight))
                    scb[r] += 1
            elif r > 0 and weight != 1:
                weight = 1.0
                o.append(("load", scb[r] - 1, weight))
        ops.append(o)
        R *= Rf
    return dict(R=np.array(Rs), zpp_edge=np.array(edges), zpp=np.array(zpps), ops=ops, weight_end=weight,
                zp=zp, dzp=dzp, dt_dzp=dt_dzp, homogeneous=(zp - p["ReionMaxHeatingRedshift"]) >= -0.0001)


# ---- stellar spectra and sum_lyn (XRayHeatingFunctions.c:545-728, ComputeTs.c:918-1011) ----
def stellar_spectra(tables_dir, pop2):
    rows = np.loadtxt(os.path.join(tables_dir, "stellar_spectra.dat"), max_rows=NSPEC_MAX - 1)
    n = np.zeros(NSPEC_MAX, int)
    N0 = np.zeros(NSPEC_MAX, np.float32)
    al = np.zeros(NSPEC_MAX, np.float32)
    n[1:] = rows[:, 0].astype(int)
    N0[1:] = rows[:, 1].astype(np.float32)
    al[1:] = rows[:, 2].astype(np.float32)
    nu = np.zeros(NSPEC_MAX, np.float32)
    for i in range(1, NSPEC_MAX):
        nu[i] = f32(4.0 / 3.0 * (1.0 - 1.0 / float(n[i]) ** 2.0))
    for i in range(1, NSPEC_MAX - 1):
        a1 = float(f32(al[i] + f32(1)))
        fac = float(nu[i + 1]) ** a1 - float(nu[i]) ** a1
        N0[i] = f32(float(N0[i]) * (a1 / fac * pop2))
    return nu, N0, al


def spectral_emissivity(nu_norm, spec):
    nu, N0, al = spec
    for i in range(1, NSPEC_MAX - 1):
        if nu[i] <= nu_norm < nu[i + 1]:
            return float(N0[i]) * nu_norm ** float(al[i]) / NU_LA
    i = NSPEC_MAX - 1
    return float(N0[i]) * nu_norm ** float(al[i]) / NU_LA


def zmax(z, n):
    num = 1 - (n + 1) ** -2.0
    den = 1 - n ** -2.0
    return float(f32(float(f32(1) + f32(z)) * num / den - 1))


def nu_n(n):
    return (1.0 - n ** -2.0) / 0.75


def sum_lyn(p, pl, tables_dir):
    cz = Cosmo(p)
    spec = stellar_spectra(tables_dir, p["ReionNionPhotPerBary"])
    zp, F = pl["zp"], p["TsNumFilterSteps"]
    out = np.zeros(F)
    first_radii, first_zero, weight = True, True, pl["weight_end"]
    for r in range(F):
        prev_zpp = zp if r == 0 else pl["zpp_edge"][r - 1]
        prev_R = 0.0 if r == 0 else pl["R"][r - 1]
        zpp = (pl["zpp_edge"][r] + prev_zpp) * 0.5
        s = 0.0
        for n_ct in range(NSPEC_MAX, 1, -1):
            if zpp > zmax(zp, n_ct):
                continue
            s += FRECYCLE[n_ct] * spectral_emissivity(nu_n(n_ct) * (1 + zpp) / (1.0 + zp), spec)
        out[r] = s
        if r > 2 and out[r] == 0.0 and out[r - 1] > 0.0 and first_radii:
            e = pl["zpp_edge"]
            tmax = (prev_zpp - (pl["R"][r] - prev_R) * MPC / cz.drdz(prev_zpp) + prev_zpp) * 0.5
            tmin = (e[r - 2] - (pl["R"][r - 1] - pl["R"][r - 2]) * MPC / cz.drdz(e[r - 2]) + e[r - 2]) * 0.5
            for ii in range(1000):
                tz = tmin + (tmax - tmin) * float(f32(ii)) / (float(f32(1000)) - 1.0)
                cnt = sum(1 for n_ct in range(NSPEC_MAX, 1, -1) if not tz > zmax(zp, n_ct))
                if cnt == 0 and first_zero:
                    first_zero = False
                    weight = float(f32(ii) / f32(1000))
            out[r] = weight * out[r - 1]
            first_radii = False
    return out


def prefactors(p, zp):
    cz = Cosmo(p)
    a = p["SpecIndexXrayGal"]
    thr, soft = p["NuXrayThreshold"], p["NuXraySoftCut"]
    if abs(a - 1.0) < REL_TOL:
        L = 1.0 / ((thr * NU_over_EV) * math.log(soft / thr))
    else:
        L = 1.0 / ((soft * NU_over_EV) ** (1.0 - a) - (thr * NU_over_EV) ** (1.0 - a))
        L *= (thr * NU_over_EV) ** (-a) * (1 - a)
    L *= 1.0 / PLANCK
    pref = L / (thr * NU_over_EV) * SPEED_OF_LIGHT * (1 + zp) ** (a + 3)
    dz = float(f32(1e-10))
    return dict(const_zp_prefactor_GAL=pref, growth_factor_zp=cz.dicke(zp),
                dgrowth_factor_dzp=(cz.dicke(zp + dz) - cz.dicke(zp)) / dz, dt_dzp=cz.dtdz(zp))


# ---- cross sections and the secondary-electron tables --------------------------------------
def HI_cs(nu):
    nu = np.asarray(nu, float)
    out = np.zeros_like(nu)
    m = nu >= NUIONIZATION
    x = np.where(nu[m] == NUIONIZATION, nu[m] + TINY, nu[m])
    eps = np.sqrt(x / NUIONIZATION - 1)
    out[m] = SIGMA_HI * (NUIONIZATION / x) ** 4 * np.exp(4 - 4 * np.arctan(eps) / eps) / (1 - np.exp(-2 * np.pi / eps))
    return out


def HeII_cs(nu):
    nu = np.asarray(nu, float)
    out = np.zeros_like(nu)
    m = nu >= HeII_NUIONIZATION
    x = np.where(nu[m] == HeII_NUIONIZATION, nu[m] + TINY, nu[m])
    eps = np.sqrt(x / HeII_NUIONIZATION - 1)
    out[m] = 6.3e-18 / 4 * (HeII_NUIONIZATION / x) ** 4 * np.exp(4 - 4 * np.arctan(eps) / eps) / (
        1 - np.exp(-2 * np.pi / eps))
    return out


def HeI_cs(nu):
    nu = np.asarray(nu, float)
    out = np.zeros_like(nu)
    m = nu >= HeI_NUIONIZATION
    x = nu[m] / NU_over_EV / 13.61 - 0.4434
    y = np.sqrt(x * x + 2.136 ** 2)
    out[m] = 9.492e-16 * ((x - 1) ** 2 + 2.039 ** 2) * y ** (0.5 * 3.188 - 5.5) * (1.0 + np.sqrt(y / 1.469)) ** -3.188
    return out


class XInt:
    COLS = dict(fheat=2, nLya=4, nHI=5, nHeI=6, nHeII=7)

    def __init__(self, tables_dir):
        self.tab = {k: np.zeros((14, 258), np.float32) for k in self.COLS}
        for m, x in enumerate(X_INT_XHII):
            if x < 0.3:
                name = "log_xi_%1.1f.dat" % math.log10(float(x))
            else:
                name = "xi_%1.3f.dat" % float(x)
            d = np.loadtxt(os.path.join(tables_dir, "x_int_tables", name), skiprows=3, dtype=np.float64)
            d = d.astype(np.float32)
            self.E = d[:, 0].copy()
            for k, c in self.COLS.items():
                self.tab[k][m] = d[:, c]

    def interp(self, name, En, x, below):
        """Vectorised float interpolation of XRayHeatingFunctions.c:970-1248 (En float32 array)."""
        E, T = self.E, self.tab[name]
        En = np.asarray(En, np.float32).copy()
        top = En.astype(np.float64) > 0.999 * float(E[-1])
        En[top] = f32(float(E[-1]) * 0.999)
        low = En < E[0]
        x = f32(x)
        if float(x) > float(X_INT_XHII[-1]) * 0.999:
            x = f32(float(X_INT_XHII[-1]) * 0.999)
        elif x < X_INT_XHII[0]:
            x = f32(1.001 * float(X_INT_XHII[0]))
        Ed = En.astype(np.float64)
        n_lo = np.zeros(Ed.shape, int)
        small = (Ed < 1008.88) & ~low
        big = ~(Ed < 1008.88) & ~low
        n_lo[small] = np.trunc(np.log(Ed[small] / 10.0) / 1.98026273e-2).astype(int)
        n_lo[big] = 232 + np.trunc(np.log(Ed[big] / 1008.88) / 9.53101798e-2).astype(int)
        n_hi = n_lo + 1
        m = 13
        while x < X_INT_XHII[m]:
            m -= 1
        mh = m + 1
        dE = (E[n_hi] - E[n_lo]).astype(np.float32)
        lo = ((T[m, n_hi] - T[m, n_lo]) / dE).astype(np.float32)
        lo = (lo * (En - E[n_lo])).astype(np.float32)
        lo = (lo + T[m, n_lo]).astype(np.float32)
        hi = ((T[mh, n_hi] - T[mh, n_lo]) / dE).astype(np.float32)
        hi = (hi * (En - E[n_lo])).astype(np.float32)
        hi = (hi + T[mh, n_lo]).astype(np.float32)
        fr = ((hi - lo) / (X_INT_XHII[mh] - X_INT_XHII[m])).astype(np.float32)
        fr = (fr * (x - X_INT_XHII[m])).astype(np.float32)
        fr = (fr + lo).astype(np.float32)
        return np.where(low, f32(below), fr)


def _gl_pieces(bounds, npts):
    """Nodes and weights of npts-point Gauss-Legendre on every [bounds[i], bounds[i+1]]."""
    xg, wg = np.polynomial.legendre.leggauss(npts)
    a = np.asarray(bounds[:-1])[:, None]
    b = np.asarray(bounds[1:])[:, None]
    return ((a + b) / 2 + (b - a) / 2 * xg[None, :]).ravel(), ((b - a) / 2 * wg[None, :]).ravel()


def integrate_over_nu(xi: XInt, cz: Cosmo, p, zp, x_e, lo, hi, flag, npts=16, spec_index=None):
    """The frequency integral of integrate_over_nu to ~1e-10 (kinks at the table energies), for
    the spectral index `spec_index` (default: the galaxies' SpecIndexXrayGal; the AGN bands pass
    theirs, ComputeTs.c:842-896)."""
    alpha = p["SpecIndexXrayGal"] if spec_index is None else spec_index
    if lo >= hi:
        return 0.0
    cuts = [lo, hi]
    for thr in (NUIONIZATION, HeI_NUIONIZATION, HeII_NUIONIZATION):
        e = xi.E.astype(np.float64) * NU_over_EV + thr
        cuts += list(e[(e > lo) & (e < hi)])
        top = 0.999 * float(xi.E[-1]) * NU_over_EV + thr
        if lo < top < hi:
            cuts.append(top)
    cuts = np.unique(np.array(cuts))
    nu, w = _gl_pieces(cuts, npts)
    xf = float(f32(x_e))
    fH, fHe = cz.f_H, cz.f_He
    s = np.zeros_like(nu)
    specs = ((NUIONIZATION, fH * (1 - x_e), HI_cs), (HeI_NUIONIZATION, fHe * (1 - x_e), HeI_cs),
             (HeII_NUIONIZATION, fHe * x_e, HeII_cs))
    for thr, frac, cs in specs:
        En = ((nu - thr) / NU_over_EV).astype(np.float32)
        if flag == 0:
            s += xi.interp("fheat", En, xf, 1.0).astype(float) * PLANCK * (nu - thr) * frac * cs(nu)
        elif flag == 1:
            Fi = (xi.interp("nHI", En, xf, 0.0) + xi.interp("nHeI", En, xf, 0.0)).astype(np.float32)
            Fi = (Fi + xi.interp("nHeII", En, xf, 0.0)).astype(np.float32)
            Fi = (Fi + f32(1)).astype(np.float32)
            s += Fi.astype(float) * frac * cs(nu)
        else:
            s += xi.interp("nLya", En, xf, 0.0).astype(float) * frac * cs(nu)
    s *= (nu / (p["NuXrayThreshold"] * NU_over_EV)) ** (-alpha - 1)
    res = float(np.sum(s * w))
    if flag == 2:
        res *= SPEED_OF_LIGHT / (4 * math.pi) / NU_LA / cz.hubble(zp)
    return res


def tauX(cz: Cosmo, p, nu, x_e, zp, zpp, stored_fcoll, x_e_ave, npieces=64, npts=20):
    ZZ, snap = p["ZZ"], p["snapshot"]
    nu0 = nu / (1 + zp)
    cuts = [zp, zpp]
    for thr in (NUIONIZATION, HeI_NUIONIZATION, HeII_NUIONIZATION):
        zc = thr / nu0 - 1
        if zp < zc < zpp:
            cuts.append(zc)
    cuts = np.unique(np.array(cuts))
    fine = np.unique(np.concatenate([np.linspace(cuts[i], cuts[i + 1], npieces + 1) for i in range(len(cuts) - 1)]))
    z, w = _gl_pieces(fine, npts)
    opz = (np.float32(1) + z.astype(np.float32)).astype(np.float64)
    x = math.sqrt(cz.OMl / cz.OMm) * opz ** -1.5
    dxdz = math.sqrt(cz.OMl / cz.OMm) * opz ** -2.5 * (-1.5)
    const1 = 2 * math.sqrt(1 + cz.OMm / cz.OMl) / (3.0 * HUBBLE * cz.h)
    dtdz = const1 * dxdz * (1 + x * (x ** 2 + 1) ** -0.5) / (x + np.sqrt(x ** 2 + 1))
    n = cz.N_b0 * (1 + z) ** 3
    if snap == 0:
        fc = np.full_like(z, stored_fcoll[0])
    else:
        fc = stored_fcoll[snap - 1] + (z - ZZ[snap - 1]) * (stored_fcoll[snap] - stored_fcoll[snap - 1]) / (
            ZZ[snap] - ZZ[snap - 1])
    fc = np.maximum(fc, 0.0)
    ff = np.where(fc < 1e-20, 1.0, 1 - (p["ReionEfficiency"] * fc) / (1.0 - x_e_ave))
    ff = np.maximum(ff, 1e-4)
    nuh = nu0 * (1 + z)
    sig = cz.f_H * (1 - x_e) * HI_cs(nuh) + cz.f_He * (1 - x_e) * HeI_cs(nuh) + cz.f_He * x_e * HeII_cs(nuh)
    # integral from zpp down to zp (QAG's orientation): minus the integral over [zp, zpp]
    return -float(np.sum(SPEED_OF_LIGHT * dtdz * n * ff * sig * w))


def nu_tau_one(cz, p, zp, zpp, x_e, stored_fcoll):
    if x_e > 0.9999:
        return -1.0
    f = lambda nu: tauX(cz, p, nu, x_e, zp, zpp, stored_fcoll, x_e) - 1.0
    if f(HeI_NUIONIZATION) < 0:
        return HeI_NUIONIZATION
    return brentq(f, HeI_NUIONIZATION, 1e6 * NU_over_EV, xtol=1e-12 * HeI_NUIONIZATION, rtol=1e-12)
