"""Seeded synthetic input grids for the reionization grid path.

Not part of the hot path: it makes the deltax / stars / weighted_sfr grids that the
benchmark and the parity tests feed, identically, to the HIP path and to the CPU oracle.
torch is used for its FFT and RNG on either device; for a given (seed, device) the output is
deterministic.  Grids come back in the reference's padded real layout
(I = k + 2(N/2+1)(j + N i), src/core/misc_tools.c:244-264) with zero padding.

Recipe (one realisation):
  * g: Gaussian random field, white noise coloured in k-space with P(k) ~ k^-2.5, P(0) = 0,
    rescaled to a per-cell standard deviation of 0.5;
  * deltax = exp(g - var/2) - 1 (lognormal, <1 + deltax> ~ 1);
  * halo counts per cell are Poisson with mean lambda = nbar exp(1.5 deltax) / <exp(1.5 deltax)>,
    split over 24 log-spaced mass bins (8 per dex) of the power law dn/dM ~ M^-1.9 on
    [1e8, 1e11] Msun; each bin contributes Poisson(lambda p_b) haloes of the bin's mean mass;
  * stars = summed halo mass in 1e10 Msun/h; weighted_sfr = stars / 100 Myr in internal units;
  * nbar is set by `stars_to_matter`: mean stars per cell = stars_to_matter x the mean matter
    mass of a cell (Omega_m rho_crit V_pix), so the ionised fraction depends on the ratio, not
    on N or the box.
"""
from __future__ import annotations

import math

import numpy as np
import torch

SEC_PER_MEGAYEAR = 3.155e13
GRAVITY = 6.672e-8
HUBBLE = 3.2407789e-18


def _units(p):
    UL, UM, UV = p["UnitLength_in_cm"], p["UnitMass_in_g"], p["UnitVelocity_in_cm_per_s"]
    UT = UL / UV
    G = GRAVITY / UL ** 3 * UM * UT ** 2
    Hub = HUBBLE * UT
    rho_crit = 3 * Hub * Hub / (8 * math.pi * G)
    return UT, rho_crit


def mass_bins(m_lo=1e8, m_hi=1e11, slope=-1.9, per_dex=8):
    ndex = round(math.log10(m_hi / m_lo))
    edges = m_lo * 10.0 ** (np.arange(ndex * per_dex + 1) / per_dex)
    a = slope + 1.0  # integral of M^slope
    b = slope + 2.0  # integral of M^(slope+1)
    n_int = (edges[1:] ** a - edges[:-1] ** a) / a
    m_int = (edges[1:] ** b - edges[:-1] ** b) / b
    p = n_int / n_int.sum()
    mbar = m_int / n_int
    return p, mbar


def gaussian_field(N: int, seed: int, device="cpu", sigma=0.5, slope=-2.5) -> torch.Tensor:
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    white = torch.randn((N, N, N), generator=g, device=device, dtype=torch.float32)
    wk = torch.fft.rfftn(white)
    kx = torch.fft.fftfreq(N, device=device) * N
    kz = torch.fft.rfftfreq(N, device=device) * N
    k2 = kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + kz[None, None, :] ** 2
    amp = torch.where(k2 > 0, k2.clamp_min(1e-30) ** (slope / 4.0), torch.zeros_like(k2))
    field = torch.fft.irfftn(wk * amp, s=(N, N, N))
    field = field - field.mean()
    field = field * (sigma / field.std())
    del white, wk, amp, k2
    return field


def to_padded(x: torch.Tensor) -> torch.Tensor:
    N = x.shape[0]
    H = N // 2 + 1
    out = torch.zeros((N, N, 2 * H), dtype=torch.float32, device=x.device)
    out[:, :, :N] = x
    return out.reshape(-1)


def realisation(N: int, box: float, seed: int, device="cpu", stars_to_matter: float = 1.0, params=None,
                padded=True):
    """Return dict(deltax, stars, weighted_sfr) as float32 tensors on `device`."""
    from mxr import GENESIS  # noqa: E402  (sibling module)

    p = dict(GENESIS)
    if params:
        p.update(params)
    UT, rho_crit = _units(p)
    h = p["Hubble_h"]
    v_pix = (box / N) ** 3
    matter_cell = p["OmegaM"] * rho_crit * v_pix  # 1e10 Msun/h
    prob, mbar = mass_bins()
    mean_mass = float((prob * mbar).sum()) * h / 1e10  # 1e10 Msun/h
    nbar = stars_to_matter * matter_cell / mean_mass

    g = gaussian_field(N, seed, device)
    var = float(g.var())
    deltax = torch.exp(g - var / 2.0) - 1.0
    del g
    w = torch.exp(1.5 * deltax.double())
    lam = (nbar * w / w.mean()).float()
    del w
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed) * 7919 + 17)
    stars = torch.zeros_like(lam, dtype=torch.float64)
    for pb, mb in zip(prob, mbar):
        n = torch.poisson(lam * float(pb), generator=gen)
        stars += n.double() * float(mb * h / 1e10)
        del n
    del lam
    t100 = 100.0 / (UT / SEC_PER_MEGAYEAR)
    sfr = stars / t100
    out = dict(deltax=deltax.float(), stars=stars.float(), weighted_sfr=sfr.float())
    if padded:
        out = {k: to_padded(v) for k, v in out.items()}
    return out


def bh_grids(stars: "torch.Tensor", seed: int, bh_fraction: float = 0.05, params=None):
    """effective_bhm / effective_bhar grids beside a (padded) stars grid: a fraction of the stellar
    mass with cell-to-cell lognormal scatter (sigma 0.5 dex in e-folds), accreting over 50 Myr."""
    from mxr import GENESIS  # noqa: E402

    p = dict(GENESIS)
    if params:
        p.update(params)
    UT, _ = _units(p)
    gen = torch.Generator(device=stars.device)
    gen.manual_seed(int(seed) * 104729 + 3)
    scatter = torch.exp(0.5 * torch.randn(stars.shape, generator=gen, device=stars.device, dtype=torch.float64))
    bhm = (bh_fraction * stars.double() * scatter).float()
    t50 = 50.0 / (UT / SEC_PER_MEGAYEAR)
    return dict(effective_bhm=bhm, effective_bhar=(bhm.double() / t50).float())


def numpy_realisation(N: int, box: float, seed: int, **kw):
    return {k: v.cpu().numpy() for k, v in realisation(N, box, seed, device="cpu", **kw).items()}


def history_sfr(N: int, box: float, seed: int, redshifts, device="cpu", **kw):
    """sfr grids for a list of prior snapshots: seed + i, amplitude ~ 1/(1+z) normalised to
    the last redshift (the spin-temperature history set-up)."""
    z_last = redshifts[-1]
    out = []
    for i, z in enumerate(redshifts):
        r = realisation(N, box, seed + i, device=device, **kw)
        out.append(r["weighted_sfr"] * float((1 + z_last) / (1 + z)))
    return out


SOLAR_MASS = 1.989e33
SEC_PER_YEAR = 3.155e7
LXRAY_GAL = 3.16e40  # erg/s per Msun/yr (input/params/defaults.par:51)


def ts_inputs(N: int, box: float, seed: int, ZZ, snapshot: int, n_hist: int, device="cpu",
              stars_to_matter: float = 1.0, params=None):
    """Inputs of one spin-temperature step at `snapshot`: deltax and stars of that snapshot and
    n_hist padded SFR / X-ray luminosity history grids, [s] = s snapshots back.

    Snapshot i's SFR grid is the realisation with seed + i scaled by (1 + z_last) / (1 + z_i)
    (z_last = ZZ[-1]); histories before snapshot 0 are zero.  X-ray luminosity follows
    reionization.c:1982-2022 with Flag_XrayMetallicity = 0: LXrayGal x SFR[Msun/yr] / 1e40."""
    from mxr import GENESIS  # noqa: E402  (sibling module)

    p = dict(GENESIS)
    if params:
        p.update(params)
    UT = p["UnitLength_in_cm"] / p["UnitVelocity_in_cm_per_s"]
    to_msun_yr = p["UnitMass_in_g"] / SOLAR_MASS * SEC_PER_YEAR / UT
    H = N // 2 + 1
    npad = N * N * 2 * H
    cur = realisation(N, box, seed + snapshot, device=device, stars_to_matter=stars_to_matter, params=params)
    sfr_h = torch.zeros(n_hist * npad, dtype=torch.float32, device=device)
    xr_h = torch.zeros(n_hist * npad, dtype=torch.float32, device=device)
    z_last = float(ZZ[-1])
    for s in range(n_hist):
        i = snapshot - s
        if i < 0:
            break
        if s == 0:
            sfr = cur["weighted_sfr"]
        else:
            sfr = realisation(N, box, seed + i, device=device, stars_to_matter=stars_to_matter,
                              params=params)["weighted_sfr"]
        sfr = (sfr.double() * ((1.0 + z_last) / (1.0 + float(ZZ[i])))).float()
        sfr_h[s * npad:(s + 1) * npad] = sfr
        xr_h[s * npad:(s + 1) * npad] = (sfr.double() * (to_msun_yr * LXRAY_GAL / 1e40)).float()
        del sfr
    return dict(deltax=cur["deltax"], stars=cur["stars"], sfr_histories=sfr_h, xray_luminosity_histories=xr_h)


def galaxies(n: int, box: float, seed: int, n_halos: int | None = None, dead_fraction: float = 0.02,
             negative_fraction: float = 0.0) -> dict:
    """A synthetic galaxy catalogue (galaxy_t fields as structure of arrays, numpy, host).

    Galaxies cluster around n_halos centres (a few tens of kpc/h scatter), so many cells hold
    several galaxies and the order of their float sums matters.  Masses and rates are lognormal
    in internal units (1e10 Msun/h, Myr-scale times); a few galaxies are dead (Type 3), and
    negative_fraction of them carry negative stellar quantities (to exercise the clamps)."""
    rng = np.random.default_rng(seed)
    n_halos = n_halos or max(1, n // 8)
    centres = rng.uniform(0.0, box, (n_halos, 3))
    host = rng.integers(0, n_halos, n)
    pos = centres[host] + rng.normal(0.0, 0.02 * box / 64, (n, 3))
    pos = np.mod(pos, box).astype(np.float32)
    pos[pos >= np.float32(box)] = 0.0
    t = rng.choice(np.array([0, 1, 2, 3], np.int32), n, p=[0.5, 0.5 - dead_fraction, 0.0, dead_fraction])
    lg = lambda mu, s: np.exp(rng.normal(mu, s, n))
    gsm = lg(math.log(1e-3), 1.5)
    sfr = lg(math.log(1e-2), 1.5)
    out = dict(
        Pos=pos, Type=t.astype(np.int32),
        FescWeightedGSM=0.1 * gsm, FescWeightedSfr=0.1 * sfr, Sfr=sfr, GrossStellarMass=gsm,
        ColdGas=lg(math.log(1e-2), 1.0), MetalsColdGas=lg(math.log(1e-4), 1.0),
        BlackHoleMass=lg(math.log(1e-5), 2.0), EffectiveBHM=lg(math.log(1e-6), 2.0),
        EffectiveBHAR=lg(math.log(1e-8), 2.0), BHXrayEmissivity_hard=lg(math.log(1e-4), 2.0),
        BHXrayEmissivity_soft=lg(math.log(1e-4), 2.0))
    if negative_fraction > 0:
        neg = rng.uniform(size=n) < negative_fraction
        for k in ("FescWeightedGSM", "FescWeightedSfr", "Sfr", "GrossStellarMass", "EffectiveBHM"):
            out[k][neg] *= -1.0
    return out


def baryon_grids(N: int, nstore: int, seed: int = 0) -> dict:
    """Output grids for construct_baryon_grids (padded float32, numpy): histories hold nstore
    grids of earlier snapshots (random, so that the shift is visible), the rest random too."""
    rng = np.random.default_rng(seed)
    npad = N * N * 2 * (N // 2 + 1)
    g = {}
    for k in ("stars", "weighted_sfr", "effective_bhm", "effective_bhar", "sfr", "xray_luminosity",
              "BHXrayEmissivity_hard", "BHXrayEmissivity_soft"):
        g[k] = rng.uniform(-1, 1, npad).astype(np.float32)
    for k in ("sfr_histories", "xray_luminosity_histories", "bh_xray_histories_hard", "bh_xray_histories_soft"):
        g[k] = rng.uniform(-1, 1, npad * nstore).astype(np.float32)
    return g
