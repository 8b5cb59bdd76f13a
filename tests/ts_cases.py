"""Seeded inputs and state for the spin-temperature tests (snapshots z = 20 .. 8, dz = -0.3)."""
from __future__ import annotations

import numpy as np

import mxr
import synth
from conftest import CALIB

ZZ = mxr.snapshot_redshifts()
_cache = {}


def params(N, box, snap, **over):
    return mxr.ts_params(N, box, ZZ, snap, **over)


def inputs(N, box, snap, n_hist, scale=1.0):
    """deltax, stars and n_hist SFR / X-ray histories (numpy, padded); `scale` multiplies the
    histories (1: the find_HII_bubbles calibration of this N, strong heating)."""
    seed, s2m = CALIB.get((N, box, 1), (1000 + N, 0.1))
    key = (N, box, snap, n_hist)
    if key not in _cache:
        dev = "cpu"
        if N >= 128:
            import torch

            dev = "cuda" if torch.cuda.is_available() else "cpu"
        d = synth.ts_inputs(N, box, seed, ZZ, snap, n_hist, device=dev, stars_to_matter=s2m)
        _cache[key] = {k: v.cpu().numpy() for k, v in d.items()}
    d = {k: v.copy() for k, v in _cache[key].items()}
    if scale != 1.0:
        for k in ("sfr_histories", "xray_luminosity_histories"):
            d[k] = (d[k].astype(np.float64) * scale).astype(np.float32)
    return d


def state(N, deltax, seed=7, xe0=2.0e-4, T0=40.0):
    """x_e_box_prev (padded) and Tk_box (real) with some spread."""
    rng = np.random.default_rng(seed)
    H = N // 2 + 1
    dx = deltax.reshape(N, N, 2 * H)[:, :, :N]
    xe = np.zeros((N, N, 2 * H), np.float32)
    xe[:, :, :N] = (xe0 * (1.0 + 0.05 * rng.standard_normal((N, N, N)))).astype(np.float32)
    Tk = (T0 * (1.0 + 0.3 * np.clip(dx, -0.9, 3.0))).astype(np.float32)
    return xe.reshape(-1), Tk.reshape(-1)


def rel_err(a, b, floor=0.0):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), floor if floor else 1e-300)


SOLAR_LUM = 3.826e33


def agn_histories(N, d, n_hist, level=1.0, seed=11):
    """AGN soft / hard X-ray emissivity histories for the inputs d (padded float32 grids, 1e10 L_sun
    per cell, [s] = s snapshots back): each snapshot's cells at or above the 99th percentile of its
    positive SFR host an AGN whose hard emissivity follows that SFR with a lognormal scatter, and
    whose soft emissivity is half the hard one with its own scatter.  `level` is the ratio of the
    AGN's hard-band emissivity to the galaxies' X-ray emissivity, volume-averaged per snapshot."""
    rng = np.random.default_rng(seed)
    H = N // 2 + 1
    npad = N * N * 2 * H
    sfr = d["sfr_histories"].reshape(-1, npad)[:n_hist]
    xr = d["xray_luminosity_histories"].reshape(-1, npad)[:n_hist]
    hard = np.zeros((n_hist, npad), np.float32)
    soft = np.zeros((n_hist, npad), np.float32)
    real = np.zeros((N, N, 2 * H), bool)
    real[:, :, :N] = True
    real = real.reshape(-1)
    for s in range(n_hist):
        v = np.where(real, sfr[s].astype(np.float64), 0.0)
        if not np.any(v > 0):
            continue
        host = v >= np.percentile(v[v > 0], 99)
        e = np.where(host, v * rng.lognormal(0.0, 0.5, npad), 0.0)
        want = level * float(np.sum(np.where(real, xr[s], 0.0), dtype=np.float64)) * 1e40 / (1e10 * SOLAR_LUM)
        e *= want / e.sum()
        hard[s] = e.astype(np.float32)
        soft[s] = (0.5 * e * rng.lognormal(0.0, 0.3, npad)).astype(np.float32)
    return soft.reshape(-1), hard.reshape(-1)
