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
