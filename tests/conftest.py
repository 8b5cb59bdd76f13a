"""Shared test set-up: the `gpu` marker, import paths, and seeded input fixtures."""
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "meraxes-devs_qyxnew_amd")
for p in (PKG, os.path.join(REPO, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libmxr.so")


# Seeded test configurations: (N, box, ReionFilterType) -> (seed, stars_to_matter), the ratio
# bisected on the oracle so that the seed's realisation ends with a global xH near one half
# (the ionised fraction depends strongly on the rare massive haloes of each realisation).
CALIB = {
    (16, 20.0, 1): (1016, 0.27824),
    (32, 40.0, 1): (1032, 0.23336),
    (64, 67.8, 1): (1064, 0.04474),
    (64, 67.8, 0): (1064, 0.61529),
    (64, 67.8, 2): (1064, 0.89551),
    (96, 100.0, 1): (1096, 0.01253),
    (128, 125.0, 1): (1128, 0.19495),
    (160, 150.0, 1): (1160, 0.02558),
    (256, 200.0, 1): (1256, 0.020055),
}

_cache = {}


def realisation(N, box, ftype=1, seed=None, s2m=None):
    """Synthetic inputs of a calibrated configuration (or an explicit seed / ratio)."""
    cs, cr = CALIB.get((N, box, ftype), (1000 + N, 0.1))
    seed = cs if seed is None else seed
    s2m = cr if s2m is None else s2m
    key = (N, box, seed, s2m)
    if key not in _cache:
        import synth

        _cache[key] = synth.numpy_realisation(N, box, seed, stars_to_matter=s2m)
    return {k: v.copy() for k, v in _cache[key].items()}


@pytest.fixture
def rng():
    return np.random.default_rng(12345)
