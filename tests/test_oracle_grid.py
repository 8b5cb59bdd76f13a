"""CPU pins of the oracle's input-grid preparation (read_grid__gbptrees, smooth_grid,
subsample_grid; read_grids*.c): the normalisation without resampling, a constant field through
the smoothing, a plane wave damped by exactly the real-space top-hat window, grid selection in
the file, and the header checks."""
import math

import numpy as np
import pytest

import mxr
import oracle_lib as O
from grid_files import density_grids, write_gbptrees


def _real(x, N):
    return x.reshape(N, N, 2 * (N // 2 + 1))[:, :, :N]


def test_density_without_resampling(tmp_path):
    N, box = 16, 20.0
    g = density_grids(N, 1)
    f = tmp_path / "g.dat"
    write_gbptrees(f, g, box)
    p = mxr.grid_params(box, npart=1000, part_mass=0.5, hubble_h=0.7)
    out = _real(O.read_grid_gbptrees(str(f), mxr.MXR_GRID_DENSITY, N, p), N)
    mean_inv = np.float32(box ** 3 * 0.7 / 1000 / 0.5)
    want = np.maximum((g[0] * mean_inv).astype(np.float64) - 1.0, -1.0).astype(np.float32)
    assert np.array_equal(out, want)
    # the velocity grids come through untouched; grid 2 is the y component
    v = _real(O.read_grid_gbptrees(str(f), mxr.MXR_GRID_Y_VELOCITY, N, p), N)
    assert np.array_equal(v, g[2])


def test_floors_of_the_two_readers():
    N = 8
    hi = np.zeros((N, N, N), np.float32)
    hi[0, 0, 0] = 1.0
    for trees, floor in [(mxr.MXR_TREES_GBPTREES, np.float32(-1.0)),
                         (mxr.MXR_TREES_VELOCIRAPTOR, np.float32(-1.0 + float(np.float32(1e-5))))]:
        p = mxr.grid_params(10.0, npart=8, part_mass=1.0, hubble_h=1.0, trees_id=trees)
        out = _real(O.prepare_grid(N, hi, N, mxr.MXR_GRID_DENSITY, 10.0, p), N)
        assert out[1, 1, 1] == floor
        # VELOCIraptor divides by NPart * PartMass at once, gbpTrees by each in turn
        assert out[0, 0, 0] == pytest.approx(1000 / 8 - 1, rel=1e-6)


def test_constant_field_survives_smoothing():
    N, n_hi, box = 16, 32, 30.0
    hi = np.full((n_hi, n_hi, n_hi), 3.0, np.float32)
    p = mxr.grid_params(box)
    out = _real(O.prepare_grid(N, hi, n_hi, mxr.MXR_GRID_X_VELOCITY, box, p), N)
    np.testing.assert_allclose(out, 3.0, rtol=1e-6)


def test_plane_wave_takes_the_top_hat_window():
    # f = 1 + A cos(2 pi m x / n_hi): the top-hat of radius R = L / N / 2 multiplies the mode by
    # W(kR) = 3 (sin kR - kR cos kR) / kR^3, and subsampling keeps every (n_hi/N)-th x
    N, n_hi, box, m, A = 16, 32, 30.0, 3, 0.4
    x = np.arange(n_hi)
    hi = (1 + A * np.cos(2 * math.pi * m * x / n_hi))[:, None, None] * np.ones((n_hi, n_hi, n_hi))
    p = mxr.grid_params(box)
    out = _real(O.prepare_grid(N, hi.astype(np.float32), n_hi, mxr.MXR_GRID_Z_VELOCITY, box, p), N)
    kR = m * 2 * math.pi / box * (box / N / 2)
    W = 3 * (math.sin(kR) - kR * math.cos(kR)) / kR ** 3
    want = 1 + A * W * np.cos(2 * math.pi * m * (2 * np.arange(N)) / n_hi)
    np.testing.assert_allclose(out[:, 3, 5], want, atol=2e-6)
    assert np.allclose(out, out[:, :1, :1], atol=2e-6)


def test_header_checks(tmp_path):
    N = 8
    g = density_grids(N, 2)
    p = mxr.grid_params(10.0)
    bad = tmp_path / "three.dat"
    write_gbptrees(bad, g[:3], 10.0, n_grids=3)
    with pytest.raises(mxr.MxrError):
        O.read_grid_gbptrees(str(bad), 0, N, p)
    odd = tmp_path / "odd.dat"
    write_gbptrees(odd, density_grids(12, 3), 10.0)
    with pytest.raises(mxr.MxrError):
        O.read_grid_gbptrees(str(odd), 0, N, p)
