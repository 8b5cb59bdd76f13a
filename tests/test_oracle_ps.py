"""CPU pins of the oracle's 21-cm power spectrum (Compute_PS, ComputePowerSpectrum.c:18-255):
the bin count against log(N)/log(1.35), the bin averages and counts against a numpy brute force
over the half spectrum, single-cosine fields against the closed form, and the A^2 scaling."""
import math

import numpy as np
import pytest

import oracle_lib as O


def _kmag_half(N, box_size):
    """float32 |k| of every half-spectrum mode (n_x, n_y full, n_z <= N/2), as numpy computes it."""
    dk = np.float32(2.0 * math.pi / box_size)
    n = np.arange(N)
    kf = (np.where(n > N // 2, n - N, n).astype(np.float32) * dk).astype(np.float32)
    kz = (np.arange(N // 2 + 1).astype(np.float32) * dk).astype(np.float32)
    kx, ky, kzz = np.meshgrid(kf, kf, kz, indexing="ij")
    k2 = (kx * kx + ky * ky).astype(np.float32) + kzz * kzz
    return np.sqrt(k2.astype(np.float64)).astype(np.float32)


def _bins(N, box_size):
    """Bin index of every half-spectrum mode (-1 beyond the last ceiling), float ceilings."""
    dk = np.float32(2.0 * math.pi / box_size)
    kmax = np.float32(dk * np.float32(N))
    ceil = [np.float32(dk)]
    while True:
        nxt = np.float32(ceil[-1] * np.float32(1.35))
        if not nxt < kmax:
            break
        ceil.append(nxt)
    edges = np.array([0.0] + ceil, np.float32)
    km = _kmag_half(N, box_size)
    b = np.searchsorted(edges, km, side="right") - 1
    b[b >= len(ceil)] = -1
    return km, b, len(ceil)


@pytest.mark.parametrize("N,box,h", [(16, 20.0, 0.678), (32, 40.0, 0.7), (64, 67.8, 0.678), (128, 500.0, 1.0)])
def test_ps_length_closed_form(N, box, h):
    # ceilings delta_k 1.35^i below N delta_k: ceil(log N / log 1.35) bins (no float tie at these N)
    assert O.ps_length(N, box, h) == math.ceil(math.log(N) / math.log(1.35))


@pytest.mark.parametrize("N,box,h", [(16, 20.0, 0.678), (32, 40.0, 0.7)])
def test_bin_averages_and_counts(N, box, h):
    rng = np.random.default_rng(N)
    dT = rng.uniform(5.0, 30.0, N ** 3).astype(np.float32)
    PS_k, PS_data, PS_error = O.power_spectrum(dT, N, box, h)
    km, b, nb = _bins(N, box / h)
    assert nb == len(PS_k)
    counts = np.array([(b == i).sum() for i in range(nb)])
    kave = np.array([km[b == i].astype(np.float64).mean() for i in range(nb)])
    np.testing.assert_allclose(PS_k, kave, rtol=2e-7)
    # PS_error = PS_data / sqrt(count); the first bin holds only the DC mode, zero after the
    # mean is subtracted, so its ratio is 0/0
    assert counts[0] == 1 and PS_data[0] == 0.0
    np.testing.assert_allclose((PS_data[1:] / PS_error[1:]).astype(np.float64) ** 2, counts[1:], rtol=1e-5)


def _cos_field(N, m, axis, A=20.0, eps=0.1):
    x = np.arange(N)
    c = np.cos(2 * math.pi * m * x / N)
    shape = [1, 1, 1]
    shape[axis] = N
    f = A * (1.0 + eps * c.reshape(shape)) * np.ones((N, N, N))
    return f.astype(np.float32).reshape(-1)


@pytest.mark.parametrize("axis,nmodes", [(0, 2), (2, 1)])
def test_single_cosine_closed_form(axis, nmodes):
    # dT = A (1 + eps cos(2 pi m i / N)) along one axis: <dT> = A, and the unnormalised r2c of
    # dd = eps cos(.) V/N^3 A holds eps A V / 2 at k = +-m delta_k.  Along x both n_x = m and
    # N - m lie in the half spectrum; along z only n_z = m does.
    N, box, h, m, A, eps = 32, 40.0, 0.7, 3, 20.0, 0.1
    box_size = box / h
    V = float(np.float32(box_size) ** 3)
    PS_k, PS_data, PS_error = O.power_spectrum(_cos_field(N, m, axis, A, eps), N, box, h)
    km, b, nb = _bins(N, box_size)
    k = float(np.float32(np.float32(m) * np.float32(2 * math.pi / box_size)))
    ib = int(b[m, 0, 0])
    count = int((b == ib).sum())
    want = nmodes * k ** 3 * (eps * A * V / 2) ** 2 / (2 * math.pi ** 2 * V) / count
    assert PS_data[ib] == pytest.approx(want, rel=1e-5)
    others = np.delete(PS_data, ib)
    assert np.all(np.abs(others) < 1e-9 * want)


def test_amplitude_invariance_and_scaling():
    # dd depends on dT only through dT/<dT> times <dT>: PS(s dT) = s^2 PS(dT)
    N, box, h = 16, 20.0, 0.678
    rng = np.random.default_rng(7)
    dT = rng.uniform(1.0, 3.0, N ** 3).astype(np.float32)
    a = O.power_spectrum(dT, N, box, h)
    b = O.power_spectrum((dT * np.float32(4.0)).astype(np.float32), N, box, h)
    np.testing.assert_allclose(b[0], a[0], rtol=0)
    np.testing.assert_allclose(b[1], 16.0 * a[1], rtol=2e-5)
