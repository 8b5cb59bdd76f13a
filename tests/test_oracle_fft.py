"""Pins the oracle's FFT (the stand-in for FFTW's fftwf r2c/c2r) without trusting it:
brute-force DFT sums, closed forms, Parseval and round trips."""
import numpy as np
import pytest

import oracle_lib as O


def dft_matrix(n, sign):
    k = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / n)


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 8, 10, 12, 15, 16, 20, 24])
@pytest.mark.parametrize("sign", [-1, 1])
def test_c2c_brute_force(n, sign, rng):
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = dft_matrix(n, sign) @ x
    y64 = O.c2c1d(x.astype(np.complex128), sign)
    np.testing.assert_allclose(y64, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    y32 = O.c2c1d(x.astype(np.complex64), sign)
    np.testing.assert_allclose(y32, ref, rtol=0, atol=2e-6 * np.abs(ref).max())


@pytest.mark.parametrize("n", [48, 80, 96, 160, 256, 512])
def test_c2c_mixed_radix_against_numpy(n, rng):
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    np.testing.assert_allclose(O.c2c1d(x, -1), np.fft.fft(x), atol=1e-11 * np.sqrt(n))
    np.testing.assert_allclose(O.c2c1d(x, +1), np.fft.ifft(x) * n, atol=1e-11 * np.sqrt(n))


def padded(x):
    N = x.shape[0]
    H = N // 2 + 1
    out = np.zeros((N, N, 2 * H), x.dtype)
    out[:, :, :N] = x
    return out.reshape(-1)


def unpad(y, N):
    H = N // 2 + 1
    return y.reshape(N, N, 2 * H)[:, :, :N]


@pytest.mark.parametrize("N", [4, 6, 8])
def test_r2c3d_brute_force(N, rng):
    x = rng.standard_normal((N, N, N))
    Fm = dft_matrix(N, -1)
    full = np.einsum("ai,bj,ck,ijk->abc", Fm, Fm, Fm, x)
    got = O.r2c3d(padded(x), N).reshape(N, N, N // 2 + 1)
    np.testing.assert_allclose(got, full[:, :, : N // 2 + 1], atol=1e-11)


@pytest.mark.parametrize("N", [8, 16, 32])
def test_round_trip_is_n_cubed(N, rng):
    x = rng.standard_normal((N, N, N))
    spec = O.r2c3d(padded(x), N)
    back = unpad(O.c2r3d(spec, N), N)
    np.testing.assert_allclose(back, x * N ** 3, atol=1e-9 * N ** 3)
    spec32 = O.r2c3d(padded(x.astype(np.float32)), N)
    back32 = unpad(O.c2r3d(spec32, N), N)
    np.testing.assert_allclose(back32 / N ** 3, x, atol=3e-6)


def test_delta_and_plane_wave():
    N = 16
    H = N // 2 + 1
    d = np.zeros((N, N, N))
    d[0, 0, 0] = 1.0
    np.testing.assert_allclose(O.r2c3d(padded(d), N), np.ones(N * N * H), atol=1e-14)
    # one half-spectrum mode (0 < kz < N/2) -> 2 cos(2 pi k.n / N)
    kx, ky, kz = 3, 5, 2
    spec = np.zeros((N, N, H), np.complex128)
    spec[kx, ky, kz] = 1.0
    x = unpad(O.c2r3d(spec.reshape(-1), N), N)
    n = np.arange(N)
    phase = 2 * np.pi * (kx * n[:, None, None] + ky * n[None, :, None] + kz * n[None, None, :]) / N
    np.testing.assert_allclose(x, 2 * np.cos(phase), atol=1e-12)


def test_parseval(rng):
    N = 16
    H = N // 2 + 1
    x = rng.standard_normal((N, N, N))
    X = O.r2c3d(padded(x), N).reshape(N, N, H)
    w = np.full(H, 2.0)
    w[0] = 1.0
    w[-1] = 1.0
    assert np.isclose((np.abs(X) ** 2 * w).sum() / N ** 3, (x ** 2).sum(), rtol=1e-12)


def test_fp32_close_to_fp64(rng):
    N = 32
    x = rng.standard_normal((N, N, N))
    a = O.r2c3d(padded(x), N)
    b = O.r2c3d(padded(x.astype(np.float32)), N)
    assert np.abs(a - b).max() / np.abs(a).max() < 1e-5
