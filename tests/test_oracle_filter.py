"""Pins the oracle's filter() (src/core/reionization.c:3373-3433) with closed forms of the
three windows, evaluated independently in float64 here."""
import numpy as np
import pytest

import mxr
import oracle_lib as O

N, BOX = 32, 50.0


def kR_grid(R):
    H = N // 2 + 1
    n = np.arange(N)
    k1 = np.where(n > N // 2, n - N, n) * 2 * np.pi / BOX
    kz = np.arange(H) * 2 * np.pi / BOX
    kmag = np.sqrt(k1[:, None, None] ** 2 + k1[None, :, None] ** 2 + kz[None, None, :] ** 2)
    return (kmag * R).reshape(-1)


def window(ftype, R, dtype=np.complex64):
    ones = np.ones(N * N * (N // 2 + 1), dtype)
    return O.filter_box(ones, N, BOX, R, ftype).real


@pytest.mark.parametrize("ftype", [0, 1, 2])
def test_dc_mode_untouched(ftype):
    assert window(ftype, 5.0)[0] == 1.0


def test_sharp_k_cutoff():
    R = 4.7
    x = 0.413566994 * kR_grid(R)
    w = window(1, R)
    assert set(np.unique(w)) <= {0.0, 1.0}
    far = np.abs(x - 1) > 1e-5
    np.testing.assert_array_equal(w[far], (x[far] <= 1).astype(np.float32))
    assert 0 < w.sum() < w.size


def test_real_space_top_hat():
    R = 3.1
    x = kR_grid(R)
    ref = np.where(x > 1e-4, 3 * (np.sin(x) - x * np.cos(x)) / np.maximum(x, 1e-30) ** 3, 1.0)
    w = window(0, R)
    big = np.abs(ref) > 1e-3
    np.testing.assert_allclose(w[big], ref[big], rtol=2e-4)
    np.testing.assert_allclose(w[~big], ref[~big], atol=2e-6)
    small = (x > 1e-4) & (x < 0.05)
    if small.any():
        np.testing.assert_allclose(w[small], 1 - x[small] ** 2 / 10, atol=1e-6)


def test_gaussian():
    R = 2.2
    x = 0.643 * kR_grid(R)
    np.testing.assert_allclose(window(2, R), np.exp(-x ** 2 / 2), rtol=1e-5, atol=1e-30)


@pytest.mark.parametrize("ftype", [0, 1, 2])
def test_fp64_window_matches_fp32(ftype):
    R = 3.3
    a = window(ftype, R)
    b = window(ftype, R, np.complex128)
    x = 0.413566994 * kR_grid(R)
    keep = (np.abs(x - 1) > 1e-5) if ftype == 1 else np.ones_like(x, bool)
    np.testing.assert_allclose(a[keep], b[keep], rtol=2e-4, atol=2e-6)


def test_undefined_filter_type():
    with pytest.raises(mxr.MxrError) as e:
        window(3, 1.0)
    assert e.value.code == mxr.MXR_ERR_FILTER


# a radius on the sharp-k edge of one mode at N = 64, 67.8 cMpc/h: found by a search over the
# modes whose float |k| differs between the reference's separately rounded sum
# (kx^2 + ky^2) + kz^2 and the same sum contracted into FMAs, fma(kz, kz, fma(kx, kx, ky^2))
EDGE_N, EDGE_BOX, EDGE_R, EDGE_MODE = 64, 67.8, 2.596230983734131, (0, 1, 10)


def _f32_kmag(mode, N, box, fused):
    f32 = np.float32
    dk = f32(2 * np.pi / f32(box))
    kx, ky, kz = (f32(f32(n - N if n > N // 2 else n) * dk) for n in mode)
    if fused:
        s = f32(np.float64(kz) * kz + np.float64(f32(np.float64(kx) * kx + np.float64(f32(ky * ky)))))
    else:
        s = f32(f32(f32(kx * kx) + f32(ky * ky)) + f32(kz * kz))
    return f32(np.sqrt(np.float64(s)))


def _keep(kmag, R):
    kR = np.float32(kmag * np.float32(R))
    return not (np.float32(np.float64(kR) * 0.413566994) > np.float32(1.0))


def test_sharp_k_edge_rounds_unfused():
    """The oracle keeps the edge mode, as the separately rounded |k| does; the FMA-contracted
    |k| would drop it, so this radius tells the two roundings apart."""
    assert _keep(_f32_kmag(EDGE_MODE, EDGE_N, EDGE_BOX, False), EDGE_R)
    assert not _keep(_f32_kmag(EDGE_MODE, EDGE_N, EDGE_BOX, True), EDGE_R)
    H = EDGE_N // 2 + 1
    ones = np.ones(EDGE_N * EDGE_N * H, np.complex64)
    w = O.filter_box(ones, EDGE_N, EDGE_BOX, EDGE_R, 1).real.reshape(EDGE_N, EDGE_N, H)
    i, j, k = EDGE_MODE
    assert w[i, j, k] == 1.0 and w[i, EDGE_N - j, k] == 1.0
