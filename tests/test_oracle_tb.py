"""CPU pins of the oracle's brightness box with peculiar velocities (BrightnessTemperature.c)
and of velocity_gradient (reionization.c:3435-3450): closed forms and invariants only."""
import math

import numpy as np
import pytest

import mxr
import oracle_lib as O

MPC = 3.086e24
HUBBLE = 3.2407789e-18


def _hubble(p):
    opz = np.float32(1.0) + np.float32(p.redshift)
    return HUBBLE * p.Hubble_h * math.sqrt(p.OmegaM * float(opz) ** 3 + p.OmegaR * float(opz) ** 4 + p.OmegaLambda)


def _pad(x):
    N = x.shape[0]
    H = N // 2 + 1
    out = np.zeros((N, N, 2 * H), np.float32)
    out[:, :, :N] = x
    return out.reshape(-1)


def _unpad(x, N):
    H = N // 2 + 1
    return x.reshape(N, N, 2 * H)[:, :, :N]


def _fields(N, seed=1, xH=None):
    rng = np.random.default_rng(seed)
    dx = (0.3 * rng.standard_normal((N, N, N))).astype(np.float32)
    xh = rng.uniform(0.2, 1.0, (N, N, N)).astype(np.float32) if xH is None else np.full((N, N, N), xH, np.float32)
    TS = rng.uniform(5.0, 60.0, (N, N, N)).astype(np.float32)
    return _pad(dx), xh.reshape(-1), TS.reshape(-1)


def _vel_in(v_conv, p):
    """Input velocities (VELOCIRAPTOR units) whose comoving conversion is v_conv."""
    opz = 1.0 + float(np.float32(p.redshift))
    return (v_conv / (opz * opz * 1e5 / MPC)).astype(np.float32)


def test_velocity_gradient_differentiates_plane_waves():
    N, box = 32, 60.0
    z = np.arange(N)
    for m in (1, 3, 7):
        v = np.broadcast_to(np.sin(2 * np.pi * m * z / N), (N, N, N)).astype(np.float32)
        spec = np.zeros(N * N * (N // 2 + 1) * 2, np.float32)
        vp = _pad(v)
        O.lib().oracle_r2c3d_f32(N, vp.ctypes.data, spec.ctypes.data)
        g = O.velocity_gradient(spec.view(np.complex64), N, box)
        g = g.view(np.float32).copy()
        O.lib().oracle_c2r3d_f32(N, g.ctypes.data)
        d = _unpad(g, N) / N ** 3
        want = (2 * np.pi * m / box) * np.cos(2 * np.pi * m * z / N)
        np.testing.assert_allclose(d, np.broadcast_to(want, (N, N, N)), atol=2e-5 * 2 * np.pi * m / box)
        # a k_z factor only: x and y waves have no z-derivative
        vx = _pad(np.broadcast_to(np.sin(2 * np.pi * m * z / N)[:, None, None], (N, N, N)).astype(np.float32))
        O.lib().oracle_r2c3d_f32(N, vx.ctypes.data, spec.ctypes.data)
        gx = O.velocity_gradient(spec.view(np.complex64), N, box).view(np.float32).copy()
        O.lib().oracle_c2r3d_f32(N, gx.ctypes.data)
        assert np.abs(_unpad(gx, N)).max() / N ** 3 < 1e-5


@pytest.mark.parametrize("trees", [mxr.MXR_TREES_VELOCIRAPTOR, mxr.MXR_TREES_GBPTREES, mxr.MXR_TREES_VELOCIRAPTOR_AUG, 7])
def test_velocity_units(trees):
    N = 16
    p = mxr.tb_params(N, 8.0, True, box=40.0, pecvel=1, trees_id=trees)
    dx, xh, TS = _fields(N)
    rng = np.random.default_rng(2)
    v = _pad(rng.uniform(-300, 300, (N, N, N)).astype(np.float32))
    _, _, vc = O.brightness_temperature(p, xh, dx, TS, v, return_vel=True)
    opz = 1.0 + float(np.float32(p.redshift))
    vin = _unpad(v, N).astype(np.float64)
    f = {mxr.MXR_TREES_GBPTREES: math.sqrt(1 / opz) / 1000.0, mxr.MXR_TREES_VELOCIRAPTOR: opz,
         mxr.MXR_TREES_VELOCIRAPTOR_AUG: opz}.get(trees, 1.0)
    want = vin * f * 1e5 / MPC * opz
    np.testing.assert_allclose(_unpad(vc, N), want, rtol=3e-7)


def _plane_vel(N, p, amp_frac):
    """Comoving velocity sin wave along z whose gradient peaks at amp_frac H(z) (per Mpc/h,
    times Hubble_h as the reference applies)."""
    Hz = float(np.float32(_hubble(p)))
    k = 2 * np.pi / p.BoxSize
    A = amp_frac * Hz / (k * p.Hubble_h)
    z = np.arange(N)
    v = np.broadcast_to(A * np.sin(2 * np.pi * z / N), (N, N, N))
    grad = np.broadcast_to(p.Hubble_h * A * k * np.cos(2 * np.pi * z / N), (N, N, N))
    return _pad(_vel_in(v, p)), grad, Hz


@pytest.mark.parametrize("amp", [0.1, 0.5])
def test_mode1_closed_form_with_cap(amp):
    N = 32
    p = mxr.tb_params(N, 8.0, True, box=50.0, pecvel=1)
    dx, xh, TS = _fields(N)
    v, grad, Hz = _plane_vel(N, p, amp)
    base, _ = O.brightness_temperature(mxr.tb_params(N, 8.0, True, box=50.0, pecvel=0), xh, dx, TS)
    got, ave = O.brightness_temperature(p, xh, dx, TS, v)
    cap = abs(np.float32(0.2) * np.float32(Hz))
    g = np.clip(grad.reshape(-1), -cap, cap)
    want = base.astype(np.float64) / (g / Hz + 1.0)
    np.testing.assert_allclose(got, want, rtol=2e-4)
    assert (np.abs(grad) > cap).any() == (amp > 0.2)
    assert ave == pytest.approx(np.mean(got, dtype=np.float64), rel=1e-12)


def test_mode2_optical_depth_closed_form():
    N = 32
    p = mxr.tb_params(N, 8.0, True, box=50.0, pecvel=2)
    dx, xh, TS = _fields(N)
    v, grad, Hz = _plane_vel(N, p, 0.4)
    z = float(np.float32(p.redshift))
    T_rad = 2.728 * (1 + z)
    # tau itself: T0 xH (1+delta) (1+z) / (1000 T_S)
    mx = mxr.tb_params(N, 8.0, False, box=50.0)
    t0, _ = O.brightness_temperature(mx, xh, dx)
    tau = t0.astype(np.float64) * (1 + z) / (1000.0 * TS)
    got, _ = O.brightness_temperature(p, xh, dx, TS, v)
    gc = np.abs(grad.reshape(-1) / Hz + 1.0)
    want = (1 - np.exp(-tau / gc)) * 1000.0 * (TS - T_rad) / (1 + z)
    np.testing.assert_allclose(got, want, rtol=3e-4, atol=1e-6)


def test_mode3_zero_velocity_equals_mode2():
    N = 16
    dx, xh, TS = _fields(N)
    v = np.zeros(N * N * 2 * (N // 2 + 1), np.float32)
    a, _ = O.brightness_temperature(mxr.tb_params(N, 8.0, True, box=30.0, pecvel=2), xh, dx, TS, v)
    b, _ = O.brightness_temperature(mxr.tb_params(N, 8.0, True, box=30.0, pecvel=3), xh, dx, TS, v)
    np.testing.assert_allclose(b, a, rtol=2e-6, atol=1e-9)


def test_mode3_half_cell_shift_averages_neighbours():
    """A uniform velocity displacing every sub-cell by half a cell away from the observer: each
    cell keeps half of its signal and receives half of the previous cell's."""
    N = 16
    p = mxr.tb_params(N, 8.0, True, box=30.0, pecvel=3)
    dx, xh, TS = _fields(N, xH=1.0)
    Hz = float(np.float32(_hubble(p)))
    cell = p.BoxSize / p.Hubble_h / N
    v = _pad(_vel_in(np.full((N, N, N), 0.5 * cell * Hz), p))
    a, _ = O.brightness_temperature(mxr.tb_params(N, 8.0, True, box=30.0, pecvel=2), xh, dx, TS, v)
    b, _ = O.brightness_temperature(p, xh, dx, TS, v)
    a = a.reshape(N, N, N).astype(np.float64)
    want = 0.5 * (a + np.roll(a, 1, axis=2))
    # 50 float additions per cell: absolute rounding ~1e-6 of the largest |signal|
    np.testing.assert_allclose(b.reshape(N, N, N), want, rtol=1e-4, atol=1e-6 * np.abs(a).max())


def test_mode3_conserves_each_line_of_sight():
    N = 32
    p = mxr.tb_params(N, 8.0, True, box=50.0, pecvel=3)
    dx, xh, TS = _fields(N, xH=1.0)
    rng = np.random.default_rng(5)
    Hz = float(np.float32(_hubble(p)))
    cell = p.BoxSize / p.Hubble_h / N
    z = np.arange(N)
    phase = rng.uniform(0, 2 * np.pi, (N, N, 1))
    disp = 2.3 * cell * np.sin(2 * np.pi * z / N + phase)  # up to 2.3 cells either way
    v = _pad(_vel_in(disp * Hz, p))
    a, _ = O.brightness_temperature(mxr.tb_params(N, 8.0, True, box=50.0, pecvel=2), xh, dx, TS, v)
    b, _ = O.brightness_temperature(p, xh, dx, TS, v)
    sa = a.reshape(N, N, N).astype(np.float64).sum(axis=2)
    sb = b.reshape(N, N, N).astype(np.float64).sum(axis=2)
    mass = np.abs(a.reshape(N, N, N)).astype(np.float64).sum(axis=2)
    assert np.all(np.abs(sb - sa) <= 2e-6 * mass)  # float rounding of the redistributed sums
    assert np.abs(b - a).max() > 1e-3 * np.abs(a).max()  # and the signal did move
