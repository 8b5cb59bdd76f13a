"""GPU parity of the brightness box with peculiar velocities (mxr_brightness_temperature,
Flag_IncludePecVelsFor21cm 1..3) and of mxr_velocity_gradient against the oracle.

The velocity conversion and velocity_gradient are element-wise float arithmetic: bit for bit.
delta_T differs from the oracle only through the FFT rounding of dv/dz (our transform against
the oracle's): within 1e-4 relative per cell at the 99.9th percentile and 1e-5 of the mean
|delta_T| on the volume mean.  The worst cell may reach 2e-3: the optical-depth form (PecVels 2, 3) divides by
|dv/dz / H + 1|, which amplifies the gradient's rounding where that factor is small."""
import math

import numpy as np
import pytest

import mxr
import oracle_lib as O

pytestmark = pytest.mark.gpu

MPC = 3.086e24
HUBBLE = 3.2407789e-18


def _pad(x):
    N = x.shape[0]
    H = N // 2 + 1
    out = np.zeros((N, N, 2 * H), np.float32)
    out[:, :, :N] = x
    return out.reshape(-1)


def _case(N, box, seed, pecvel, trees=mxr.MXR_TREES_VELOCIRAPTOR):
    p = mxr.tb_params(N, 8.0, True, box=box, pecvel=pecvel, trees_id=trees)
    rng = np.random.default_rng(seed)
    dx = _pad((0.3 * rng.standard_normal((N, N, N))).astype(np.float32))
    xh = rng.uniform(0.0, 1.0, (N, N, N)).astype(np.float32)
    xh[rng.uniform(size=xh.shape) < 0.2] = 0.0  # ionised cells
    TS = rng.uniform(5.0, 60.0, N ** 3).astype(np.float32)
    opz = float(np.float32(1.0) + np.float32(p.redshift))
    Hz = HUBBLE * p.Hubble_h * math.sqrt(p.OmegaM * opz ** 3 + p.OmegaLambda)
    cell = box / p.Hubble_h / N
    i, j, k = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    disp = cell * (2.3 * np.sin(2 * np.pi * k / N + 0.7 * i) + 0.6 * np.cos(2 * np.pi * 3 * k / N + 1.3 * j)
                   + 0.2 * rng.standard_normal((N, N, N)))
    v = (disp * Hz / (opz * opz * 1e5 / MPC)).astype(np.float32)  # VELOCIRAPTOR units
    if trees == mxr.MXR_TREES_GBPTREES:
        v = (disp * Hz / (math.sqrt(1 / opz) / 1000.0 * 1e5 / MPC * opz)).astype(np.float32)
    return p, dx, xh.reshape(-1), TS, _pad(v)


def _check(N, got, ref, ave, ave_r, tag):
    e = np.abs(got.astype(np.float64) - ref) / np.maximum(np.abs(ref), 1e-3 * np.abs(ref).max())
    p999, mx = float(np.quantile(e, 0.999)), float(e.max())
    print(f"{tag}: delta_T rel err p99.9 {p999:.2e} max {mx:.2e}, mean {ave:.6e} vs {ave_r:.6e}", flush=True)
    assert p999 <= 1e-4 and mx <= 2e-3
    # delta_T takes both signs (T_S above or below T_CMB): the mean is bounded against mean |delta_T|
    assert abs(ave - ave_r) <= 1e-5 * max(abs(ave_r), float(np.abs(ref).mean()))


@pytest.mark.parametrize("pecvel", [1, 2, 3])
@pytest.mark.parametrize("N,box", [(32, 50.0), (64, 100.0)])
def test_pecvel_host_vs_oracle(N, box, pecvel):
    print(f"tb N={N} pecvel={pecvel}", flush=True)
    p, dx, xh, TS, v = _case(N, box, 11 + N, pecvel)
    ref, ave_r, vref = O.brightness_temperature(p, xh, dx, TS, v, return_vel=True)
    ctx = mxr.Context(N)
    dT = np.zeros(N ** 3, np.float32)
    vg = v.copy()
    ave = ctx.brightness_temperature(p, xh, dx, dT, TS, vel=vg)
    np.testing.assert_array_equal(vg, vref)  # converted in place, bit for bit
    _check(N, dT, ref, ave, ave_r, f"N={N} pecvel={pecvel}")
    ctx.close()


def test_pecvel_gbptrees_units_and_no_spin():
    N, box = 32, 50.0
    p, dx, xh, TS, v = _case(N, box, 3, 3, trees=mxr.MXR_TREES_GBPTREES)
    ref, ave_r, vref = O.brightness_temperature(p, xh, dx, TS, v, return_vel=True)
    ctx = mxr.Context(N)
    dT = np.zeros(N ** 3, np.float32)
    vg = v.copy()
    ave = ctx.brightness_temperature(p, xh, dx, dT, TS, vel=vg)
    np.testing.assert_array_equal(vg, vref)
    _check(N, dT, ref, ave, ave_r, "GBPtrees pecvel=3")
    # without spin temperature the velocity terms do not apply (the reference skips them)
    p.Flag_IncludeSpinTemp = 0
    ref0, ave0 = O.brightness_temperature(p, xh, dx)
    vg = v.copy()
    ave = ctx.brightness_temperature(p, xh, dx, dT, None, vel=vg)
    np.testing.assert_array_equal(vg, v)
    np.testing.assert_allclose(dT, ref0, rtol=1e-6, atol=1e-6)
    ctx.close()


def test_pecvel3_device_128():
    import torch

    N, box = 128, 150.0
    print("tb N=128 pecvel=3 device", flush=True)
    p, dx, xh, TS, v = _case(N, box, 5, 3)
    ref, ave_r, vref = O.brightness_temperature(p, xh, dx, TS, v, return_vel=True)
    ctx = mxr.Context(N)
    t = {k: torch.from_numpy(a.copy()).cuda() for k, a in dict(xh=xh, dx=dx, TS=TS, v=v).items()}
    dT = torch.zeros(N ** 3, device="cuda")
    ave = ctx.brightness_temperature(p, t["xh"], t["dx"], dT, t["TS"], vel=t["v"])
    torch.cuda.synchronize()
    np.testing.assert_array_equal(t["v"].cpu().numpy(), vref)
    _check(N, dT.cpu().numpy(), ref, ave, ave_r, "N=128 pecvel=3 device")
    ctx.close()


@pytest.mark.parametrize("device", [False, True])
def test_velocity_gradient_entry_point(device):
    N, box = 32, 40.0
    rng = np.random.default_rng(9)
    H = N // 2 + 1
    spec = (rng.standard_normal((N, N, H)) + 1j * rng.standard_normal((N, N, H))).astype(np.complex64)
    ref = O.velocity_gradient(spec, N, box)
    ctx = mxr.Context(N)
    if device:
        import torch

        t = torch.from_numpy(spec.view(np.float32).reshape(-1).copy()).cuda()
        ctx.velocity_gradient(t, N, N, box)
        torch.cuda.synchronize()
        got = t.cpu().numpy().view(np.complex64).reshape(N, N, H)
    else:
        got = spec.copy()
        ctx.velocity_gradient(got, N, N, box)
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))
    with pytest.raises(mxr.MxrError):
        ctx.velocity_gradient(spec.copy(), N, N, -1.0)
    ctx.close()
