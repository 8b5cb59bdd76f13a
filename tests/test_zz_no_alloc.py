"""mxr_brightness_temperature_noalloc (the entry point that allocates nothing, with its scratch in
a fixed device array) against mxr_brightness_temperature, bit for bit, on two calls in a row,
the second with a smaller grid, with and without the spin-temperature factor."""
import numpy as np
import pytest

import mxr

pytestmark = pytest.mark.gpu


def _inputs(N, seed):
    r = np.random.default_rng(seed)
    H = N // 2 + 1
    xH = r.uniform(0, 1, N ** 3).astype(np.float32)
    dx = np.zeros((N, N, 2 * H), np.float32)
    dx[:, :, :N] = r.lognormal(0, 0.5, (N, N, N)) - 1
    TS = r.uniform(3, 300, N ** 3).astype(np.float32)
    return xH, dx.reshape(-1), TS


@pytest.mark.parametrize("spin", [False, True])
def test_noalloc_matches_context_entry_point(spin):
    import torch

    for N in (128, 64):  # the second call smaller: the scratch keeps the first call's partials
        print(f"noalloc brightness N={N} spin={spin}", flush=True)
        xH, dx, TS = _inputs(N, 11 + N)
        p = mxr.tb_params(N, 8.0, spin)
        ctx = mxr.Context(N)
        ref = np.zeros(N ** 3, np.float32)
        ave_ref = ctx.brightness_temperature(p, xH, dx, ref, TS if spin else None)
        ctx.close()
        dxH, ddx, dTS = (torch.from_numpy(a).cuda() for a in (xH, dx, TS))
        dT = torch.full((N ** 3,), -7.0, device="cuda")
        ave = torch.zeros(1, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        mxr.brightness_temperature_noalloc(p, dxH, ddx, dT, ave, dTS if spin else None)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(dT.cpu().numpy(), ref)
        assert float(ave.item()) == ave_ref


def test_noalloc_rejects_what_it_cannot_do():
    import torch

    N = 32
    xH, dx, TS = (torch.from_numpy(a).cuda() for a in _inputs(N, 3))
    dT = torch.zeros(N ** 3, device="cuda")
    ave = torch.zeros(1, dtype=torch.float64, device="cuda")
    with pytest.raises(mxr.MxrError) as e:  # velocity terms need a context
        mxr.brightness_temperature_noalloc(mxr.tb_params(N, 8.0, True, pecvel=1), xH, dx, dT, ave, TS)
    assert e.value.code == mxr.MXR_ERR_UNSUPPORTED
    with pytest.raises(mxr.MxrError) as e:  # host memory: refused before any launch
        mxr.brightness_temperature_noalloc(mxr.tb_params(N, 8.0, False), xH.cpu().numpy(), dx.cpu().numpy(),
                                           dT.cpu().numpy(), ave.cpu().numpy())
    assert e.value.code == mxr.MXR_ERR_INVALID_ARG
    with pytest.raises(mxr.MxrError) as e:  # beyond the largest dimension
        mxr.brightness_temperature_noalloc(mxr.tb_params(4096, 8.0, False), xH, dx, dT, ave)
    assert e.value.code == mxr.MXR_ERR_INVALID_ARG
