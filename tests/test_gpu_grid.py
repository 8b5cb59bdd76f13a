"""GPU parity of the input-grid preparation (mxr_read_grid_gbptrees, mxr_prepare_grid) against
the oracle.  Without resampling the work is element-wise float arithmetic: bit for bit.  With
smoothing the two FP32 FFT round trips differ by rounding only: within 2e-6 of the field's
maximum per cell."""
import numpy as np
import pytest

import mxr
import oracle_lib as O
from grid_files import density_grids, write_gbptrees

pytestmark = pytest.mark.gpu


def _real(x, N):
    return x.reshape(N, N, 2 * (N // 2 + 1))[:, :, :N]


@pytest.mark.parametrize("prop", [mxr.MXR_GRID_DENSITY, mxr.MXR_GRID_Z_VELOCITY])
@pytest.mark.parametrize("N,n_hi", [(32, 32), (32, 64), (64, 128), (64, 256)])
def test_read_gbptrees_vs_oracle(tmp_path, N, n_hi, prop):
    box = 50.0
    f = tmp_path / "grid.dat"
    write_gbptrees(f, density_grids(n_hi, n_hi + prop), box)
    p = mxr.grid_params(box, npart=n_hi ** 3, part_mass=1.0)
    ref = O.read_grid_gbptrees(str(f), prop, N, p)
    got = np.zeros_like(ref)
    mxr.Context(N).read_grid_gbptrees(str(f), prop, p, got)
    if n_hi == N:
        assert np.array_equal(_real(got, N), _real(ref, N))
    else:
        err = np.abs(_real(got, N).astype(np.float64) - _real(ref, N)).max() / np.abs(_real(ref, N)).max()
        print(f"N={N} n_hi={n_hi} prop={prop}: max err {err:.2e} of max |field|", flush=True)
        assert err <= 2e-6
    # padding entries are zero
    assert not np.any(got.reshape(N, N, -1)[:, :, N:])


def test_prepare_grid_device_velociraptor():
    import torch
    N, n_hi, box = 64, 128, 60.0
    hi = density_grids(n_hi, 7)[0]
    p = mxr.grid_params(box, npart=n_hi ** 3, part_mass=1.0, trees_id=mxr.MXR_TREES_VELOCIRAPTOR)
    ref = O.prepare_grid(N, hi, n_hi, mxr.MXR_GRID_DENSITY, box, p)
    out = torch.zeros(ref.size, device="cuda")
    mxr.Context(N).prepare_grid(torch.from_numpy(hi).cuda(), n_hi, mxr.MXR_GRID_DENSITY, box, p, out)
    got = out.cpu().numpy()
    err = np.abs(_real(got, N).astype(np.float64) - _real(ref, N)).max() / np.abs(_real(ref, N)).max()
    assert err <= 2e-6


def test_errors(tmp_path):
    N = 32
    p = mxr.grid_params(40.0)
    odd = tmp_path / "odd.dat"
    write_gbptrees(odd, density_grids(48, 1), 40.0)
    out = np.zeros(N * N * 2 * (N // 2 + 1), np.float32)
    ctx = mxr.Context(N)
    with pytest.raises(mxr.MxrError) as e:
        ctx.read_grid_gbptrees(str(odd), 0, p, out)
    assert e.value.code == mxr.MXR_ERR_INVALID_ARG
    # a 192^3 grid is a multiple of 32 but not a dimension the FFT engine smooths
    with pytest.raises(mxr.MxrError) as e:
        ctx.prepare_grid(np.zeros(192 ** 3, np.float32), 192, 1, 40.0, p, out)
    assert e.value.code == mxr.MXR_ERR_UNSUPPORTED
    with pytest.raises(mxr.MxrError):
        ctx.read_grid_gbptrees(str(tmp_path / "missing.dat"), 0, p, out)


def test_smoothing_window_cache_and_unaligned_input():
    """The smoothing filter's window values are cached per (grid, R): a call at another box size
    rebuilds them, and returning to the first box gives the first result bit for bit, as does a
    fresh context.  A device input at an odd float offset is staged and gives the same result."""
    import torch
    N, n_hi = 32, 64
    hi = density_grids(n_hi, 5)[0]
    outs = []
    ctx = mxr.Context(N)
    for box in (50.0, 80.0, 50.0):
        p = mxr.grid_params(box, npart=n_hi ** 3, part_mass=1.0)
        out = np.zeros(N * N * 2 * (N // 2 + 1), np.float32)
        ctx.prepare_grid(hi, n_hi, mxr.MXR_GRID_DENSITY, box, p, out)
        outs.append(out)
    assert np.array_equal(outs[0], outs[2])
    assert not np.array_equal(outs[0], outs[1])
    p = mxr.grid_params(50.0, npart=n_hi ** 3, part_mass=1.0)
    fresh = np.zeros_like(outs[0])
    mxr.Context(N).prepare_grid(hi, n_hi, mxr.MXR_GRID_DENSITY, 50.0, p, fresh)
    assert np.array_equal(fresh, outs[0])
    buf = torch.zeros(hi.size + 1, device="cuda")
    buf[1:] = torch.from_numpy(hi.reshape(-1)).cuda()
    dev = torch.zeros(outs[0].size, device="cuda")
    ctx.prepare_grid(buf[1:], n_hi, mxr.MXR_GRID_DENSITY, 50.0, p, dev)
    assert np.array_equal(dev.cpu().numpy(), outs[0])
    ref = O.prepare_grid(N, hi, n_hi, mxr.MXR_GRID_DENSITY, 80.0, mxr.grid_params(80.0, npart=n_hi ** 3, part_mass=1.0))
    err = np.abs(_real(outs[1], N).astype(np.float64) - _real(ref, N)).max() / np.abs(_real(ref, N)).max()
    assert err <= 2e-6


def test_prepare_grid_512_to_256_vs_oracle():
    """The benchmark's smoothing size: a 512^3 density grid smoothed and subsampled to 256^3,
    within 2e-6 of the field's maximum of the oracle, padding zero."""
    import torch
    N, n_hi, box = 256, 512, 400.0
    hi = np.random.default_rng(512).uniform(0, 16, n_hi ** 3).astype(np.float32)
    p = mxr.grid_params(box, npart=n_hi ** 3, part_mass=1.0)
    ref = O.prepare_grid(N, hi, n_hi, mxr.MXR_GRID_DENSITY, box, p)
    out = torch.zeros(ref.size, device="cuda")
    mxr.Context(N).prepare_grid(torch.from_numpy(hi).cuda(), n_hi, mxr.MXR_GRID_DENSITY, box, p, out)
    got = out.cpu().numpy()
    err = np.abs(_real(got, N).astype(np.float64) - _real(ref, N)).max() / np.abs(_real(ref, N)).max()
    print(f"512 -> 256: max err {err:.2e} of max |field|", flush=True)
    assert err <= 2e-6
    assert not np.any(got.reshape(N, N, -1)[:, :, N:])
