"""BF16-storage variant of find_HII_bubbles (stars and weighted_sfr stored as bfloat16, FP32
compute).  Correctness: bit-identical to the FP32 path fed the BF16-rounded grids (the load only
widens).  Accuracy of the storage itself: flag mismatch and |dxH| against the FP64 oracle on the
original FP32 grids, asserted against the FP32 path's bounds in a strict xfail: the 8-bit
mantissa misses them, and a change that meets them shows up as an unexpected pass."""
import numpy as np
import pytest

import mxr
import oracle_lib as O
from conftest import realisation

pytestmark = pytest.mark.gpu


def _bf16_run(N, box):
    import torch

    print(f"bf16 N={N}", flush=True)
    inp = realisation(N, box, 1)
    p = mxr.hii_params(N, box, filter_type=1)
    dev = {k: torch.from_numpy(v).cuda() for k, v in inp.items()}
    bf = dict(dev, stars=dev["stars"].to(torch.bfloat16), weighted_sfr=dev["weighted_sfr"].to(torch.bfloat16))
    rounded = dict(dev, stars=bf["stars"].float(), weighted_sfr=bf["weighted_sfr"].float())
    ctx = mxr.Context(N)
    sa = mxr.HiiState(*(t.cuda() for t in (torch.ones(N ** 3), torch.full((N ** 3,), -1.0), torch.zeros(N ** 3),
                                          torch.zeros(N ** 3), torch.zeros(N ** 3))))
    sb = mxr.HiiState(*(t.cuda() for t in (torch.ones(N ** 3), torch.full((N ** 3,), -1.0), torch.zeros(N ** 3),
                                          torch.zeros(N ** 3), torch.zeros(N ** 3))))
    ga = ctx.find_hii_bubbles(p, **bf, **sa.arrays())
    gb = ctx.find_hii_bubbles(p, **rounded, **sb.arrays())
    torch.cuda.synchronize()
    for k in sa.arrays():
        np.testing.assert_array_equal(getattr(sa, k).cpu().numpy(), getattr(sb, k).cpu().numpy(), err_msg=k)
    assert ga.volume_weighted_global_xH == gb.volume_weighted_global_xH
    assert ga.volume_weighted_global_weighted_sfr == gb.volume_weighted_global_weighted_sfr
    # storage error against the FP64 oracle on the original grids
    s64 = mxr.HiiState.host(N)
    r64 = O.find_hii_bubbles(p, inp, s64, precision="f64")
    gpu = {k: getattr(sa, k).cpu().numpy() for k in sa.arrays()}
    ref = {k: np.asarray(v) for k, v in s64.arrays().items()}
    flips = int(np.sum((gpu["xH"] < 1e-5) != (ref["xH"] < 1e-5)))
    dxh = abs(ga.volume_weighted_global_xH - r64.volume_weighted_global_xH)
    print(f"bf16 N={N}: flag mismatch {flips / N ** 3 * 100:.4f}% of cells, |dxH| {dxh:.2e}", flush=True)
    ctx.close()
    return flips, dxh


@pytest.mark.parametrize("N,box", [(64, 67.8), (128, 125.0)])
def test_bf16_storage_identity(N, box):
    """BF16 storage equals the FP32 path fed the BF16-rounded grids, bit for bit."""
    _bf16_run(N, box)


@pytest.mark.xfail(strict=True, reason="BF16 storage of stars/sfr misses the FP32 path's bounds against the FP64 "
                                       "oracle: |dxH| 6e-5 at 64^3, 0.083% flags and 6.7e-4 at 128^3")
@pytest.mark.parametrize("N,box", [(64, 67.8), (128, 125.0)])
def test_bf16_storage_meets_bounds(N, box):
    """The bounds the FP32 path meets, against the FP64 oracle on the original grids: flag mismatch
    at most 0.05% of the cells and |dxH_global| at most 1e-5."""
    flips, dxh = _bf16_run(N, box)
    assert flips <= 5e-4 * N ** 3
    assert dxh <= 1e-5
