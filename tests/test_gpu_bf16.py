"""BF16-storage variant of find_HII_bubbles (stars and weighted_sfr stored as bfloat16, FP32
compute).  Correctness: bit-identical to the FP32 path fed the BF16-rounded grids (the load only
widens).  Accuracy of the storage itself: flag mismatch and |dxH| against the FP64 oracle on the
original FP32 grids, reported (the 0.05% flag bound is not expected to hold for BF16)."""
import numpy as np
import pytest

import mxr
import oracle_lib as O
from conftest import realisation

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,box", [(64, 67.8), (128, 125.0)])
def test_bf16_storage(N, box):
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
    print(f"bf16 N={N}: flag mismatch {flips / N ** 3 * 100:.4f}% of cells, "
          f"|dxH| {abs(ga.volume_weighted_global_xH - r64.volume_weighted_global_xH):.2e}", flush=True)
    assert abs(ga.volume_weighted_global_xH - r64.volume_weighted_global_xH) < 0.05
    ctx.close()
