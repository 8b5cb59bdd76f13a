"""find_HII_bubbles at 1024^3 (the largest compiled dimension) against the oracle, through a
property that does not need the oracle at 1024^3.

A 1024^3 box of 1000 cMpc/h that tiles a 128^3 box of 125 cMpc/h 8 x 8 x 8 times has the same
cell size, so the radius schedule is the same (it runs from min(R_max, L_FACTOR box) down to
the cell size; checked on the CPU below).  Its spectrum is the small box's, scaled by 8^3, on
the modes whose indices are multiples of 8, and zero elsewhere; the 1024^3 dk is exactly 1/8 of
the 128^3 one, so every kept mode's |k| and window value are the small box's bit for bit.  The
1024^3 result is therefore the 128^3 result tiled, up to FFT rounding: it is compared, cell by
cell over all 1.07e9 cells, with the tiled result of the 128^3 FP32 oracle at the parity bounds
of tests/parity.py (flag mismatch <= 0.05% of cells, each at the threshold; global xH within
1e-5; z_at_ionization identical where the flags agree; the last step's partial xH within 1e-4
at the 99.9th percentile).
"""
import numpy as np
import pytest

import mxr
import oracle_lib as O
from conftest import realisation
from parity import FLAG_FRAC, MARGIN_TOL

N_SMALL, BOX_SMALL, REP = 128, 125.0, 8
N_BIG, BOX_BIG = N_SMALL * REP, BOX_SMALL * REP


def test_tiled_boxes_share_the_radius_schedule():
    small = mxr.radii(mxr.hii_params(N_SMALL, BOX_SMALL, filter_type=1))
    big = mxr.radii(mxr.hii_params(N_BIG, BOX_BIG, filter_type=1))
    assert len(small) == len(big) > 30
    np.testing.assert_allclose(big, small, rtol=1e-12)
    assert 1024 in mxr.supported_dims()


@pytest.mark.gpu
def test_parity_1024_tiled():
    import torch

    print("parity N=1024 (tiled 128^3)", flush=True)
    inp = realisation(N_SMALL, BOX_SMALL, 1)
    p_small = mxr.hii_params(N_SMALL, BOX_SMALL, filter_type=1)
    sr = mxr.HiiState.host(N_SMALL)
    rg, margin = O.find_hii_bubbles(p_small, inp, sr, margin=True)

    Hs, Hb = N_SMALL // 2 + 1, N_BIG // 2 + 1

    def tile_padded(a):
        t = torch.from_numpy(a).cuda().view(N_SMALL, N_SMALL, 2 * Hs)[:, :, :N_SMALL]
        big = torch.zeros(N_BIG, N_BIG, 2 * Hb, dtype=torch.float32, device="cuda")
        big[:, :, :N_BIG] = t.repeat(REP, REP, REP)
        return big.view(-1)

    dev = {k: tile_padded(v) for k, v in inp.items()}
    n = N_BIG ** 3
    st = mxr.HiiState(torch.ones(n, device="cuda"), torch.full((n,), -1.0, device="cuda"),
                      torch.zeros(n, device="cuda"), torch.zeros(n, device="cuda"), torch.zeros(n, device="cuda"))
    ctx = mxr.Context(N_BIG)
    try:
        p_big = mxr.hii_params(N_BIG, BOX_BIG, filter_type=1)
        gg = ctx.find_hii_bubbles(p_big, **dev, **st.arrays())
        torch.cuda.synchronize()
        print(f"N=1024 call {gg.device_ms:.1f} ms, {len(mxr.radii(p_big))} radii", flush=True)
    finally:
        ctx.close()
    del dev

    def tile_real(a):
        return torch.from_numpy(np.ascontiguousarray(a)).cuda().view(N_SMALL, N_SMALL, N_SMALL).repeat(
            REP, REP, REP).view(-1)

    ref_xH = tile_real(sr.xH)
    ion_g, ion_r = st.xH == 0, ref_xH == 0
    mism = ion_g != ion_r
    nm = int(mism.sum())
    ok = ~mism
    stats = {"mismatch": nm, "mismatch_frac": nm / n,
             "dxH_global": abs(gg.volume_weighted_global_xH - rg.volume_weighted_global_xH)}
    if nm:
        mt = tile_real(margin.astype(np.float32))
        stats["max_margin_mismatch"] = float(mt[mism].abs().max())
        del mt
    dx = (st.xH - ref_xH).abs()[ok]
    stats["max_dxH_cell"] = float(dx.max())
    stats["p999_dxH_cell"] = float(torch.quantile(dx[::97].double(), 0.999))
    del dx, ref_xH
    zr = tile_real(sr.z_at_ionization)
    stats["z_in_diff_cells"] = int(((st.z_at_ionization != zr) & ok).sum())
    del zr
    tr = tile_real(sr.temp_kinetic_all_gas)
    rel = (st.temp_kinetic_all_gas - tr).abs() / tr.abs().clamp_min(1e-30)
    stats["max_rel_T"] = float(rel[ok].max())
    print(stats, flush=True)
    assert nm <= FLAG_FRAC * n, stats
    if nm:
        assert stats["max_margin_mismatch"] <= MARGIN_TOL, stats
    assert stats["dxH_global"] <= 1e-5 + nm / n, stats
    assert stats["p999_dxH_cell"] <= 1e-4, stats
    assert stats["max_dxH_cell"] <= 1e-2, stats
    assert stats["z_in_diff_cells"] == 0, stats
    assert stats["max_rel_T"] <= 1e-4, stats
