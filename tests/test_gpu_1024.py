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


def _tile_real(torch, a, n_small, rep):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda().view(n_small, n_small, n_small).repeat(
        rep, rep, rep).view(-1)


def _tile_padded(torch, a, n_small, rep, count=1):
    """`count` padded n_small^3 grids, each tiled rep^3 times into a padded (n_small rep)^3 grid."""
    nb = n_small * rep
    Hs, Hb = n_small // 2 + 1, nb // 2 + 1
    src = torch.from_numpy(np.ascontiguousarray(a)).cuda().view(count, n_small, n_small, 2 * Hs)[..., :n_small]
    big = torch.zeros(count, nb, nb, 2 * Hb, dtype=torch.float32, device="cuda")
    big[..., :nb] = src.repeat(1, rep, rep, rep)
    return big.view(-1)


@pytest.mark.gpu
def test_brightness_1024_tiled():
    """The 21-cm brightness box with peculiar velocities (optical-depth form, PecVels 2) at
    1024^3 against the tiled 128^3 oracle: the conversion is per cell and dv/dz comes from the
    same modes (k_z = 8 n dk / 8), so the tiled result holds up to FFT rounding; the bounds of
    test_gpu_tb.py (1e-4 per cell at the 99.9th percentile, 2e-3 at worst, the mean within 1e-5
    of mean |delta_T|)."""
    import math

    import torch

    print("brightness N=1024 (tiled 128^3)", flush=True)
    n, box, rep = N_SMALL, BOX_SMALL, REP
    p_small = mxr.tb_params(n, 8.0, True, box=box, pecvel=2)
    rng = np.random.default_rng(5)
    H = n // 2 + 1

    def pad(x):
        out = np.zeros((n, n, 2 * H), np.float32)
        out[:, :, :n] = x
        return out.reshape(-1)

    dx = pad((0.3 * rng.standard_normal((n, n, n))).astype(np.float32))
    xh = rng.uniform(0.0, 1.0, (n, n, n)).astype(np.float32)
    xh[rng.uniform(size=xh.shape) < 0.2] = 0.0
    TS = rng.uniform(5.0, 60.0, n ** 3).astype(np.float32)
    opz = float(np.float32(1.0) + np.float32(p_small.redshift))
    Hz = 3.2407789e-18 * p_small.Hubble_h * math.sqrt(p_small.OmegaM * opz ** 3 + p_small.OmegaLambda)
    cell = box / p_small.Hubble_h / n
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    disp = cell * (2.3 * np.sin(2 * np.pi * k / n + 0.7 * np.sin(2 * np.pi * i / n))
                   + 0.6 * np.cos(2 * np.pi * 3 * k / n + 1.3 * np.cos(2 * np.pi * j / n))
                   + 0.2 * rng.standard_normal((n, n, n)))
    v = pad((disp * Hz / (opz * opz * 1e5 / 3.086e24)).astype(np.float32))
    ref, ave_ref = O.brightness_temperature(p_small, xh.reshape(-1), dx, TS, vel=v)

    p_big = mxr.tb_params(N_BIG, 8.0, True, box=BOX_BIG, pecvel=2)
    dxb, vb = _tile_padded(torch, dx, n, rep), _tile_padded(torch, v, n, rep)
    xhb, TSb = _tile_real(torch, xh, n, rep), _tile_real(torch, TS, n, rep)
    dT = torch.zeros(N_BIG ** 3, dtype=torch.float32, device="cuda")
    ctx = mxr.Context(N_BIG)
    try:
        ave = ctx.brightness_temperature(p_big, xhb, dxb, dT, TSb, vel=vb)
        torch.cuda.synchronize()
    finally:
        ctx.close()
    del dxb, vb, xhb, TSb
    rt = _tile_real(torch, ref.astype(np.float32), n, rep).double()
    scale = 1e-3 * float(np.abs(ref).max())
    e = (dT.double() - rt).abs() / rt.abs().clamp_min(scale)
    p999 = float(torch.quantile(e[::97].float(), 0.999))
    mx = float(e.max())
    print(f"N=1024 delta_T rel err p99.9 {p999:.2e} max {mx:.2e}, mean {ave:.6e} vs {ave_ref:.6e}", flush=True)
    assert p999 <= 1e-4 and mx <= 2e-3
    assert abs(ave - ave_ref) <= 1e-5 * max(abs(ave_ref), float(np.abs(ref).mean()))


@pytest.mark.gpu
def test_compute_ts_1024_tiled():
    """One spin-temperature step at 1024^3 (10 shells, snapshot 3, z = 19.1: an early step whose
    shells reach back past snapshot 0, so the history plan clears and reloads) against the tiled
    128^3 oracles.  The shells' radii come from the cell size and the plan from the redshifts,
    both shared by the two boxes.  The state is a strong-heating one (cells near the 0.1 K floor,
    see test_gpu_ts.py::test_cold_cells_stop_like_the_reference), so the bounds are FP64-anchored:
    against the tiled FP64-smoothed oracle, the GPU's 99.9th-percentile error at most twice the
    FP32 oracle's own plus 2.5e-5 and its cells beyond 1e-2 at most twice the FP32 oracle's plus
    0.01%; each volume average's distance from the FP64 oracle's at most twice the FP32 oracle's
    plus 1e-5 relative (the cells past the discontinuities move the T_S mean by ~1e-4)."""
    import torch

    from ts_cases import inputs, params, state

    torch.cuda.empty_cache()  # the earlier 1024^3 tests' cached blocks
    print("ComputeTs N=1024 (tiled 128^3, 10 shells)", flush=True)
    n, box, rep, snap, F = N_SMALL, BOX_SMALL, REP, 3, 10
    p_small = params(n, box, snap, TsNumFilterSteps=F)
    p_big = params(N_BIG, BOX_BIG, snap, TsNumFilterSteps=F)
    sf = np.zeros(41)
    sf[:snap] = 1e-4 * np.arange(snap)
    t_small = mxr.ts_build_tables(p_small, 1e-3, 2e-4, sf.copy())
    t_big = mxr.ts_build_tables(p_big, 1e-3, 2e-4, sf.copy())
    assert t_small.ops() == t_big.ops() and t_small.n == t_big.n == F
    np.testing.assert_allclose(t_big.arr("R"), t_small.arr("R"), rtol=1e-12)
    depth = max(1, t_small.hist_max + 1)
    d = inputs(n, box, snap, depth, scale=1e-2)
    xe, Tk = state(n, d["deltax"])

    # GPU at 1024^3 on the tiled inputs (device grids)
    dev = {k: _tile_padded(torch, d[k], n, rep) for k in ("deltax", "stars")}
    hs = _tile_padded(torch, d["sfr_histories"], n, rep, depth)
    hx = _tile_padded(torch, d["xray_luminosity_histories"], n, rep, depth)
    xe_g = _tile_padded(torch, xe, n, rep)
    Tk_g = _tile_real(torch, Tk, n, rep)
    xe_box = torch.zeros_like(xe_g)
    TS = torch.zeros(N_BIG ** 3, dtype=torch.float32, device="cuda")
    ctx = mxr.Context(N_BIG)
    try:
        g, tab = ctx.compute_ts(p_big, dev["deltax"], dev["stars"], hs, hx, depth, xe_g, xe_box, Tk_g, TS, sf.copy())
        torch.cuda.synchronize()
        print(f"N=1024 ComputeTs {g.device_ms:.1f} ms", flush=True)
    finally:
        ctx.close()
    del dev, hs, hx, xe_box
    assert tab.homogeneous == 0 and tab.NO_LIGHT == 0

    args = (p_small, tab, d["deltax"], d["stars"], d["sfr_histories"], d["xray_luminosity_histories"], depth, xe, Tk)
    ref = O.ts_step(*args)
    ref64 = O.ts_step(*args, f64=True)
    H = n // 2 + 1
    Hb = N_BIG // 2 + 1
    xe_cells = xe_g.view(N_BIG, N_BIG, 2 * Hb)[:, :, :N_BIG].reshape(-1)

    def cells(r, k):
        a = r[k] if k != "x_e" else r["x_e_box_prev"].reshape(n, n, 2 * H)[:, :, :n].reshape(-1)
        return np.asarray(a, np.float64)

    for name, got, floor in (("TS_box", TS, 0.0), ("Tk_box", Tk_g, 0.0), ("x_e", xe_cells, 1e-6)):
        assert bool(torch.isfinite(got).all()), name
        w64 = _tile_real(torch, cells(ref64, name), n, rep)
        eg = (got.double() - w64).abs() / w64.abs().clamp_min(floor if floor else 1e-300)
        eo = np.abs(cells(ref, name) - cells(ref64, name)) / np.maximum(np.abs(cells(ref64, name)),
                                                                        floor if floor else 1e-300)
        pg = float(torch.quantile(eg[::97].float(), 0.999))
        po = float(np.percentile(eo, 99.9))
        fg, fo = int((eg > 1e-2).sum()), int((eo > 1e-2).sum()) * rep ** 3
        print(f"N=1024 {name}: vs FP64 p99.9 gpu {pg:.2e} fp32-oracle {po:.2e}; beyond 1e-2 gpu {fg} "
              f"fp32-oracle (tiled) {fo}", flush=True)
        assert pg <= 2 * po + 2.5e-5, (name, pg, po)
        assert fg <= 2 * fo + 1e-4 * N_BIG ** 3, (name, fg, fo)
    for k in ("volume_ave_TS", "volume_ave_TK", "volume_ave_xe", "volume_ave_J_alpha", "volume_ave_Xheat"):
        r32, r64 = ref["globals"][k], ref64["globals"][k]
        print(f"N=1024 {k}: gpu {getattr(g, k):.9e} fp32-oracle {r32:.9e} fp64-oracle {r64:.9e}", flush=True)
        assert abs(getattr(g, k) - r64) <= 2 * abs(r32 - r64) + 1e-5 * abs(r64), k
