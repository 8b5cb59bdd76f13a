"""GPU parity of the spin-temperature step with AGN X-ray bands (Flag_IncludeAGNXray 1-3,
mxr_compute_ts through the C ABI) against the oracle (oracle/ts.c, checked on the CPU in
test_ts_agn.py).

Bounds, as for the galaxy-only step: T_S, T_K and x_e within 1e-4 relative per cell at the
99.9th percentile against the FP32 restatement, the largest error against the FP64-smoothed
oracle no larger than twice the FP32 restatement's own plus a quarter of that tolerance; globals,
the bands' heating averages and their R_ct = 0 emissivities within 1e-5.  Zero AGN grids leave
every cell bit for bit as Flag_IncludeAGNXray = 0 leaves it; device grids, a ring history layout
and 64 shells (the largest table block) are covered.
"""
import numpy as np
import pytest

import mxr
import oracle_lib as ol
from ts_cases import ZZ, agn_histories, inputs, params, rel_err, state

pytestmark = pytest.mark.gpu

TOL = 1e-4


@pytest.fixture(scope="module")
def ctxs():
    cache = {}

    def get(N):
        if N not in cache:
            cache[N] = mxr.Context(N)
        return cache[N]

    yield get
    for c in cache.values():
        c.close()


def _cells(N, padded):
    H = N // 2 + 1
    return padded.reshape(N, N, 2 * H)[:, :, :N].reshape(-1)


def _inputs(N, box, snap, flag, scale, level, **over):
    p = params(N, box, snap, Flag_IncludeAGNXray=flag, **over)
    depth = max(1, mxr.ts_history_depth(p, 40))
    d = inputs(N, box, snap, depth, scale=scale)
    soft, hard = agn_histories(N, d, depth, level=level)
    xe, Tk = state(N, d["deltax"])
    sf = np.zeros(len(ZZ))
    sf[:snap] = 1e-3 * np.arange(snap) / max(snap, 1)
    return p, depth, d, soft, hard, xe, Tk, sf


def _gpu(ctx, p, d, depth, soft, hard, xe, Tk, sf, history_head=0):
    N = p.ReionGridDim
    xe_g, Tk_g = xe.copy(), Tk.copy()
    xe_box = np.zeros_like(xe)
    TS = np.zeros(N ** 3, np.float32)
    g, tab = ctx.compute_ts(p, d["deltax"], d["stars"], d["sfr_histories"], d["xray_luminosity_histories"], depth,
                            xe_g, xe_box, Tk_g, TS, sf.copy(), history_head=history_head,
                            bh_xray_histories_soft=soft, bh_xray_histories_hard=hard)
    return dict(x_e_box_prev=xe_g, x_e_box=xe_box, Tk_box=Tk_g, TS_box=TS, globals=g.as_dict()), tab


def _oracle(p, tab, d, depth, soft, hard, xe, Tk, f64=False):
    return ol.ts_step(p, tab, d["deltax"], d["stars"], d["sfr_histories"], d["xray_luminosity_histories"], depth, xe,
                      Tk, f64=f64, bh_soft_histories=soft, bh_hard_histories=hard)


def _compare(N, gpu, ref, tag, ref64=None):
    fields = {"TS_box": (lambda r: r["TS_box"], 0.0), "Tk_box": (lambda r: r["Tk_box"], 0.0),
              "x_e": (lambda r: _cells(N, r["x_e_box_prev"]), 1e-6)}
    for k, (get, floor) in fields.items():
        e = rel_err(get(gpu), get(ref), floor)
        p999, mx = float(np.percentile(e, 99.9)), float(e.max())
        line = f"{tag} {k}: vs FP32 p99.9 {p999:.2e} max {mx:.2e}"
        assert p999 <= TOL, (k, p999)
        if ref64 is not None:
            eg = rel_err(get(gpu), get(ref64), floor).max()
            eo = rel_err(get(ref), get(ref64), floor).max()
            line += f" | vs FP64 max: gpu {eg:.2e} fp32-oracle {eo:.2e}"
            assert eg <= 2 * eo + TOL / 4, (k, eg, eo)
        print(line, flush=True)
    gg, rg = gpu["globals"], ref["globals"]
    for k in ("volume_ave_TS", "volume_ave_TK", "volume_ave_xe", "volume_ave_J_alpha", "volume_ave_xalpha",
              "volume_ave_Xheat", "volume_ave_Xion", "volume_ave_Xheat_AGN_soft", "volume_ave_Xheat_AGN_hard",
              "XrayEmissivity_soft", "XrayEmissivity_hard"):
        assert gg[k] == pytest.approx(rg[k], rel=1e-5, abs=1e-300), k
    assert gg["XrayEmissivity_HMXB"] == pytest.approx(ref["red0"][2], rel=1e-5)
    assert gg["XrayEmissivity_soft"] == pytest.approx(ref["red0"][3], rel=1e-5, abs=1e-300)
    assert gg["XrayEmissivity_hard"] == pytest.approx(ref["red0"][4], rel=1e-5, abs=1e-300)


@pytest.mark.parametrize("N,box,snap,flag,scale,level",
                         [(32, 40.0, 14, 1, 1e-2, 1.0), (32, 40.0, 30, 2, 1e-3, 3.0), (32, 40.0, 30, 3, 1e-3, 3.0),
                          (64, 67.8, 25, 1, 1e-2, 1.0)])
def test_agn_parity_step(ctxs, N, box, snap, flag, scale, level):
    print(f"ts AGN step N={N} snap={snap} flag={flag}", flush=True)
    p, depth, d, soft, hard, xe, Tk, sf = _inputs(N, box, snap, flag, scale, level)
    gpu, tab = _gpu(ctxs(N), p, d, depth, soft, hard, xe, Tk, sf)
    assert tab.homogeneous == 0 and tab.NO_LIGHT == 0
    g = gpu["globals"]
    if flag in (1, 3):
        assert g["volume_ave_Xheat_AGN_soft"] < 0 and g["XrayEmissivity_soft"] > 0
    else:
        assert g["volume_ave_Xheat_AGN_soft"] == 0 and g["XrayEmissivity_soft"] == 0
    if flag in (1, 2):
        assert g["volume_ave_Xheat_AGN_hard"] < 0 and g["XrayEmissivity_hard"] > 0
    else:
        assert g["volume_ave_Xheat_AGN_hard"] == 0 and g["XrayEmissivity_hard"] == 0
    ref = _oracle(p, tab, d, depth, soft, hard, xe, Tk)
    ref64 = _oracle(p, tab, d, depth, soft, hard, xe, Tk, f64=True) if N == 32 else None
    _compare(N, gpu, ref, f"AGN N={N} snap={snap} flag={flag}", ref64)


def test_agn_zero_grids_bit_identical(ctxs):
    """The two-band kernels with zero AGN grids leave every cell as the galaxy-only kernels do."""
    N, box, snap = 32, 40.0, 20
    print("ts AGN zero grids", flush=True)
    p1, depth, d, soft, hard, xe, Tk, sf = _inputs(N, box, snap, 1, 1e-2, 1.0)
    z = np.zeros_like(soft)
    a, _ = _gpu(ctxs(N), p1, d, depth, z, z, xe, Tk, sf)
    p0 = params(N, box, snap, Flag_IncludeAGNXray=0)
    b, _ = _gpu(ctxs(N), p0, d, depth, None, None, xe, Tk, sf)
    for k in ("x_e_box_prev", "Tk_box", "TS_box"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    assert a["globals"]["volume_ave_TS"] == pytest.approx(b["globals"]["volume_ave_TS"], rel=1e-12)


def test_agn_device_grids_and_ring_layout(ctxs):
    """Device grids with the histories kept as a ring (history_head 2) give the host call's cells."""
    import torch

    N, box, snap = 32, 40.0, 14
    print("ts AGN device + ring", flush=True)
    p, depth, d, soft, hard, xe, Tk, sf = _inputs(N, box, snap, 1, 1e-2, 1.0)
    ref, _ = _gpu(ctxs(N), p, d, depth, soft, hard, xe, Tk, sf)
    npad = mxr.padded_len(N)
    head = 2

    def ring(h):
        a = h.reshape(depth, npad)
        out = np.empty_like(a)
        for s in range(depth):
            out[(head + s) % depth] = a[s]
        return torch.from_numpy(out.reshape(-1)).cuda()

    dev = {k: torch.from_numpy(v).cuda() for k, v in d.items() if k in ("deltax", "stars")}
    xe_g, Tk_g = torch.from_numpy(xe.copy()).cuda(), torch.from_numpy(Tk.copy()).cuda()
    xe_box = torch.zeros_like(xe_g)
    TS = torch.zeros(N ** 3, dtype=torch.float32, device="cuda")
    g, _ = ctxs(N).compute_ts(p, dev["deltax"], dev["stars"], ring(d["sfr_histories"]),
                              ring(d["xray_luminosity_histories"]), depth, xe_g, xe_box, Tk_g, TS, sf.copy(),
                              history_head=head, bh_xray_histories_soft=ring(soft), bh_xray_histories_hard=ring(hard))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(TS.cpu().numpy(), ref["TS_box"])
    np.testing.assert_array_equal(Tk_g.cpu().numpy(), ref["Tk_box"])
    np.testing.assert_array_equal(xe_g.cpu().numpy(), ref["x_e_box_prev"])
    assert g.volume_ave_Xheat_AGN_hard == pytest.approx(ref["globals"]["volume_ave_Xheat_AGN_hard"], rel=1e-12)


def test_agn_64_shells(ctxs):
    """64 shells with both bands: the largest table block (130 KB of LDS)."""
    N, box, snap = 32, 40.0, 30
    print("ts AGN 64 shells", flush=True)
    p, depth, d, soft, hard, xe, Tk, sf = _inputs(N, box, snap, 1, 1e-2, 1.0, TsNumFilterSteps=64)
    gpu, tab = _gpu(ctxs(N), p, d, depth, soft, hard, xe, Tk, sf)
    assert tab.n == 64
    ref = _oracle(p, tab, d, depth, soft, hard, xe, Tk)
    _compare(N, gpu, ref, "AGN 64 shells")


def test_agn_missing_histories_is_an_error(ctxs):
    N, box, snap = 32, 40.0, 14
    p, depth, d, soft, hard, xe, Tk, sf = _inputs(N, box, snap, 2, 1e-2, 1.0)
    with pytest.raises(mxr.MxrError) as e:
        _gpu(ctxs(N), p, d, depth, soft, None, xe, Tk, sf)
    assert e.value.code == mxr.MXR_ERR_INVALID_ARG
