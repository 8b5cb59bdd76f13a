"""GPU parity of the spin-temperature step (mxr_compute_ts through the C ABI) against the oracle.

Each step feeds the HIP path and oracle/ts.c the same inputs and state; the oracle uses the
grid-independent tables the library reports for that call (checked on the CPU in
test_ts_tables.py).  Bound: T_S, T_K and x_e within 1e-4 relative per cell, globals within
1e-5, the R_ct = 0 averages within 1e-9.
"""
import numpy as np
import pytest

import mxr
import oracle_lib as ol
from ts_cases import ZZ, inputs, params, rel_err, state

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


def _step(ctx, p, d, depth, xe, Tk, sf):
    N = p.ReionGridDim
    xe_g, Tk_g = xe.copy(), Tk.copy()
    xe_box = np.zeros_like(xe)
    TS = np.zeros(N ** 3, np.float32)
    g, tab = ctx.compute_ts(p, d["deltax"], d["stars"], d["sfr_histories"], d["xray_luminosity_histories"], depth,
                            xe_g, xe_box, Tk_g, TS, sf)
    return dict(x_e_box_prev=xe_g, x_e_box=xe_box, Tk_box=Tk_g, TS_box=TS, globals=g.as_dict()), tab


def _compare(N, gpu, ref, tab, tag):
    stats = {}
    for k in ("TS_box", "Tk_box"):
        e = rel_err(gpu[k], ref[k])
        stats[k] = (float(np.percentile(e, 99.9)), float(e.max()))
    e = rel_err(_cells(N, gpu["x_e_box_prev"]), _cells(N, ref["x_e_box_prev"]), floor=1e-6)
    stats["x_e"] = (float(np.percentile(e, 99.9)), float(e.max()))
    print(tag, {k: "p99.9 %.2e max %.2e" % v for k, v in stats.items()}, flush=True)
    for k, (p999, mx) in stats.items():
        assert mx <= TOL, (k, p999, mx)
    np.testing.assert_array_equal(gpu["x_e_box"], gpu["x_e_box_prev"])
    gg, rg = gpu["globals"], ref["globals"]
    keys = ["volume_ave_TS", "volume_ave_TK", "volume_ave_xe"]
    if not tab.homogeneous:
        keys += ["volume_ave_J_alpha", "volume_ave_xalpha", "volume_ave_Xheat", "volume_ave_Xion"]
        assert tab.collapse_fraction == pytest.approx(ref["red0"][0], rel=1e-9, abs=1e-300)
        assert tab.x_e_ave == pytest.approx(ref["red0"][1], rel=1e-9)
    for k in keys:
        assert gg[k] == pytest.approx(rg[k], rel=1e-5, abs=1e-30), k


def _case(ctx, N, box, snap, scale=1.0, seed=7, **over):
    p = params(N, box, snap, **over)
    depth = max(1, mxr.ts_history_depth(p, 40))
    d = inputs(N, box, snap, depth, scale=scale)
    xe, Tk = state(N, d["deltax"], seed=seed)
    sf = np.zeros(len(ZZ))
    sf[:snap] = 1e-3 * np.arange(snap) / max(snap, 1)
    gpu, tab = _step(ctx, p, d, depth, xe, Tk, sf)
    ref = ol.ts_step(p, tab, d["deltax"], d["stars"], d["sfr_histories"], d["xray_luminosity_histories"], depth,
                     xe, Tk)
    return gpu, ref, tab


def test_homogeneous_branch(ctxs):
    for N, box in [(16, 20.0), (64, 67.8)]:
        print(f"ts homogeneous N={N}", flush=True)
        gpu, ref, tab = _case(ctxs(N), N, box, 0)
        assert tab.homogeneous == 1
        _compare(N, gpu, ref, tab, f"homog N={N}")


@pytest.mark.parametrize("N,box,snap,scale", [(32, 40.0, 14, 1.0), (32, 40.0, 14, 1e-3), (32, 40.0, 40, 1e-3),
                                              (64, 67.8, 25, 1e-2), (64, 67.8, 40, 1.0)])
def test_parity_step(ctxs, N, box, snap, scale):
    print(f"ts step N={N} snap={snap} scale={scale}", flush=True)
    gpu, ref, tab = _case(ctxs(N), N, box, snap, scale)
    assert tab.homogeneous == 0 and tab.NO_LIGHT == 0
    _compare(N, gpu, ref, tab, f"step N={N} snap={snap} scale={scale}")


@pytest.mark.parametrize("ftype", [0, 2])
def test_parity_filter_types(ctxs, ftype):
    print(f"ts filter type {ftype}", flush=True)
    gpu, ref, tab = _case(ctxs(32), 32, 40.0, 20, 1e-2, TsHeatingFilterType=ftype)
    _compare(32, gpu, ref, tab, f"filter {ftype}")


def test_parity_non_pow2(ctxs):
    print("ts N=96", flush=True)
    gpu, ref, tab = _case(ctxs(96), 96, 100.0, 30, 1e-2)
    _compare(96, gpu, ref, tab, "N=96")


def test_sequence(ctxs):
    """Snapshots 0..4 in order: homogeneous start, then the clear-and-reload history edge cases
    of the first inhomogeneous snapshots; every step checked against the oracle on the same
    input state."""
    N, box = 32, 40.0
    ctx = ctxs(N)
    sf = np.zeros(len(ZZ))
    xe = Tk = None
    for snap in range(5):
        print(f"ts sequence snap={snap}", flush=True)
        p = params(N, box, snap)
        depth = max(1, mxr.ts_history_depth(p, 40))
        d = inputs(N, box, snap, depth, scale=1e-2)
        if xe is None:
            xe, Tk = state(N, d["deltax"])
        gpu, tab = _step(ctx, p, d, depth, xe, Tk, sf)
        ref = ol.ts_step(p, tab, d["deltax"], d["stars"], d["sfr_histories"], d["xray_luminosity_histories"], depth,
                         xe, Tk)
        _compare(N, gpu, ref, tab, f"seq snap={snap}")
        if snap > 0:
            assert sf[snap] == tab.collapse_fraction
        xe, Tk = gpu["x_e_box_prev"], gpu["Tk_box"]


def test_device_memory_matches_host(ctxs):
    import torch

    N, box, snap = 32, 40.0, 30
    ctx = ctxs(N)
    p = params(N, box, snap)
    depth = mxr.ts_history_depth(p, 40)
    d = inputs(N, box, snap, depth, scale=1e-2)
    xe, Tk = state(N, d["deltax"])
    sf = np.zeros(len(ZZ))
    host, _ = _step(ctx, p, d, depth, xe, Tk, sf.copy())
    dev = {k: torch.from_numpy(v).cuda() for k, v in d.items()}
    xe_d, Tk_d = torch.from_numpy(xe.copy()).cuda(), torch.from_numpy(Tk.copy()).cuda()
    xe_box = torch.zeros_like(xe_d)
    TS = torch.zeros(N ** 3, dtype=torch.float32, device="cuda")
    g, _ = ctx.compute_ts(p, dev["deltax"], dev["stars"], dev["sfr_histories"], dev["xray_luminosity_histories"],
                          depth, xe_d, xe_box, Tk_d, TS, sf.copy())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(TS.cpu().numpy(), host["TS_box"])
    np.testing.assert_array_equal(Tk_d.cpu().numpy(), host["Tk_box"])
    np.testing.assert_array_equal(xe_d.cpu().numpy(), host["x_e_box_prev"])
    assert g.volume_ave_TS == host["globals"]["volume_ave_TS"]


def test_deterministic_and_bounded_256(ctxs):
    """At the benchmark size (256^3, 40 shells): two calls on the same input are identical, and
    the state stays physical (0 <= x_e <= 1, T_K, T_S > 0, finite)."""
    N, box, snap = 256, 200.0, 40
    print("ts N=256", flush=True)
    ctx = ctxs(N)
    p = params(N, box, snap)
    depth = mxr.ts_history_depth(p, 40)
    d = inputs(N, box, snap, depth, scale=1e-2)
    xe, Tk = state(N, d["deltax"])
    sf = np.zeros(len(ZZ))
    a, tab = _step(ctx, p, d, depth, xe, Tk, sf.copy())
    b, _ = _step(ctx, p, d, depth, xe, Tk, sf.copy())
    for k in ("TS_box", "Tk_box", "x_e_box_prev"):
        np.testing.assert_array_equal(a[k], b[k])
    x = _cells(N, a["x_e_box_prev"])
    assert np.all(np.isfinite(a["TS_box"])) and np.all(a["TS_box"] > 0)
    assert np.all(a["Tk_box"] > 0) and np.all((x >= 0) & (x <= 1))
    assert tab.n == 40 and tab.hist_max >= 1


def test_too_few_histories_is_an_error(ctxs):
    N, box, snap = 32, 40.0, 30
    ctx = ctxs(N)
    p = params(N, box, snap)
    depth = mxr.ts_history_depth(p, 40)
    d = inputs(N, box, snap, depth, scale=1e-2)
    xe, Tk = state(N, d["deltax"])
    with pytest.raises(mxr.MxrError) as e:
        _step(ctx, p, d, 1, xe, Tk, np.zeros(len(ZZ)))
    assert e.value.code == mxr.MXR_ERR_INVALID_ARG
