"""Ring-buffer histories: mxr_construct_baryon_grids with a history head keeps each history as a
ring (no slot moves), and mxr_compute_ts reads a ring through mxr_ts_grids.history_head.  Both
must give, bit for bit, what the reference's shifted layout gives: the ring read in logical order
(slot (head + s) % n) equals the shifted layout's slot s on every real cell, and ComputeTs on a
rotated history with its head equals ComputeTs on the reference layout."""
import numpy as np
import pytest

import mxr
import synth
from ts_cases import ZZ, inputs, params, state

pytestmark = pytest.mark.gpu


def _real(x, N):
    return x.reshape(-1, N, N, 2 * (N // 2 + 1))[..., :N]


@pytest.mark.parametrize("device", [False, True])
def test_baryon_ring_equals_shifted_layout(device):
    N, box, nstore = 32, 40.0, 4
    p = mxr.baryon_params(N, box, spin_temp=True, bh=True, agn_xray=mxr.MXR_AGN_XRAY_BOTH, nstore=nstore,
                          BlackHoleMassLimitReion=1e-5)
    g0 = synth.baryon_grids(N, nstore, seed=11)
    shifted = {k: v.copy() for k, v in g0.items()}
    ring = {k: v.copy() for k, v in g0.items()}
    head = np.zeros(1, np.int32)
    ctx = mxr.Context(N)
    if device:
        import torch
        ring = {k: torch.from_numpy(v).cuda() for k, v in ring.items()}
    for snap in range(6):
        gals = synth.galaxies(20000, box, seed=100 + snap)
        b1 = ctx.construct_baryon_grids(p, gals, shifted)
        b2 = ctx.construct_baryon_grids(p, gals, ring, history_head=head)
        assert b1 == b2
        assert head[0] == (-(snap + 1)) % nstore
        got = {k: (v.cpu().numpy() if device else v) for k, v in ring.items()}
        for k in mxr.BARYON_GRIDS:
            if "histories" in k:
                a = _real(got[k], N)
                b = _real(shifted[k], N)
                for s in range(nstore):
                    assert np.array_equal(a[(head[0] + s) % nstore], b[s]), (snap, k, s)
            else:
                assert np.array_equal(got[k], shifted[k]), (snap, k)


@pytest.mark.parametrize("device", [False, True])
def test_compute_ts_reads_a_ring(device):
    N, box, snap = 32, 40.0, 30
    p = params(N, box, snap)
    depth = max(2, mxr.ts_history_depth(p, 40))
    d = inputs(N, box, snap, depth, scale=1e-2)
    xe0, Tk0 = state(N, d["deltax"])
    sf0 = np.zeros(len(ZZ))
    sf0[:snap] = 1e-3 * np.arange(snap) / snap
    npad = N * N * 2 * (N // 2 + 1)
    h = 3 % depth

    def rotate(hist):
        r = np.empty_like(hist).reshape(depth, npad)
        src = hist.reshape(depth, npad)
        for s in range(depth):
            r[(h + s) % depth] = src[s]
        return r.reshape(-1)

    ctx = mxr.Context(N)
    outs = []
    for head, hs, hx in [(0, d["sfr_histories"], d["xray_luminosity_histories"]),
                         (h, rotate(d["sfr_histories"]), rotate(d["xray_luminosity_histories"]))]:
        xe, Tk = xe0.copy(), Tk0.copy()
        xe_box = np.zeros_like(xe)
        TS = np.zeros(N ** 3, np.float32)
        sf = sf0.copy()
        arrs = [d["deltax"], d["stars"], hs, hx, xe, xe_box, Tk, TS]
        if device:
            import torch
            arrs = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrs]
        g, _ = ctx.compute_ts(p, arrs[0], arrs[1], arrs[2], arrs[3], depth, arrs[4], arrs[5], arrs[6], arrs[7], sf,
                              want_tables=False, history_head=head)
        res = [a.cpu().numpy() if device else a for a in arrs[4:]]
        outs.append((g.as_dict(), res))
    (g0, r0), (g1, r1) = outs
    for a, b in zip(r0, r1):
        assert np.array_equal(a, b)
    g0.pop("device_ms")
    g1.pop("device_ms")
    assert g0 == g1
