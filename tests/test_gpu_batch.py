"""Batched find_HII_bubbles (mxr_find_hii_bubbles_batched): realisations of different grid
dimensions, boxes and ReionEfficiency, run as one captured hipGraph, on per-realisation streams
and one at a time, must give exactly the fields and globals of separate single calls."""
import numpy as np
import pytest

import mxr
from conftest import realisation

pytestmark = pytest.mark.gpu

CASES = [(32, 40.0, 1.0), (64, 67.8, 1.3), (96, 100.0, 0.7), (32, 36.0, 1.9), (64, 70.0, 0.6), (16, 20.0, 1.0)]


def _inputs(torch):
    out = []
    for N, box, zeta in CASES:
        inp = {k: torch.from_numpy(v).cuda() for k, v in realisation(N, box).items()}
        p = mxr.hii_params(N, box, filter_type=1, zeta=zeta)
        out.append((N, p, inp))
    return out


def _state(torch, N):
    n = N ** 3
    return dict(xH=torch.ones(n, device="cuda"), z_at_ionization=torch.full((n,), -1.0, device="cuda"),
                J_21_at_ionization=torch.zeros(n, device="cuda"), r_bubble=torch.zeros(n, device="cuda"),
                temp_kinetic_all_gas=torch.zeros(n, device="cuda"))


def _fields(st):
    return {k: v.cpu().numpy() for k, v in st.items()}


def test_batch_modes_match_single_calls():
    import torch

    cases = _inputs(torch)
    ref = []
    for N, p, inp in cases:
        ctx = mxr.Context(N)
        st = _state(torch, N)
        g = ctx.find_hii_bubbles(p, **inp, **st)
        torch.cuda.synchronize()
        ref.append((_fields(st), g))
        ctx.close()
    b = mxr.Batch([c[0] for c in cases])
    for mode in (mxr.MXR_BATCH_SERIAL, mxr.MXR_BATCH_STREAMS, mxr.MXR_BATCH_GRAPH, mxr.MXR_BATCH_GRAPH):
        print(f"batch mode {mode}", flush=True)
        states = [_state(torch, c[0]) for c in cases]
        arrays = [dict(c[2], **s) for c, s in zip(cases, states)]
        gl = b.find_hii_bubbles([c[1] for c in cases], arrays, mode)
        torch.cuda.synchronize()
        for i, (st, (rf, rg)) in enumerate(zip(states, ref)):
            f = _fields(st)
            for k in rf:
                np.testing.assert_array_equal(f[k], rf[k], err_msg=f"mode {mode} realisation {i} {k}")
            assert gl[i].volume_weighted_global_xH == rg.volume_weighted_global_xH
            assert gl[i].mass_weighted_global_temp_kinetic_all_gas == rg.mass_weighted_global_temp_kinetic_all_gas
            assert gl[i].n_radii == rg.n_radii
        print(f"batch mode {mode}: {gl[0].device_ms:.2f} ms", flush=True)
    # a changed parameter forces a new capture; the graph's results follow it
    params = [c[1] for c in cases]
    params[2] = mxr.hii_params(96, 100.0, filter_type=1, zeta=1.5)
    states = [_state(torch, c[0]) for c in cases]
    arrays = [dict(c[2], **s) for c, s in zip(cases, states)]
    gl = b.find_hii_bubbles(params, arrays, mxr.MXR_BATCH_GRAPH)
    ctx = mxr.Context(96)
    st = _state(torch, 96)
    g = ctx.find_hii_bubbles(params[2], **cases[2][2], **st)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(_fields(states[2])["xH"], _fields(st)["xH"])
    assert gl[2].volume_weighted_global_xH == g.volume_weighted_global_xH
    ctx.close()
    b.close()


def test_batch_rejects_host_memory():
    N, box = 16, 20.0
    inp = realisation(N, box)
    st = mxr.HiiState.host(N)
    b = mxr.Batch([N])
    with pytest.raises(mxr.MxrError) as e:
        b.find_hii_bubbles([mxr.hii_params(N, box)], [dict(inp, **st.arrays())])
    assert e.value.code == mxr.MXR_ERR_INVALID_ARG
    b.close()
