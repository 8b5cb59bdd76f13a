"""The launch-shape autotuner on the device: it leaves the caller's grids alone, writes a
wisdom entry a new context loads back, and the tuned path keeps parity -- bit for bit where the
chosen shapes change no arithmetic (tile widths, live tiles), within the oracle bounds where
the z pass switches to direct summation."""
import numpy as np
import pytest

import mxr
import oracle_lib as O
from conftest import realisation
from parity import check, compare

pytestmark = pytest.mark.gpu


def _state(N):
    return mxr.HiiState.host(N)


def _run(ctx, p, inp, st):
    arrays = dict(inp)
    arrays.update(st.arrays())
    return ctx.find_hii_bubbles(p, **arrays)


def _fields(st):
    return {k: np.asarray(v).copy() for k, v in st.arrays().items()}


def test_autotune_host_grids_and_wisdom_roundtrip(tmp_path):
    N, box = 64, 67.8
    inp = realisation(N, box, 1)
    before = {k: v.copy() for k, v in inp.items()}
    p = mxr.hii_params(N, box, filter_type=1)
    wis = str(tmp_path / "wisdom.json")
    ctx = mxr.Context(N)
    default = ctx.tuning().as_dict()
    sd = _state(N)
    gd = _run(ctx, p, inp, sd)
    print("autotune N=64", flush=True)
    t = ctx.autotune(p, wisdom=wis, **inp)
    print("chosen", t.as_dict(), flush=True)
    for k, v in inp.items():
        np.testing.assert_array_equal(v, before[k])  # inputs untouched
    assert ctx.tuning().as_dict() == t.as_dict()
    assert t.zdirect_kmax == default["zdirect_kmax"] and t.zdirect_rpb == default["zdirect_rpb"]  # N % 256 != 0
    st = _state(N)
    g = _run(ctx, p, inp, st)
    a, b = _fields(st), _fields(sd)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)  # tile shapes change no arithmetic
    assert g.volume_weighted_global_xH == gd.volume_weighted_global_xH
    ctx2 = mxr.Context(N)
    assert ctx2.tuning().as_dict() == default
    assert ctx2.load_wisdom(p, wis) == mxr.MXR_OK
    assert ctx2.tuning().as_dict() == t.as_dict()
    p0 = mxr.hii_params(N, box, filter_type=0)
    assert ctx2.load_wisdom(p0, wis) == mxr.MXR_ERR_NO_WISDOM  # other filter type: no entry
    ctx.close()
    ctx2.close()


def test_autotune_256_device_grids_parity(tmp_path):
    import torch

    N, box = 256, 200.0
    inp = realisation(N, box, 1)
    dinp = {k: torch.from_numpy(v).cuda() for k, v in inp.items()}
    p = mxr.hii_params(N, box, filter_type=1)
    ctx = mxr.Context(N)
    print("autotune N=256", flush=True)
    t = ctx.autotune(p, wisdom=str(tmp_path / "w.json"), **dinp)
    print("chosen", t.as_dict(), flush=True)
    sg, sr = _state(N), _state(N)
    gg = _run(ctx, p, inp, sg)
    rg, margin = O.find_hii_bubbles(p, inp, sr, margin=True)
    st = compare(N, {k: np.asarray(v) for k, v in sg.arrays().items()},
                 {k: np.asarray(v) for k, v in sr.arrays().items()}, gg, rg, margin)
    print(st)
    check(st)
    ctx.close()


def test_batch_autotune_then_graph_matches_single_calls(tmp_path):
    import torch

    cases = [(64, 67.8, 1.0), (96, 100.0, 0.8), (64, 70.0, 1.4)]
    inputs = [{k: torch.from_numpy(v).cuda() for k, v in realisation(N, box).items()} for N, box, _ in cases]
    params = [mxr.hii_params(N, box, filter_type=1, zeta=z) for N, box, z in cases]
    wis = str(tmp_path / "w.json")

    def state(N):
        n = N ** 3
        return dict(xH=torch.ones(n, device="cuda"), z_at_ionization=torch.full((n,), -1.0, device="cuda"),
                    J_21_at_ionization=torch.zeros(n, device="cuda"), r_bubble=torch.zeros(n, device="cuda"),
                    temp_kinetic_all_gas=torch.zeros(n, device="cuda"))

    ref = []
    for (N, _, _), p, inp in zip(cases, params, inputs):
        ctx = mxr.Context(N)
        st = state(N)
        ctx.find_hii_bubbles(p, **inp, **st)
        torch.cuda.synchronize()
        ref.append({k: v.cpu().numpy() for k, v in st.items()})
        ctx.close()
    b = mxr.Batch([c[0] for c in cases])
    print("batch autotune", flush=True)
    b.autotune(params, inputs, wisdom=wis)
    b2 = mxr.Batch([c[0] for c in cases])
    assert b2.load_wisdom(params, wis) == mxr.MXR_OK
    for bb in (b, b2):
        for rep in range(2):  # second launch replays the instantiated graph
            states = [state(c[0]) for c in cases]
            bb.find_hii_bubbles(params, [dict(i, **s) for i, s in zip(inputs, states)], mxr.MXR_BATCH_GRAPH)
            torch.cuda.synchronize()
            for i, (s, r) in enumerate(zip(states, ref)):
                for k in r:
                    np.testing.assert_array_equal(s[k].cpu().numpy(), r[k], err_msg=f"{i} {k} rep {rep}")
    b.close()
    b2.close()
