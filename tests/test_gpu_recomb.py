"""Inhomogeneous recombinations in find_HII_bubbles (Flag_IncludeRecombinations) on the GPU.

  * The MHR00 tables the library builds (src/core/recombinations.c:208-374) against an
    independent quadrature: scipy's adaptive QUADPACK of the oracle's integrands (each pinned in
    test_oracle_recomb.py) at relative tolerance 1e-10, with the PDF normalisation A(z) computed
    the reference's way (1 / integral at z = 2 .. 61, then a natural spline in z) from the same
    quadrature.  The reference integrates with GSL QAG at relative tolerance 0.01 (A: 0.001);
    the bound here is 2e-3 relative.
  * A three-snapshot sequence (N_rec and Gamma12 carried between calls, z falling 10 -> 8.5)
    against the FP32 restatement, which interpolates the same tables with its own splines.
  * The fused, dense and plain paths bit-identical with recombinations, and the batched graph.

Reference: src/core/find_HII_bubbles.c:128-130, 216-245, 389-393, 409-414, 438-441, 467-506,
612-700.
"""
import math

import numpy as np
import pytest
from scipy import integrate
from scipy.interpolate import CubicSpline

import mxr
import oracle_lib as O
from conftest import realisation
from parity import check, compare

pytestmark = pytest.mark.gpu

N_Z, N_T, N_G = 200, 100, 150


@pytest.fixture(scope="module")
def ctxs():
    cache = {}

    def get(N, path=mxr.MXR_PATH_FUSED):
        if N not in cache:
            cache[N] = mxr.Context(N)
        cache[N].set_path(path)
        return cache[N]

    yield get
    for c in cache.values():
        c.close()


def _params(N, box, **kw):
    kw.setdefault("Flag_IncludeRecombinations", 1)
    kw.setdefault("Y_He", 0.24)
    return mxr.hii_params(N, box, **kw)


@pytest.fixture(scope="module")
def tables(ctxs):
    p = _params(32, 40.0)
    print("building recombination tables", flush=True)
    t = ctxs(32).recombination_tables(p)
    O.rec_set_tables(*t)
    return p, t


def _quad(f, a, b, epsabs=0.0):
    return integrate.quad(f, a, b, epsabs=epsabs, epsrel=1e-10, limit=400)[0]


def _a_spline(p):
    """A(z) the reference's way: 1 / integral of the unnormalised PDF at z = 2 .. 61 (integrated
    here over ln Delta in pieces), natural cubic spline in z, clamped to [2, 62]."""
    zs = 2.0 + np.arange(60, dtype=np.float32).astype(np.float64)
    inv = []
    for z in zs:
        f = lambda l, z=z: O.rec_probe(p, 3, math.exp(l), z) * math.exp(l)
        edges = [math.log(1e-25), -5.0, -1.0, 0.0, 1.0, 3.0, 6.0, math.log(1e25)]
        # the pieces far out in the tails are ~0 (the integral itself is O(1)): absolute floor
        inv.append(1.0 / sum(_quad(f, a, b, 1e-15) for a, b in zip(edges[:-1], edges[1:])))
    sp = CubicSpline(zs, inv, bc_type="natural")
    return lambda z: float(sp(min(max(z, 2.0), 62.0)))


def test_tables_against_quadrature(tables):
    p, (x, ln_rate, clump, ln_res) = tables
    g = np.arange(N_G, dtype=np.float32)
    np.testing.assert_array_equal(x, (np.float32(-10.0) + g * np.float32(0.1)).astype(np.float64))
    assert np.all(np.isfinite(ln_rate)) and np.all(np.isfinite(clump)) and np.all(np.isfinite(ln_res))
    A = _a_spline(p)
    rng = np.random.default_rng(5)
    picks = [(0, 0, 0), (199, 99, 149), (40, 67, 100), (10, 0, 149), (150, 99, 0), (25, 33, 50)]
    picks += [tuple(int(v) for v in (rng.integers(N_Z), rng.integers(N_T), rng.integers(N_G))) for _ in range(30)]
    lo, hi = math.log(0.01), math.log(200.0)
    worst = [0.0, 0.0, 0.0]
    for zc, tc, gc in picks:
        z = float(np.float32(zc) * np.float32(0.2))
        T4 = 10.0 ** (float(np.float32(tc) * np.float32(0.03) + np.float32(2.0)) - 4.0)
        g12 = math.exp(x[gc])
        a = A(z)
        ints = [_quad(lambda l, k=k: O.rec_probe(p, 4, l, z, g12, T4, a)[k], lo, hi) for k in range(4)]
        want = (ints[0], ints[1] / ints[2] / ints[2], ints[3])
        got = (math.exp(ln_rate[zc, tc, gc]), clump[zc, tc, gc], math.exp(ln_res[zc, tc, gc]))
        for k in range(3):
            rel = abs(got[k] / want[k] - 1)
            worst[k] = max(worst[k], rel)
            assert rel < 2e-3, (zc, tc, gc, k, got[k], want[k])
    print(f"table rel. error vs quadrature: rate {worst[0]:.2e} clumping {worst[1]:.2e} "
          f"residual {worst[2]:.2e}", flush=True)


def _state(N):
    return mxr.HiiState.host(N, recombinations=True)


def _inputs(N, box, ftype=1):
    """The seeded realisation with weighted_sfr scaled so that the first-crossing Gamma_12 spans
    the tables' ln Gamma_12 range (-10 .. 4.9) instead of sitting above it."""
    inp = realisation(N, box, ftype)
    inp["weighted_sfr"] = (inp["weighted_sfr"] * np.float32(3e-4)).astype(np.float32)
    return inp


def _run_gpu(ctx, params, inputs, state):
    arrays = dict(inputs)
    arrays.update(state.arrays())
    return ctx.find_hii_bubbles(params, **arrays)


def _fields(st):
    return {k: np.asarray(v).copy() for k, v in st.arrays().items()}


SNAPSHOTS = ((9.5, 10.0), (9.0, 9.5), (8.5, 9.0))


def _flips(a, b):
    """Cells where two results disagree on the ionisation flag or the first-crossing radius."""
    return ((a["xH"] == 0) != (b["xH"] == 0)) | (a["r_bubble"] != b["r_bubble"])


def _rec_stats(N, g, r, margin, history):
    """Recombination fields where the two sides agree on the ionisation radius, now and in every
    earlier snapshot (`history`: cells that ever flipped, whose N_rec carries the difference)."""
    H = N // 2 + 1
    ok = (g["xH"] == 0) == (r["xH"] == 0)
    rb_diff = ok & (g["r_bubble"] != r["r_bubble"])
    same = ok & ~rb_diff & ~history
    out = {"r_bubble_flips": int(rb_diff.sum()), "history_flips": int(history.sum())}
    real = lambda a: a.reshape(N, N, 2 * H)[:, :, :N].reshape(-1)
    nrec_g, nrec_r = real(g["N_rec"]), real(r["N_rec"])
    for k, (a, b) in dict(Gamma12=(g["Gamma12"], r["Gamma12"]), N_rec=(nrec_g, nrec_r),
                          residual_xH=(g["residual_xH"], r["residual_xH"]),
                          clumping_factor=(g["clumping_factor"], r["clumping_factor"]),
                          t_resp=(g["t_resp"], r["t_resp"])).items():
        a, b = a[same], b[same]
        rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-30 if k != "N_rec" else 1e-12)
        out[f"p999_rel_{k}"] = float(np.quantile(rel, 0.999)) if rel.size else 0.0
        out[f"max_rel_{k}"] = float(rel.max()) if rel.size else 0.0
    return out


REC_FIELDS = ("Gamma12", "N_rec", "residual_xH", "clumping_factor", "t_resp")


def _real_view(N, k, a):
    H = N // 2 + 1
    return a.reshape(N, N, 2 * H)[:, :, :N].reshape(-1) if k == "N_rec" else a


def _fp64_anchor(N, g, r, r64, history):
    """Per-field |GPU - FP64| and |FP32 restatement - FP64| (max, 99.9th percentile) over the
    cells where all three have agreed on the ionisation flag and radius in every snapshot."""
    agree = ~history
    out = {}
    for k in REC_FIELDS:
        a, b, c = (_real_view(N, k, d[k])[agree].astype(np.float64) for d in (g, r, r64))
        fin = np.abs(c) < 1e29  # t_resp is 1e30 where nothing recombines or ionises
        ea, eb = np.abs(a - c)[fin], np.abs(b - c)[fin]
        out[k] = dict(max_gpu=float(ea.max()), max_f32=float(eb.max()), p999_gpu=float(np.quantile(ea, 0.999)),
                      p999_f32=float(np.quantile(eb, 0.999)), scale=float(np.abs(c[fin]).max()))
    return out


@pytest.mark.parametrize("N,box,ftype,uvb", [(64, 67.8, 1, 1), (64, 67.8, 0, 2), (128, 125.0, 1, 1)])
def test_sequence_vs_oracle(ctxs, tables, N, box, ftype, uvb):
    """Three snapshots with N_rec and Gamma12 carried over: every call's state, recombination
    grids and globals against the FP32 restatement on the same tables, and each recombination
    field's worst cell against the FP64 oracle."""
    inp = _inputs(N, box, ftype)
    sg, sr, s64 = _state(N), _state(N), _state(N)
    hist_r = np.zeros(N ** 3, bool)    # GPU vs FP32 restatement
    hist_64 = np.zeros(N ** 3, bool)   # any of the three vs another
    for z, zp in SNAPSHOTS:
        print(f"recombinations N={N} filter={ftype} UVB={uvb} z={z}", flush=True)
        p = _params(N, box, filter_type=ftype, redshift=z, prev_redshift=zp, ReionUVBFlag=uvb)
        gg = _run_gpu(ctxs(N), p, inp, sg)
        rg, margin = O.find_hii_bubbles(p, inp, sr, margin=True)
        O.find_hii_bubbles(p, inp, s64, precision="f64")
        g, r, r64 = _fields(sg), _fields(sr), _fields(s64)
        st = compare(N, g, r, gg, rg, margin, exclude=hist_r)
        hist_r |= _flips(g, r)
        hist_64 |= hist_r | _flips(g, r64) | _flips(r, r64)
        rs = _rec_stats(N, g, r, margin, hist_r)
        anchor = _fp64_anchor(N, g, r, r64, hist_64)
        gam = g["Gamma12"][g["Gamma12"] > 0] * p.Hubble_h ** 2
        print(st, rs, "ln Gamma_12 h^2 percentiles 1/50/99:", np.log(np.percentile(gam, [1, 50, 99])), flush=True)
        print("max |error| vs FP64 (GPU, FP32 restatement, max |FP64|):", anchor, flush=True)
        tol = 1e-4 if ftype == 1 else 5e-4
        check(st, p999_tol=tol)
        # a cell whose largest ionising radius sits at the threshold may cross one radius
        # earlier or later on the other side (like the J_21 radius flips)
        assert rs["r_bubble_flips"] <= max(2, 5e-4 * N ** 3), rs
        assert rs["history_flips"] <= max(4, 1e-3 * N ** 3), rs
        # the recombination fields inherit the filtered fields' FP32 rounding: the same 99.9th
        # percentile bound as the partial xH against the FP32 restatement (N_rec grows with
        # 1 - xH, whose relative error is largest in barely ionised cells); against the FP64
        # oracle the 99.9th-percentile error within twice the FP32 restatement's own, and the
        # worst cell within twice the restatement's worst or 1e-4 of the field's largest value
        for k in REC_FIELDS:
            assert rs[f"p999_rel_{k}"] <= tol, (k, rs)
            a = anchor[k]
            assert a["p999_gpu"] <= 2 * a["p999_f32"] + 1e-7 * a["scale"], (k, a)
            assert a["max_gpu"] <= max(2 * a["max_f32"], 1e-4 * a["scale"]), (k, a)
        n = N ** 3
        slack = (st["mismatch"] + rs["r_bubble_flips"]) / n
        for k in ("N_rec", "Gamma12", "residual_xH", "clumping_factor", "r_bubble"):
            for w in ("volume", "mass"):
                name = f"{w}_weighted_global_{k}"
                a, b = getattr(gg, name), getattr(rg, name)
                assert abs(a - b) <= (1e-4 + 20 * slack) * abs(b) + 1e-30, (name, a, b)
        assert rg.volume_weighted_global_r_bubble > 0 and rg.volume_weighted_global_Gamma12 > 0
    assert rg.volume_weighted_global_N_rec > 0  # recombinations accumulated over the steps


def test_paths_bit_identical(ctxs, tables):
    N, box = 64, 67.8
    inp = _inputs(N, box)
    outs = {}
    for path in (mxr.MXR_PATH_FUSED, mxr.MXR_PATH_FUSED_DENSE, mxr.MXR_PATH_PLAIN):
        st = _state(N)
        for z, zp in SNAPSHOTS[:2]:
            g = _run_gpu(ctxs(N, path), _params(N, box, redshift=z, prev_redshift=zp, ReionUVBFlag=2), inp, st)
        outs[path] = (_fields(st), g)
    ref, rg = outs[mxr.MXR_PATH_FUSED]
    for path in (mxr.MXR_PATH_FUSED_DENSE, mxr.MXR_PATH_PLAIN):
        o, g = outs[path]
        for k in ref:
            assert int((o[k] != ref[k]).sum()) == 0, f"path {path} field {k}"
        assert g.volume_weighted_global_N_rec == rg.volume_weighted_global_N_rec
    ctxs(N, mxr.MXR_PATH_FUSED)


def test_spin_bh_recombinations_vs_oracle(ctxs, tables):
    """All optional grids at once: x_e, the BH grids and N_rec (seven filtered fields)."""
    from test_gpu_hii import _with_bh

    N, box = 64, 67.8
    inp = _with_bh(_inputs(N, box), spin_N=N)
    sg, sr = _state(N), _state(N)
    hist = np.zeros(N ** 3, bool)
    for z, zp in SNAPSHOTS[:2]:
        p = _params(N, box, redshift=z, prev_redshift=zp, spin_temp=True, bh=True, ReionUVBFlag=2,
                    Flag_TemperatureDependentRec=1)
        gg = _run_gpu(ctxs(N), p, inp, sg)
        rg, margin = O.find_hii_bubbles(p, inp, sr, margin=True)
        g, r = _fields(sg), _fields(sr)
        st = compare(N, g, r, gg, rg, margin, exclude=hist)
        hist |= _flips(g, r)
        rs = _rec_stats(N, g, r, margin, hist)
        print(st, rs, flush=True)
        check(st)
        for k in ("Gamma12", "N_rec", "residual_xH", "clumping_factor", "t_resp"):
            assert rs[f"p999_rel_{k}"] <= 1e-4, (k, rs)


def test_batched_graph_matches_single(ctxs, tables):
    """Two realisations with recombinations in one captured graph (tables built before the
    capture), twice: bit-identical to single calls."""
    import torch

    dims = [64, 32]
    boxes = [67.8, 40.0]
    ps = [_params(N, b, redshift=9.0, prev_redshift=9.5) for N, b in zip(dims, boxes)]
    inps = [_inputs(N, b) for N, b in zip(dims, boxes)]
    singles = []
    for N, p, inp in zip(dims, ps, inps):
        st = _state(N)
        st.N_rec[:] = 0.05  # a carried-over recombination grid
        _run_gpu(ctxs(N), p, inp, st)
        singles.append(_fields(st))
    b = mxr.Batch(dims)
    try:
        for launch in range(2):
            arrays = []
            for N, inp in zip(dims, inps):
                st = _state(N)
                st.N_rec[:] = 0.05
                arrays.append(dict({k: torch.from_numpy(v).cuda() for k, v in inp.items()},
                                   **{k: torch.from_numpy(np.asarray(v).copy()).cuda() for k, v in st.arrays().items()}))
            b.find_hii_bubbles(ps, arrays, mxr.MXR_BATCH_GRAPH)
            torch.cuda.synchronize()
            for s, a in zip(singles, arrays):
                for k in s:
                    assert np.array_equal(a[k].cpu().numpy(), s[k]), (launch, k)
    finally:
        b.close()
