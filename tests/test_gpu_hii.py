"""GPU parity of find_HII_bubbles (libmxr.so through its C ABI) against the oracle.

All tests here need an MI355X and run on it; inputs are the seeded synthetic grids of
synth.py, fed identically to the HIP path and to the CPU restatement.
"""
import numpy as np
import pytest

import mxr
import oracle_lib as O
from conftest import realisation
from parity import check, compare

pytestmark = pytest.mark.gpu


def _state(N):
    st = mxr.HiiState.host(N)
    return st


def _run_gpu(ctx, params, inputs, state):
    arrays = dict(inputs)
    arrays.update(state.arrays())
    return ctx.find_hii_bubbles(params, **arrays)


def _as_dict(st):
    return {k: np.asarray(v) for k, v in st.arrays().items()}


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


@pytest.mark.parametrize("N,box,ftype", [(16, 20.0, 1), (32, 40.0, 1), (64, 67.8, 0), (64, 67.8, 1), (64, 67.8, 2),
                                         (96, 100.0, 1), (128, 125.0, 1), (160, 150.0, 1)])
def test_parity_vs_oracle(ctxs, N, box, ftype):
    print(f"parity N={N} box={box} filter={ftype}", flush=True)
    inp = realisation(N, box, ftype)
    p = mxr.hii_params(N, box, filter_type=ftype)
    sg, sr = _state(N), _state(N)
    gg = _run_gpu(ctxs(N), p, inp, sg)
    rg, margin = O.find_hii_bubbles(p, inp, sr, margin=True)
    st = compare(N, _as_dict(sg), _as_dict(sr), gg, rg, margin)
    print(st)
    check(st, p999_tol=1e-4 if ftype == 1 else 5e-4)
    assert gg.n_radii == rg.n_radii
    assert abs(gg.mass_weighted_global_xH - rg.mass_weighted_global_xH) <= 1e-5 + st["mismatch"] * 4 / N ** 3
    assert abs(gg.volume_weighted_global_weighted_sfr / rg.volume_weighted_global_weighted_sfr - 1) < 1e-12
    assert abs(gg.volume_weighted_global_temp_kinetic_all_gas / rg.volume_weighted_global_temp_kinetic_all_gas
               - 1) < 1e-4 + st["mismatch"] / N ** 3 * 10


@pytest.mark.parametrize("zdirect_kmax", [64, 0])
def test_parity_256(ctxs, zdirect_kmax):
    """256^3 sharp-k against the oracle, with the z pass by direct summation on the radii that
    keep few k_z (64) and by the row FFT on every radius (0)."""
    N, box = 256, 200.0
    print(f"parity N=256 zdirect_kmax={zdirect_kmax}", flush=True)
    inp = realisation(N, box, 1)
    p = mxr.hii_params(N, box, filter_type=1)
    ctx = ctxs(N)
    old = ctx.set_tuning(zdirect_kmax=zdirect_kmax)
    try:
        sg, sr = _state(N), _state(N)
        gg = _run_gpu(ctx, p, inp, sg)
    finally:
        ctx.restore_tuning(old)
    rg, margin = O.find_hii_bubbles(p, inp, sr, margin=True)
    st = compare(N, _as_dict(sg), _as_dict(sr), gg, rg, margin)
    print(st)
    check(st)


def test_direct_z_pass_error_vs_fp64_256(ctxs):
    """The direct-summation z pass is no less accurate than a row FFT: at 256^3 its per-cell
    partial-xH error against the FP64 oracle stays within twice the FP32 restatement's, and so
    does the GPU's row-FFT form."""
    N, box = 256, 200.0
    print("direct z pass vs FP64 N=256", flush=True)
    inp = realisation(N, box, 1)
    p = mxr.hii_params(N, box, filter_type=1)
    ctx = ctxs(N)
    res = {}
    for kmax in (64, 0):
        old = ctx.set_tuning(zdirect_kmax=kmax)
        try:
            st = _state(N)
            res[kmax] = (_run_gpu(ctx, p, inp, st), np.asarray(st.xH).copy())
        finally:
            ctx.restore_tuning(old)
    s32, s64 = _state(N), _state(N)
    g32 = O.find_hii_bubbles(p, inp, s32)
    g64, margin = O.find_hii_bubbles(p, inp, s64, precision="f64", margin=True)
    m32 = (s32.xH == 0) != (s64.xH == 0)
    for kmax, (gg, xh) in res.items():
        both = (xh > 0) & (s32.xH > 0) & (s64.xH > 0)
        e_gpu = np.abs(xh[both] - s64.xH[both])
        e_32 = np.abs(s32.xH[both] - s64.xH[both])
        mg = (xh == 0) != (s64.xH == 0)
        print(f"zdirect_kmax={kmax}: xH error vs FP64 max {e_gpu.max():.3e} p99.9 {np.quantile(e_gpu, 0.999):.3e} "
              f"flags {int(mg.sum())} | f32 restatement max {e_32.max():.3e} p99.9 {np.quantile(e_32, 0.999):.3e} "
              f"flags {int(m32.sum())}", flush=True)
        assert e_gpu.max() <= 2 * e_32.max() + 1e-6
        assert np.quantile(e_gpu, 0.999) <= 2 * np.quantile(e_32, 0.999) + 1e-6
        assert mg.sum() <= max(2, 2 * m32.sum())
        assert np.all(np.abs(margin[mg]) < 1e-3)
        assert abs(gg.volume_weighted_global_xH - g64.volume_weighted_global_xH) <= \
            2 * abs(g32.volume_weighted_global_xH - g64.volume_weighted_global_xH) + 1e-6


def test_spin_temp_grid(ctxs):
    """Flag_IncludeSpinTemp: x_e_box is a 4th filtered grid and Tk_box sets the temperatures."""
    N, box = 64, 67.8
    inp = realisation(N, box, 1)
    r = np.random.default_rng(3)
    H = N // 2 + 1
    xe = np.zeros((N, N, 2 * H), np.float32)
    xe[:, :, :N] = r.uniform(0.0, 0.3, (N, N, N)).astype(np.float32)
    inp["x_e_box"] = xe.reshape(-1)
    inp["Tk_box"] = r.uniform(5.0, 50.0, N ** 3).astype(np.float32)
    p = mxr.hii_params(N, box, filter_type=1, spin_temp=True)
    sg, sr = _state(N), _state(N)
    gg = _run_gpu(ctxs(N), p, inp, sg)
    rg, margin = O.find_hii_bubbles(p, inp, sr, margin=True)
    st = compare(N, _as_dict(sg), _as_dict(sr), gg, rg, margin)
    print(st)
    check(st)


@pytest.mark.parametrize("ftype", [0, 1, 2])
def test_paths_bit_identical(ctxs, ftype):
    """PLAIN, FUSED and FUSED_DENSE compute the same filtered fields: identical state grids."""
    N, box = 64, 67.8
    inp = realisation(N, box, ftype)
    p = mxr.hii_params(N, box, filter_type=ftype)
    outs = {}
    for path in (mxr.MXR_PATH_FUSED, mxr.MXR_PATH_FUSED_DENSE, mxr.MXR_PATH_PLAIN):
        st = _state(N)
        g = _run_gpu(ctxs(N, path), p, inp, st)
        outs[path] = (_as_dict(st), g)
    ref, rg = outs[mxr.MXR_PATH_FUSED]
    for path in (mxr.MXR_PATH_FUSED_DENSE, mxr.MXR_PATH_PLAIN):
        o, g = outs[path]
        for k in ref:
            nd = int((o[k] != ref[k]).sum())
            assert nd == 0, f"path {path} field {k}: {nd} cells differ"
        assert abs(g.volume_weighted_global_xH - rg.volume_weighted_global_xH) < 1e-12
    ctxs(N, mxr.MXR_PATH_FUSED)


def test_pruned_equals_dense_512(ctxs):
    """At the benchmark size the mode-skipping path with the row-FFT z pass equals the dense
    fused path bit for bit; with the direct-summation z pass it differs only by rounding: flag
    mismatch within the parity bound, global xH within 1e-5 plus the flipped cells' share."""
    import synth
    import torch

    N, box = 512, 400.0
    print("512 pruned vs dense", flush=True)
    inp = {k: v.contiguous() for k, v in synth.realisation(N, box, 42, device="cuda", stars_to_matter=0.05).items()}
    p = mxr.hii_params(N, box, filter_type=1)
    n = N ** 3
    res = []
    for path, kmax in ((mxr.MXR_PATH_FUSED, 0), (mxr.MXR_PATH_FUSED_DENSE, 0), (mxr.MXR_PATH_FUSED, 64)):
        st = mxr.HiiState(torch.ones(n, device="cuda"), torch.full((n,), -1.0, device="cuda"),
                          torch.zeros(n, device="cuda"), torch.zeros(n, device="cuda"), torch.zeros(n, device="cuda"))
        ctx = ctxs(N, path)
        old = ctx.set_tuning(zdirect_kmax=kmax)
        try:
            g = _run_gpu(ctx, p, inp, st)
        finally:
            ctx.restore_tuning(old)
        torch.cuda.synchronize()
        res.append((st, g))
    (a, ga), (b, gb), (d, gd) = res
    for k in ("xH", "z_at_ionization", "J_21_at_ionization", "temp_kinetic_all_gas"):
        assert torch.equal(getattr(a, k), getattr(b, k)), k
    for st, g in ((a, ga), (d, gd)):
        xh = st.xH
        assert float(xh.min()) >= 0.0 and float(xh.max()) <= 1.0
        assert abs(g.volume_weighted_global_xH - float(xh.double().mean())) < 1e-9
    flips = int(((a.xH == 0) != (d.xH == 0)).sum())
    agree = (a.xH == 0) == (d.xH == 0)
    z_diff = int((a.z_at_ionization[agree] != d.z_at_ionization[agree]).sum())
    dxh = abs(ga.volume_weighted_global_xH - gd.volume_weighted_global_xH)
    print(f"512 xH {ga.volume_weighted_global_xH} direct {gd.volume_weighted_global_xH} flips {flips} "
          f"({flips / n:.2e}) z_in diffs {z_diff} ms fft {ga.device_ms:.2f} dense {gb.device_ms:.2f} "
          f"direct {gd.device_ms:.2f}", flush=True)
    assert flips <= 5e-4 * n
    assert dxh <= 1e-5 + flips / n
    ctxs(N, mxr.MXR_PATH_FUSED)


def test_device_pointers_match_host(ctxs):
    import torch

    N, box = 64, 67.8
    inp = realisation(N, box, 1)
    p = mxr.hii_params(N, box, filter_type=1)
    sh = _state(N)
    gh = _run_gpu(ctxs(N), p, inp, sh)
    dinp = {k: torch.from_numpy(v).cuda() for k, v in inp.items()}
    n = N ** 3
    sd = mxr.HiiState(torch.ones(n, device="cuda"), torch.full((n,), -1.0, device="cuda"),
                      torch.zeros(n, device="cuda"), torch.zeros(n, device="cuda"), torch.zeros(n, device="cuda"))
    gd = _run_gpu(ctxs(N), p, dinp, sd)
    for k, v in sh.arrays().items():
        np.testing.assert_array_equal(getattr(sd, k).cpu().numpy(), v, err_msg=k)
    assert gd.volume_weighted_global_xH == gh.volume_weighted_global_xH


def test_two_snapshots_z_in_persists(ctxs):
    """z_at_ionization set at the first crossing and carried to the next call (z decreasing)."""
    N, box = 64, 67.8
    inp = realisation(N, box, 1)
    sg, sr = _state(N), _state(N)
    for z, zeta in ((9.0, 0.5), (8.0, 1.0)):
        p = mxr.hii_params(N, box, filter_type=1, redshift=z, zeta=zeta)
        gg = _run_gpu(ctxs(N), p, inp, sg)
        rg, margin = O.find_hii_bubbles(p, inp, sr, margin=True)
        st = compare(N, _as_dict(sg), _as_dict(sr), gg, rg, margin)
        check(st)
    z = np.asarray(sg.z_at_ionization)
    assert set(np.unique(z)) <= {-1.0, 9.0, 8.0}
    assert (z == 9.0).sum() > 0 and (z == 8.0).sum() > 0


def test_monotone_in_efficiency(ctxs):
    N, box = 64, 67.8
    inp = realisation(N, box, 1)
    prev = 1.1
    for zeta in (0.5, 1.0, 2.0):
        st = _state(N)
        g = _run_gpu(ctxs(N), mxr.hii_params(N, box, zeta=zeta), inp, st)
        assert g.volume_weighted_global_xH <= prev
        prev = g.volume_weighted_global_xH
        ion_now = np.asarray(st.xH) == 0
        if zeta > 0.5:
            assert np.all(ion_now[ion_prev])
        ion_prev = ion_now


def test_error_codes(ctxs):
    N = 32
    inp = realisation(N, 40.0, 1)
    st = _state(N)
    for over, code in ((dict(ReionFilterType=3), mxr.MXR_ERR_FILTER),
                       (dict(Flag_IncludeRecombinations=1), mxr.MXR_ERR_INVALID_ARG),  # no N_rec etc.
                       (dict(Flag_IncludeRecombinations=1, Y_He=0.0), mxr.MXR_ERR_INVALID_ARG),
                       (dict(ReionGridDim=64), mxr.MXR_ERR_INVALID_ARG)):
        p = mxr.hii_params(N, 40.0, **over)
        with pytest.raises(mxr.MxrError) as e:
            _run_gpu(ctxs(N), p, inp, st)
        assert e.value.code == code


@pytest.mark.parametrize("ftype", [0, 1, 2])
def test_filter_entry_point(ctxs, ftype):
    """mxr_filter against the oracle's filter on a random half spectrum."""
    N, box, R = 64, 67.8, 7.3
    r = np.random.default_rng(ftype)
    H = N // 2 + 1
    spec = (r.standard_normal(N * N * H) + 1j * r.standard_normal(N * N * H)).astype(np.complex64)
    ref = O.filter_box(spec, N, box, R, ftype)
    got = spec.copy()
    ctxs(N).filter(got, 0, N, N, box, R, ftype)
    if ftype == 1:
        np.testing.assert_array_equal(got, ref)
    else:
        np.testing.assert_allclose(got, ref, rtol=2e-6, atol=2e-7)


def test_sharp_k_edge_matches_oracle(ctxs):
    """A radius on the sharp-k edge of mode (0, 1, 10) at 64^3, 67.8 cMpc/h: the separately rounded
    |k| (the reference's and the oracle's) keeps it and an FMA-contracted one drops it.  The
    filter entry point matches the oracle bit for bit (its sharp-k test is the one the fused
    passes' window and kept extents evaluate)."""
    from test_oracle_filter import EDGE_BOX, EDGE_MODE, EDGE_N, EDGE_R

    N, H = EDGE_N, EDGE_N // 2 + 1
    r = np.random.default_rng(21)
    spec = (r.standard_normal(N * N * H) + 1j * r.standard_normal(N * N * H)).astype(np.complex64)
    ref = O.filter_box(spec, N, EDGE_BOX, EDGE_R, 1)
    i, j, k = EDGE_MODE
    assert ref.reshape(N, N, H)[i, j, k] != 0
    got = spec.copy()
    ctxs(N).filter(got, 0, N, N, EDGE_BOX, EDGE_R, 1)
    np.testing.assert_array_equal(got, ref)


def test_brightness_temperature(ctxs):
    N = 64
    r = np.random.default_rng(5)
    H = N // 2 + 1
    xH = r.uniform(0, 1, N ** 3).astype(np.float32)
    dx = np.zeros((N, N, 2 * H), np.float32)
    dx[:, :, :N] = r.lognormal(0, 0.5, (N, N, N)) - 1
    dx = dx.reshape(-1)
    TS = r.uniform(3, 300, N ** 3).astype(np.float32)
    for spin in (False, True):
        p = mxr.tb_params(N, 8.0, spin)
        ref, ave_r = O.brightness_temperature(p, xH, dx, TS if spin else None)
        got = np.zeros(N ** 3, np.float32)
        ave = ctxs(N).brightness_temperature(p, xH, dx, got, TS if spin else None)
        np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-6)
        assert abs(ave / ave_r - 1) < 1e-9


@pytest.mark.parametrize("ftype", [0, 1])
def test_error_vs_fp64_oracle(ctxs, ftype):
    """Anchor on the FP64 oracle: the GPU's per-cell error is no larger than the FP32
    restatement's own error (both are FP32 FFTs of the same fields)."""
    N, box = 64, 67.8
    inp = realisation(N, box, ftype)
    p = mxr.hii_params(N, box, filter_type=ftype)
    sg, s32, s64 = _state(N), _state(N), _state(N)
    gg = _run_gpu(ctxs(N), p, inp, sg)
    g32 = O.find_hii_bubbles(p, inp, s32)
    g64, margin = O.find_hii_bubbles(p, inp, s64, precision="f64", margin=True)
    both = (np.asarray(sg.xH) > 0) & (s32.xH > 0) & (s64.xH > 0)
    e_gpu = np.abs(np.asarray(sg.xH)[both] - s64.xH[both])
    e_32 = np.abs(s32.xH[both] - s64.xH[both])
    print("xH error vs FP64: gpu max", e_gpu.max(), "p99.9", np.quantile(e_gpu, 0.999), "| f32 restatement max",
          e_32.max(), "p99.9", np.quantile(e_32, 0.999))
    assert e_gpu.max() <= 2 * e_32.max() + 1e-6
    assert np.quantile(e_gpu, 0.999) <= 2 * np.quantile(e_32, 0.999) + 1e-6
    mg = (np.asarray(sg.xH) == 0) != (s64.xH == 0)
    m32 = (s32.xH == 0) != (s64.xH == 0)
    assert mg.sum() <= max(2, 2 * m32.sum())
    assert np.all(np.abs(margin[mg]) < 1e-3)
    assert abs(gg.volume_weighted_global_xH - g64.volume_weighted_global_xH) <= \
        2 * abs(g32.volume_weighted_global_xH - g64.volume_weighted_global_xH) + 1e-6


# ---- BH feedback: effective_bhm / effective_bhar as 5th and 6th filtered grids ----

def _with_bh(inp, seed=5, spin_N=None):
    import synth
    import torch

    out = dict(inp)
    out.update({k: v.numpy() for k, v in synth.bh_grids(torch.from_numpy(inp["stars"]), seed).items()})
    if spin_N:
        N = spin_N
        r = np.random.default_rng(3)
        H = N // 2 + 1
        xe = np.zeros((N, N, 2 * H), np.float32)
        xe[:, :, :N] = r.uniform(0.0, 0.3, (N, N, N)).astype(np.float32)
        out["x_e_box"] = xe.reshape(-1)
        out["Tk_box"] = r.uniform(5.0, 50.0, N ** 3).astype(np.float32)
    return out


@pytest.mark.parametrize("ftype,spin", [(1, False), (0, False), (1, True)])
def test_bh_parity_vs_oracle(ctxs, ftype, spin):
    N, box = 64, 67.8
    print(f"BH parity filter={ftype} spin={spin}", flush=True)
    inp = _with_bh(realisation(N, box, ftype), spin_N=N if spin else None)
    p = mxr.hii_params(N, box, filter_type=ftype, spin_temp=spin, bh=True)
    sg, sr = _state(N), _state(N)
    gg = _run_gpu(ctxs(N), p, inp, sg)
    rg, margin = O.find_hii_bubbles(p, inp, sr, margin=True)
    st = compare(N, _as_dict(sg), _as_dict(sr), gg, rg, margin)
    print(st)
    check(st, p999_tol=1e-4 if ftype == 1 else 5e-4)
    assert gg.volume_weighted_global_effective_bhar == pytest.approx(rg.volume_weighted_global_effective_bhar,
                                                                     rel=1e-12)
    assert gg.volume_weighted_global_effective_bhar > 0


@pytest.mark.parametrize("spin", [False, True])
def test_bh_paths_bit_identical(ctxs, spin):
    N, box = 64, 67.8
    inp = _with_bh(realisation(N, box, 1), spin_N=N if spin else None)
    p = mxr.hii_params(N, box, filter_type=1, spin_temp=spin, bh=True)
    outs = {}
    for path in (mxr.MXR_PATH_FUSED, mxr.MXR_PATH_FUSED_DENSE, mxr.MXR_PATH_PLAIN):
        st = _state(N)
        g = _run_gpu(ctxs(N, path), p, inp, st)
        outs[path] = (_as_dict(st), g)
    ref, rg = outs[mxr.MXR_PATH_FUSED]
    for path in (mxr.MXR_PATH_FUSED_DENSE, mxr.MXR_PATH_PLAIN):
        o, g = outs[path]
        for k in ref:
            assert int((o[k] != ref[k]).sum()) == 0, f"path {path} field {k}"
        assert g.volume_weighted_global_effective_bhar == rg.volume_weighted_global_effective_bhar
    ctxs(N, mxr.MXR_PATH_FUSED)


def test_bh_parity_256_direct(ctxs):
    """256^3 with BH grids: the direct-summation z pass on the radii that keep few k_z."""
    N, box = 256, 200.0
    print("BH parity N=256", flush=True)
    inp = _with_bh(realisation(N, box, 1))
    p = mxr.hii_params(N, box, filter_type=1, bh=True)
    sg, sr = _state(N), _state(N)
    gg = _run_gpu(ctxs(N), p, inp, sg)
    rg, margin = O.find_hii_bubbles(p, inp, sr, margin=True)
    st = compare(N, _as_dict(sg), _as_dict(sr), gg, rg, margin)
    print(st)
    check(st)
