"""The benchmark's workloads against the oracle, each result also against the FP64 oracle.

  * plain one-kernel-per-step path at 256^3, 200 cMpc/h, sharp-k (the comparator form);
  * 512^3, 400 cMpc/h, sharp-k, the fused path the benchmark times, and the same call with
    stars / weighted_sfr stored as bfloat16;
  * a batch of independent realisations with N drawn from {96, 128, 160}, boxes 100-150 cMpc/h
    and ReionEfficiency in [0.5, 2], run as one captured graph, replayed, and every realisation
    checked against the oracle.

Bounds (tests/parity.py): ionised-flag mismatch <= 0.05% of cells, every mismatch within 1e-3 of
the threshold, |dxH_global| <= 1e-5 plus the flipped cells' share (1/N^3 each).  Against the FP64
oracle the GPU's flag mismatch and global-xH error may not exceed twice the FP32 restatement's
own (both take FP32 FFTs of the same fields) plus the same absolute bounds.
"""
import numpy as np
import pytest

import mxr
import oracle_lib as O
from parity import check, compare

pytestmark = pytest.mark.gpu


def _as_dict(st):
    return {k: np.asarray(v) for k, v in st.arrays().items()}


def _fp64_stats(N, gpu_xH, gpu_glob, s32, g32, s64, g64, margin64):
    """Flag mismatch and global-xH error of the GPU and of the FP32 restatement, both against
    the FP64 oracle; asserts the GPU within the bounds above."""
    n = N ** 3
    ion64 = s64.xH == 0
    mg = (gpu_xH == 0) != ion64
    m32 = (s32.xH == 0) != ion64
    st = {"N": N, "gpu_flags_vs_fp64": int(mg.sum()), "fp32_oracle_flags_vs_fp64": int(m32.sum()),
          "gpu_dxH_vs_fp64": abs(gpu_glob.volume_weighted_global_xH - g64.volume_weighted_global_xH),
          "fp32_oracle_dxH_vs_fp64": abs(g32.volume_weighted_global_xH - g64.volume_weighted_global_xH),
          "xH_fp64": g64.volume_weighted_global_xH}
    if mg.any():
        st["gpu_max_margin_mismatch"] = float(np.abs(margin64[mg]).max())
    print("vs FP64 oracle:", st, flush=True)
    assert st["gpu_flags_vs_fp64"] <= max(2, 5e-4 * n, 2 * st["fp32_oracle_flags_vs_fp64"])
    assert st.get("gpu_max_margin_mismatch", 0.0) <= 1e-3
    assert st["gpu_dxH_vs_fp64"] <= max(1e-5, 2 * st["fp32_oracle_dxH_vs_fp64"]) + st["gpu_flags_vs_fp64"] / n
    return st


@pytest.mark.timeout(600)
def test_plain_path_256_vs_oracle():
    """The plain one-kernel-per-step path at 256^3 (copy, filter, three c2r passes, clamp and
    criterion per radius) against the FP32 restatement and the FP64 oracle."""
    from conftest import realisation

    N, box = 256, 200.0
    print("plain path N=256 vs oracle", flush=True)
    inp = realisation(N, box, 1)
    p = mxr.hii_params(N, box, filter_type=1)
    ctx = mxr.Context(N, path=mxr.MXR_PATH_PLAIN)
    sg = mxr.HiiState.host(N)
    gg = ctx.find_hii_bubbles(p, **inp, **sg.arrays())
    ctx.close()
    s32, s64 = mxr.HiiState.host(N), mxr.HiiState.host(N)
    g32, margin = O.find_hii_bubbles(p, inp, s32, margin=True)
    g64, margin64 = O.find_hii_bubbles(p, inp, s64, precision="f64", margin=True)
    st = compare(N, _as_dict(sg), _as_dict(s32), gg, g32, margin)
    print("vs FP32 restatement:", st, flush=True)
    check(st)
    assert gg.n_radii == g32.n_radii == 44
    assert 0.05 < g32.volume_weighted_global_xH < 0.95
    _fp64_stats(N, np.asarray(sg.xH), gg, s32, g32, s64, g64, margin64)


# ---- 512^3: the benchmark's own call --------------------------------------------------------

SEED_512, S2M_512 = 2024, 0.010438


@pytest.fixture(scope="module")
def w512():
    """Inputs of the 512^3 benchmark workload and the FP32 and FP64 oracle runs on them."""
    import synth
    import torch

    N, box = 512, 400.0
    print("512: inputs and both oracles", flush=True)
    dev = synth.realisation(N, box, SEED_512, device="cuda", stars_to_matter=S2M_512)
    inp = {k: v.cpu().numpy() for k, v in dev.items()}
    del dev
    torch.cuda.empty_cache()
    p = mxr.hii_params(N, box, filter_type=1)
    s32, s64 = mxr.HiiState.host(N), mxr.HiiState.host(N)
    g32, margin32 = O.find_hii_bubbles(p, inp, s32, margin=True)
    print("512: FP32 oracle done", flush=True)
    g64, margin64 = O.find_hii_bubbles(p, inp, s64, precision="f64", margin=True)
    print("512: FP64 oracle done", flush=True)
    return dict(N=N, box=box, p=p, inp=inp, s32=s32, g32=g32, m32=margin32, s64=s64, g64=g64, m64=margin64)


@pytest.mark.timeout(900)
def test_parity_512_headline_vs_oracles(w512):
    """The benchmark's workload (512^3, 400 cMpc/h, sharp-k, seed 2024, stars-to-matter 0.010438,
    about half ionised) through the fused path, cell by cell against the FP32 restatement, and
    its flags and global xH against the FP64 oracle."""
    N, p, inp = w512["N"], w512["p"], w512["inp"]
    print("parity 512 headline vs oracles", flush=True)
    ctx = mxr.Context(N)
    sg = mxr.HiiState.host(N)
    gg = ctx.find_hii_bubbles(p, **inp, **sg.arrays())
    ctx.close()
    st = compare(N, _as_dict(sg), _as_dict(w512["s32"]), gg, w512["g32"], w512["m32"])
    print("vs FP32 restatement:", st, flush=True)
    assert 0.3 < w512["g32"].volume_weighted_global_xH < 0.7
    check(st)
    assert gg.n_radii == 44
    assert abs(gg.mass_weighted_global_xH - w512["g32"].mass_weighted_global_xH) <= 1e-5 + 4 * st["mismatch"] / N ** 3
    _fp64_stats(N, np.asarray(sg.xH), gg, w512["s32"], w512["g32"], w512["s64"], w512["g64"], w512["m64"])


@pytest.mark.timeout(900)
@pytest.mark.xfail(strict=True, reason="BF16 storage of stars/sfr (8-bit mantissa) flips about 0.09% of the flags "
                                       "at 512^3 and moves global xH by ~8e-4 against FP64, outside 0.05% and 1e-5")
def test_bf16_storage_512_meets_bounds(w512):
    """The 512^3 call with stars and weighted_sfr stored as bfloat16 (FP32 compute) against the
    FP64 oracle on the original FP32 grids: the flag and global-xH bounds of the FP32 path."""
    import torch

    N, p, inp = w512["N"], w512["p"], w512["inp"]
    print("bf16 512 vs FP64 oracle", flush=True)
    dev = {k: torch.from_numpy(v).cuda() for k, v in inp.items()}
    dev["stars"] = dev["stars"].to(torch.bfloat16)
    dev["weighted_sfr"] = dev["weighted_sfr"].to(torch.bfloat16)
    n = N ** 3
    st = mxr.HiiState(torch.ones(n, device="cuda"), torch.full((n,), -1.0, device="cuda"),
                      torch.zeros(n, device="cuda"), torch.zeros(n, device="cuda"), torch.zeros(n, device="cuda"))
    ctx = mxr.Context(N)
    gb = ctx.find_hii_bubbles(p, **dev, **st.arrays())
    torch.cuda.synchronize()
    ctx.close()
    xh = st.xH.cpu().numpy()
    del dev, st
    torch.cuda.empty_cache()
    ion64 = w512["s64"].xH == 0
    flips = int(((xh == 0) != ion64).sum())
    dxh = abs(gb.volume_weighted_global_xH - w512["g64"].volume_weighted_global_xH)
    print(f"bf16 512 vs FP64: flag mismatch {flips} ({flips / n * 100:.4f}% of cells), |dxH| {dxh:.3e}", flush=True)
    assert np.isfinite(gb.volume_weighted_global_xH)
    assert flips <= 5e-4 * n
    assert dxh <= 1e-5


# ---- batches of uneven realisations -----------------------------------------------------------

# stars-to-matter ratios of the seeded single-realisation tests at these dimensions
BATCH_S2M = {96: 0.01253, 128: 0.19495, 160: 0.02558}


@pytest.mark.timeout(600)
def test_batch_uneven_vs_oracle():
    """Nine realisations, three each of N = 96, 128, 160, with their own seed, box (100-150
    cMpc/h) and ReionEfficiency (0.5-2), in one captured graph; the graph is then replayed on
    reset state.  Every realisation of both launches is checked against the FP32 restatement,
    and against the FP64 oracle."""
    import synth
    import torch

    rng = np.random.default_rng(77)
    dims = [96, 128, 160] * 3
    boxes = rng.uniform(100.0, 150.0, len(dims))
    zetas = rng.uniform(0.5, 2.0, len(dims))
    params, arrays, host = [], [], []
    for i, N in enumerate(dims):
        p = mxr.hii_params(N, float(boxes[i]), filter_type=1, zeta=float(zetas[i]))
        inp = synth.numpy_realisation(N, float(boxes[i]), 5000 + i, stars_to_matter=BATCH_S2M[N])
        m = N ** 3
        st = dict(xH=torch.ones(m, device="cuda"), z_at_ionization=torch.full((m,), -1.0, device="cuda"),
                  J_21_at_ionization=torch.zeros(m, device="cuda"), r_bubble=torch.zeros(m, device="cuda"),
                  temp_kinetic_all_gas=torch.zeros(m, device="cuda"))
        params.append(p)
        arrays.append(dict({k: torch.from_numpy(v).cuda() for k, v in inp.items()}, **st))
        host.append(inp)
    b = mxr.Batch(dims)
    launches = []
    for launch in range(2):
        for a in arrays:
            a["xH"].fill_(1.0)
            a["z_at_ionization"].fill_(-1.0)
            a["J_21_at_ionization"].zero_()
            a["r_bubble"].zero_()
            a["temp_kinetic_all_gas"].zero_()
        gl = b.find_hii_bubbles(params, arrays, mxr.MXR_BATCH_GRAPH)
        torch.cuda.synchronize()
        launches.append(([{k: a[k].cpu().numpy() for k in mxr.HII_OUT} for a in arrays], gl))
        print(f"batch launch {launch}: {gl[0].device_ms:.2f} ms", flush=True)
    b.close()
    for i, N in enumerate(dims):
        p, inp = params[i], host[i]
        s32, s64 = mxr.HiiState.host(N), mxr.HiiState.host(N)
        g32, margin = O.find_hii_bubbles(p, inp, s32, margin=True)
        g64, margin64 = O.find_hii_bubbles(p, inp, s64, precision="f64", margin=True)
        print(f"realisation {i}: N={N} box={boxes[i]:.1f} zeta={zetas[i]:.3f} xH {g32.volume_weighted_global_xH:.4f}",
              flush=True)
        for launch, (fields, gl) in enumerate(launches):
            st = compare(N, fields[i], _as_dict(s32), gl[i], g32, margin)
            print(f"  launch {launch}:", st, flush=True)
            check(st)
            assert gl[i].n_radii == g32.n_radii
            _fp64_stats(N, fields[i]["xH"], gl[i], s32, g32, s64, g64, margin64)
        for k in mxr.HII_OUT:
            np.testing.assert_array_equal(launches[0][0][i][k], launches[1][0][i][k], err_msg=f"replay {i} {k}")
