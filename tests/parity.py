"""Parity statistics between a GPU find_HII_bubbles result and the oracle's.

Bounds (stated once, used by every parity test):
  * ionised-flag mismatch (xH == 0 on one side only) <= 0.05% of cells, or <= 2 cells on grids
    too small for a percentage to mean anything;
  * every mismatched cell sits at the threshold: |margin| <= MARGIN_TOL, where margin is the
    oracle's max over radii of (LHS/RHS - 1) of the ionisation criterion;
  * |global xH difference| <= 1e-5 + (mismatched cells) / N^3 -- a flipped cell moves the
    volume mean by at most 1/N^3;
  * cells whose flag agrees: z_at_ionization identical, temp_kinetic_all_gas within 1e-4
    relative; the last step's partial xH (1 - f_coll, an FP32 value taken from c2r output whose
    absolute FFT rounding scales with the largest |stars| in the box) within 1e-4 absolute at
    the 99.9th percentile (5e-4 for the multiplicative top-hat and Gaussian windows) and 1e-2
    at worst;
  * J_21_at_ionization is set at the first ionising radius; a cell whose crossing radius sits at
    the threshold can cross one radius earlier or later on the other side, so cells differing by
    more than 1e-3 relative count like flag mismatches (<= 0.05%).
"""
from __future__ import annotations

import numpy as np

MARGIN_TOL = 1e-3
FLAG_FRAC = 5e-4


def compare(N, gpu: dict, ref: dict, gglob, rglob, margin=None, flag_frac=FLAG_FRAC, min_cells=2, exclude=None):
    """`exclude` (optional mask): cells that disagreed in an earlier call of a sequence, whose
    carried state (z_at_ionization, J_21, N_rec) keeps that difference; they are left out of the
    per-cell comparisons of agreeing cells."""
    n = N ** 3
    ion_g = gpu["xH"] == 0
    ion_r = ref["xH"] == 0
    mism = ion_g != ion_r
    nm = int(mism.sum())
    stats = {"N": N, "mismatch": nm, "mismatch_frac": nm / n}
    stats["dxH_global"] = abs(gglob.volume_weighted_global_xH - rglob.volume_weighted_global_xH)
    stats["xH_global_ref"] = rglob.volume_weighted_global_xH
    if margin is not None and nm:
        stats["max_margin_mismatch"] = float(np.abs(margin[mism]).max())
    ok = ~mism
    if exclude is not None:
        ok &= ~exclude
        stats["history_excluded"] = int(exclude.sum())
    dx = np.abs(gpu["xH"][ok] - ref["xH"][ok]) if ok.any() else np.zeros(1)
    stats["max_dxH_cell"] = float(dx.max())
    stats["p999_dxH_cell"] = float(np.quantile(dx, 0.999))
    zg, zr = gpu["z_at_ionization"][ok], ref["z_at_ionization"][ok]
    stats["z_in_diff_cells"] = int((zg != zr).sum())
    jg, jr = gpu["J_21_at_ionization"][ok], ref["J_21_at_ionization"][ok]
    rel = np.abs(jg - jr) / np.maximum(np.abs(jr), 1e-30) if ok.any() else np.zeros(1)
    stats["max_rel_J21"] = float(rel.max())
    stats["J21_radius_flips"] = int((rel > 1e-3).sum())
    tg, tr = gpu["temp_kinetic_all_gas"][ok], ref["temp_kinetic_all_gas"][ok]
    stats["max_rel_T"] = float((np.abs(tg - tr) / np.maximum(np.abs(tr), 1e-30)).max()) if ok.any() else 0.0
    return stats


def check(stats, flag_frac=FLAG_FRAC, min_cells=2, p999_tol=1e-4):
    N = stats["N"]
    n = N ** 3
    assert stats["mismatch"] <= max(min_cells, flag_frac * n), stats
    if stats["mismatch"]:
        assert stats.get("max_margin_mismatch", 0.0) <= MARGIN_TOL, stats
    assert stats["dxH_global"] <= 1e-5 + stats["mismatch"] / n, stats
    assert stats["p999_dxH_cell"] <= p999_tol, stats
    assert stats["max_dxH_cell"] <= 1e-2, stats
    assert stats["z_in_diff_cells"] == 0, stats
    assert stats["J21_radius_flips"] <= max(min_cells, flag_frac * n), stats
    assert stats["max_rel_T"] <= 1e-4, stats
