"""Parity statistics between a GPU find_HII_bubbles result and the oracle's.

Bounds (stated once, used by every parity test):
  * ionised-flag mismatch (xH == 0 on one side only) <= 0.05% of cells, or <= 2 cells on grids
    too small for a percentage to mean anything;
  * every mismatched cell sits at the threshold: |margin| <= MARGIN_TOL, where margin is the
    oracle's max over radii of (LHS/RHS - 1) of the ionisation criterion;
  * |global xH difference| <= 1e-5 + (mismatched cells) / N^3 -- a flipped cell moves the
    volume mean by at most 1/N^3;
  * cells whose flag agrees: xH within 1e-4 absolute (partial ionisation of the last step is an
    FP32 value computed from FFT output), z_at_ionization identical, J_21_at_ionization within
    1e-4 relative, temp_kinetic_all_gas within 1e-4 relative.
"""
from __future__ import annotations

import numpy as np

MARGIN_TOL = 1e-3
FLAG_FRAC = 5e-4


def compare(N, gpu: dict, ref: dict, gglob, rglob, margin=None, flag_frac=FLAG_FRAC, min_cells=2):
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
    stats["max_dxH_cell"] = float(np.abs(gpu["xH"][ok] - ref["xH"][ok]).max()) if ok.any() else 0.0
    zg, zr = gpu["z_at_ionization"][ok], ref["z_at_ionization"][ok]
    stats["z_in_diff_cells"] = int((zg != zr).sum())
    jg, jr = gpu["J_21_at_ionization"][ok], ref["J_21_at_ionization"][ok]
    stats["max_rel_J21"] = float((np.abs(jg - jr) / np.maximum(np.abs(jr), 1e-30)).max()) if ok.any() else 0.0
    tg, tr = gpu["temp_kinetic_all_gas"][ok], ref["temp_kinetic_all_gas"][ok]
    stats["max_rel_T"] = float((np.abs(tg - tr) / np.maximum(np.abs(tr), 1e-30)).max()) if ok.any() else 0.0
    return stats


def check(stats, flag_frac=FLAG_FRAC, min_cells=2):
    N = stats["N"]
    n = N ** 3
    assert stats["mismatch"] <= max(min_cells, flag_frac * n), stats
    if stats["mismatch"]:
        assert stats.get("max_margin_mismatch", 0.0) <= MARGIN_TOL, stats
    assert stats["dxH_global"] <= 1e-5 + stats["mismatch"] / n, stats
    assert stats["max_dxH_cell"] <= 1e-4, stats
    assert stats["z_in_diff_cells"] == 0, stats
    assert stats["max_rel_J21"] <= 1e-4, stats
    assert stats["max_rel_T"] <= 1e-4, stats
