"""Writes gbpTrees grid files (read_grids-gbptrees.c:38-86 layout) for the grid tests."""
import numpy as np


def write_gbptrees(path, grids, box, n_grids=4, n_cell=None, ma_scheme=1):
    """grids: list of n^3 float32 arrays (x slowest); written with 32-byte identifiers."""
    n = grids[0].shape[0] if n_cell is None else n_cell[0]
    nc = np.array(n_cell if n_cell is not None else [n, n, n], np.int32)
    with open(path, "wb") as f:
        f.write(nc.tobytes())
        f.write(np.array([box, box, box], np.float64).tobytes())
        f.write(np.array([n_grids, ma_scheme], np.int32).tobytes())
        for k, g in enumerate(grids):
            name = ("rho", "v_x", "v_y", "v_z")[k % 4].encode().ljust(32, b"\0")
            f.write(name)
            f.write(np.ascontiguousarray(g, np.float32).tobytes())


def density_grids(n, seed, mean=8.0):
    """A positive density-like grid (counts times particle mass) and three velocity grids."""
    rng = np.random.default_rng(seed)
    rho = rng.gamma(2.0, mean / 2.0, (n, n, n)).astype(np.float32)
    vel = [rng.normal(0.0, 100.0, (n, n, n)).astype(np.float32) for _ in range(3)]
    return [rho] + vel
