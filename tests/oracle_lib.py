"""ctypes binding of oracle/liboracle.so -- the CPU restatement used as the checker.

Test infrastructure: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg only.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "meraxes-devs_qyxnew_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

import mxr  # noqa: E402

ORACLE_DIR = os.path.join(REPO, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        build()
    L = ctypes.CDLL(ORACLE_SO)
    fp = ctypes.c_void_p
    L.oracle_c2c1d_f32.argtypes = [ctypes.c_int, fp, ctypes.c_int]
    L.oracle_c2c1d_f64.argtypes = [ctypes.c_int, fp, ctypes.c_int]
    L.oracle_r2c3d_f32.argtypes = [ctypes.c_int, fp, fp]
    L.oracle_r2c3d_f64.argtypes = [ctypes.c_int, fp, fp]
    L.oracle_c2r3d_f32.argtypes = [ctypes.c_int, fp]
    L.oracle_c2r3d_f64.argtypes = [ctypes.c_int, fp]
    L.oracle_filter_f32.argtypes = [fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_float,
                                    ctypes.c_int]
    L.oracle_filter_f64.argtypes = [fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_int]
    L.oracle_RtoM.argtypes = [ctypes.POINTER(mxr.HiiParams), ctypes.c_double]
    L.oracle_RtoM.restype = ctypes.c_double
    L.oracle_hii_radii.argtypes = [ctypes.POINTER(mxr.HiiParams), ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    L.oracle_T_RECFAST.argtypes = [ctypes.c_char_p, ctypes.c_float]
    L.oracle_T_RECFAST.restype = ctypes.c_double
    L.oracle_spline_natural.argtypes = [fp, fp, ctypes.c_int, ctypes.c_double]
    L.oracle_spline_natural.restype = ctypes.c_double
    L.oracle_hubble_time.argtypes = [ctypes.POINTER(mxr.HiiParams)]
    L.oracle_hubble_time.restype = ctypes.c_double
    L.oracle_rho_crit.argtypes = [ctypes.POINTER(mxr.HiiParams)]
    L.oracle_rho_crit.restype = ctypes.c_double
    for name in ("oracle_find_hii_bubbles_f32", "oracle_find_hii_bubbles_f64"):
        getattr(L, name).argtypes = [ctypes.POINTER(mxr.HiiParams), ctypes.POINTER(mxr.HiiGrids),
                                     ctypes.POINTER(mxr.HiiGlobals), fp]
    L.oracle_find_hii_bubbles_fftw.argtypes = [ctypes.POINTER(mxr.HiiParams), ctypes.POINTER(mxr.HiiGrids),
                                               ctypes.POINTER(mxr.HiiGlobals), ctypes.POINTER(ctypes.c_double)]
    L.oracle_fully_ionized_T.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float]
    L.oracle_fully_ionized_T.restype = ctypes.c_float
    L.oracle_partially_ionized_T.argtypes = [ctypes.c_float, ctypes.c_float]
    L.oracle_partially_ionized_T.restype = ctypes.c_float
    L.oracle_ts_probe.argtypes = [ctypes.POINTER(mxr.TsParams), ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_double)]
    L.oracle_ts_probe.restype = ctypes.c_int
    L.oracle_fftw_available.restype = ctypes.c_int
    L.oracle_fftw_library.restype = ctypes.c_char_p
    L.oracle_hii_sample_f32.argtypes = [ctypes.POINTER(mxr.HiiParams), ctypes.POINTER(mxr.HiiGrids), ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.oracle_brightness_temperature.argtypes = [ctypes.POINTER(mxr.TbParams), ctypes.POINTER(mxr.TbGrids),
                                                ctypes.POINTER(ctypes.c_double)]
    L.oracle_velocity_gradient.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double]
    L.oracle_velocity_gradient.restype = ctypes.c_int
    L.oracle_num_threads.restype = ctypes.c_int
    L.oracle_prepare_grid.argtypes = [ctypes.c_int, fp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                      ctypes.POINTER(mxr.GridParams), fp]
    L.oracle_prepare_grid.restype = ctypes.c_int
    L.oracle_read_grid_gbptrees.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(mxr.GridParams),
                                            fp]
    L.oracle_read_grid_gbptrees.restype = ctypes.c_int
    L.oracle_construct_baryon_grids.argtypes = [ctypes.POINTER(mxr.BaryonParams), ctypes.POINTER(mxr.Galaxies),
                                                ctypes.POINTER(mxr.BaryonGrids), ctypes.POINTER(ctypes.c_long)]
    L.oracle_construct_baryon_grids.restype = ctypes.c_int
    L.oracle_ps_length.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
    L.oracle_ps_length.restype = ctypes.c_int
    L.oracle_power_spectrum.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, fp, fp, fp, fp, ctypes.c_int]
    L.oracle_power_spectrum.restype = ctypes.c_int
    _dp = ctypes.POINTER(ctypes.c_double)
    L.oracle_rec_set_tables.argtypes = [fp, fp, fp, fp]
    L.oracle_rec_set_tables.restype = ctypes.c_int
    L.oracle_splined_recombination.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp]
    L.oracle_splined_recombination.restype = ctypes.c_int
    L.oracle_dtdz.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_float]
    L.oracle_dtdz.restype = ctypes.c_double
    L.oracle_rec_probe.argtypes = [ctypes.POINTER(mxr.HiiParams), ctypes.c_int, _dp, _dp]
    L.oracle_rec_probe.restype = ctypes.c_int
    for name in ("oracle_ts_step_f32", "oracle_ts_step_f64"):
        getattr(L, name).argtypes = [ctypes.POINTER(mxr.TsParams), ctypes.POINTER(mxr.TsTables),
                                     ctypes.POINTER(mxr.TsGrids), ctypes.POINTER(mxr.TsGlobals), fp, fp]
        getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def _p(a):
    return a.ctypes.data


def c2c1d(x: np.ndarray, sign: int) -> np.ndarray:
    dt = np.complex64 if x.dtype == np.complex64 else np.complex128
    y = np.ascontiguousarray(x.astype(dt))
    fn = lib().oracle_c2c1d_f32 if dt == np.complex64 else lib().oracle_c2c1d_f64
    assert fn(len(y), _p(y), sign) == 0
    return y


def r2c3d(padded: np.ndarray, N: int) -> np.ndarray:
    H = N // 2 + 1
    if padded.dtype == np.float32:
        out = np.zeros(N * N * H, np.complex64)
        assert lib().oracle_r2c3d_f32(N, _p(np.ascontiguousarray(padded)), _p(out)) == 0
    else:
        out = np.zeros(N * N * H, np.complex128)
        assert lib().oracle_r2c3d_f64(N, _p(np.ascontiguousarray(padded, np.float64)), _p(out)) == 0
    return out


def c2r3d(spec: np.ndarray, N: int) -> np.ndarray:
    """Returns the padded real output (2*Nc reals)."""
    y = np.ascontiguousarray(spec.copy())
    if y.dtype == np.complex64:
        assert lib().oracle_c2r3d_f32(N, _p(y)) == 0
        return y.view(np.float32)
    assert lib().oracle_c2r3d_f64(N, _p(y)) == 0
    return y.view(np.float64)


def filter_box(spec: np.ndarray, N: int, box: float, R: float, ftype: int) -> np.ndarray:
    y = np.ascontiguousarray(spec.copy())
    if y.dtype == np.complex64:
        rc = lib().oracle_filter_f32(_p(y), 0, N, N, box, R, ftype)
    else:
        rc = lib().oracle_filter_f64(_p(y), 0, N, N, box, R, ftype)
    if rc:
        raise mxr.MxrError(rc, "oracle_filter")
    return y


def radii(params) -> list:
    n = lib().oracle_hii_radii(ctypes.byref(params), None, 0)
    buf = (ctypes.c_double * n)()
    lib().oracle_hii_radii(ctypes.byref(params), buf, n)
    return list(buf)


def find_hii_bubbles(params, inputs: dict, state: "mxr.HiiState", precision="f32", margin=False):
    N = params.ReionGridDim
    arrays = dict(inputs)
    arrays.update(state.arrays())
    g = mxr.hii_grids(**arrays)
    out = mxr.HiiGlobals()
    m = np.zeros(N ** 3, np.float64) if margin else None
    fn = lib().oracle_find_hii_bubbles_f32 if precision == "f32" else lib().oracle_find_hii_bubbles_f64
    rc = fn(ctypes.byref(params), ctypes.byref(g), ctypes.byref(out), _p(m) if margin else None)
    if rc:
        raise mxr.MxrError(rc, "oracle_find_hii_bubbles")
    return (out, m) if margin else out


def rec_set_tables(lnGamma, ln_rate, clumping, ln_residual):
    """The recombination tables the oracle's find_HII_bubbles interpolates (float64, the
    layout mxr_recombination_tables reports)."""
    arrs = [np.ascontiguousarray(a, np.float64) for a in (lnGamma, ln_rate, clumping, ln_residual)]
    assert arrs[0].size == 150 and all(a.size == 200 * 100 * 150 for a in arrs[1:])
    rc = lib().oracle_rec_set_tables(*[_p(a) for a in arrs])
    if rc:
        raise mxr.MxrError(rc, "oracle_rec_set_tables")


def splined_recombination(z_eff: float, gamma12_bg: float, temp: float):
    """recombinations.c:35-82 -> (recombination rate, residual_xH, clumping factor)."""
    r, x, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    rc = lib().oracle_splined_recombination(z_eff, gamma12_bg, temp, ctypes.byref(r), ctypes.byref(x),
                                            ctypes.byref(c))
    if rc:
        raise mxr.MxrError(rc, "oracle_splined_recombination")
    return r.value, x.value, c.value


def rec_probe(params, what: int, *args):
    """oracle_rec_probe: 0 neutral_fraction(n, T4, G12); 1 Gamma_SS(Gbg, D, T4, z); 2 (C, beta)(z);
    3 PDF without A (D, z); 4 the four integrands (lnD, z, G12, T4, A); 5 No; 6 dtdz(z)."""
    a = (ctypes.c_double * 8)(*args)
    out = (ctypes.c_double * 4)()
    rc = lib().oracle_rec_probe(ctypes.byref(params), what, a, out)
    if rc:
        raise mxr.MxrError(rc, "oracle_rec_probe")
    if what == 2:
        return out[0], out[1]
    if what == 4:
        return tuple(out[:4])
    return out[0]


def ts_probe(params, what: int, *args):
    """oracle_ts_probe: 0 get_Ts -> (T_S, x_alpha); 1 xcoll; 2 alpha_A; 3 dT_comp; 4 hubble; 5 dtdz."""
    a = (ctypes.c_double * 8)(*args)
    out = (ctypes.c_double * 2)()
    rc = lib().oracle_ts_probe(ctypes.byref(params), what, a, out)
    if rc:
        raise mxr.MxrError(rc, "oracle_ts_probe")
    return (out[0], out[1]) if what == 0 else out[0]


def fftw_available() -> bool:
    return bool(lib().oracle_fftw_available())


def find_hii_bubbles_fftw(params, inputs: dict, state: "mxr.HiiState"):
    """The FP32 restatement with its transforms on FFTW3's float API (libmkl_rt); returns
    (globals, seconds of the whole call)."""
    arrays = dict(inputs)
    arrays.update(state.arrays())
    g = mxr.hii_grids(**arrays)
    out = mxr.HiiGlobals()
    secs = ctypes.c_double()
    rc = lib().oracle_find_hii_bubbles_fftw(ctypes.byref(params), ctypes.byref(g), ctypes.byref(out),
                                            ctypes.byref(secs))
    if rc:
        raise mxr.MxrError(rc, "oracle_find_hii_bubbles_fftw")
    return out, secs.value


def hii_sample(params, inputs: dict, state, n_radii: int):
    """Returns (radii done, total seconds, seconds of the forward transforms)."""
    arrays = dict(inputs)
    arrays.update(state.arrays())
    g = mxr.hii_grids(**arrays)
    secs, fwd = ctypes.c_double(), ctypes.c_double()
    n = lib().oracle_hii_sample_f32(ctypes.byref(params), ctypes.byref(g), n_radii, ctypes.byref(secs),
                                    ctypes.byref(fwd))
    if n < 0:
        raise mxr.MxrError(n, "oracle_hii_sample_f32")
    return n, secs.value, fwd.value


def brightness_temperature(params, xH, deltax, TS_box=None, vel=None, return_vel=False):
    """The oracle's brightness box on copies of the inputs; with return_vel also the velocity grid
    after the in-place unit conversion."""
    N = params.ReionGridDim
    dT = np.zeros(N ** 3, np.float32)
    v = None if vel is None else np.ascontiguousarray(vel, np.float32).copy()
    g = mxr.TbGrids()
    g.mem = 0
    g.xH, g.deltax, g.TS_box, g.delta_T = _p(xH), _p(deltax), (_p(TS_box) if TS_box is not None else None), _p(dT)
    g.vel = _p(v) if v is not None else None
    ave = ctypes.c_double()
    rc = lib().oracle_brightness_temperature(ctypes.byref(params), ctypes.byref(g), ctypes.byref(ave))
    if rc:
        raise mxr.MxrError(rc, "oracle_brightness_temperature")
    return (dT, ave.value, v) if return_vel else (dT, ave.value)


def velocity_gradient(box: np.ndarray, N: int, box_size: float) -> np.ndarray:
    """velocity_gradient on a copy of a complex64 half spectrum (N, N, N/2+1)."""
    out = np.ascontiguousarray(box, np.complex64).copy()
    rc = lib().oracle_velocity_gradient(out.ctypes.data, N, N, box_size)
    if rc:
        raise mxr.MxrError(rc, "oracle_velocity_gradient")
    return out


def prepare_grid(N: int, hi: np.ndarray, n_hi: int, prop: int, file_box: float, params) -> np.ndarray:
    out = np.zeros(N * N * 2 * (N // 2 + 1), np.float32)
    rc = lib().oracle_prepare_grid(N, _p(np.ascontiguousarray(hi, np.float32)), n_hi, prop, file_box,
                                   ctypes.byref(params), _p(out))
    if rc:
        raise mxr.MxrError(rc, "oracle_prepare_grid")
    return out


def read_grid_gbptrees(fname: str, prop: int, N: int, params) -> np.ndarray:
    out = np.zeros(N * N * 2 * (N // 2 + 1), np.float32)
    rc = lib().oracle_read_grid_gbptrees(fname.encode(), prop, N, ctypes.byref(params), _p(out))
    if rc:
        raise mxr.MxrError(rc, "oracle_read_grid_gbptrees")
    return out


def construct_baryon_grids(params, galaxies: dict, grids: dict) -> int:
    """The oracle's construct_baryon_grids on host arrays; grids are written in place."""
    g = mxr.Galaxies()
    g.mem = 0
    g.n = len(galaxies["Type"])
    for f in mxr.GALAXY_FIELDS:
        a = galaxies.get(f)
        setattr(g, f, _p(a) if a is not None else None)
    o = mxr.BaryonGrids()
    o.mem = 0
    for f in mxr.BARYON_GRIDS:
        a = grids.get(f)
        setattr(o, f, _p(a) if a is not None else None)
    below = ctypes.c_long()
    rc = lib().oracle_construct_baryon_grids(ctypes.byref(params), ctypes.byref(g), ctypes.byref(o), ctypes.byref(below))
    if rc:
        raise mxr.MxrError(rc, "oracle_construct_baryon_grids")
    return below.value


def ps_length(N: int, box_size: float, hubble_h: float) -> int:
    return lib().oracle_ps_length(N, box_size, hubble_h)


def power_spectrum(delta_T: np.ndarray, N: int, box_size: float, hubble_h: float):
    """The oracle's Compute_PS of a real grid: (PS_k, PS_data, PS_error) float32 arrays."""
    nb = ps_length(N, box_size, hubble_h)
    dT = np.ascontiguousarray(delta_T, np.float32).reshape(-1)
    assert dT.size == N ** 3
    out = [np.zeros(nb, np.float32) for _ in range(3)]
    rc = lib().oracle_power_spectrum(N, box_size, hubble_h, _p(dT), *[_p(o) for o in out], nb)
    if rc:
        raise mxr.MxrError(rc, "oracle_power_spectrum")
    return tuple(out)


def ts_step(params, tables, deltax, stars, sfr_histories, xray_histories, n_hist, x_e_box_prev, Tk_box,
            want_smoothed=False, f64=False, bh_soft_histories=None, bh_hard_histories=None):
    """One oracle spin-temperature step on copies of the state.  Returns a dict with the new
    x_e_box_prev, x_e_box, Tk_box, TS_box, the globals, red0 (collapse fraction, <x_e>, <L_X>,
    <AGN soft>, <AGN hard>) and (optionally) the smoothed shells [G, F, N^3] (SFR, X-ray, then
    the AGN bands params.Flag_IncludeAGNXray includes, soft before hard)."""
    N = params.ReionGridDim
    npad, nreal = mxr.padded_len(N), N ** 3
    xe = np.array(x_e_box_prev, np.float32, copy=True)
    Tk = np.array(Tk_box, np.float32, copy=True)
    xe_out = np.zeros(npad, np.float32)
    TS = np.zeros(nreal, np.float32)
    arrs = [np.ascontiguousarray(a, np.float32) for a in (deltax, stars, sfr_histories, xray_histories)]
    bh = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (bh_soft_histories, bh_hard_histories)]
    g = mxr.TsGrids()
    g.mem = mxr.MXR_MEM_HOST
    g.deltax, g.stars, g.sfr_histories, g.xray_luminosity_histories = (_p(a) for a in arrs)
    g.bh_xray_histories_soft = _p(bh[0]) if bh[0] is not None else None
    g.bh_xray_histories_hard = _p(bh[1]) if bh[1] is not None else None
    flag = params.Flag_IncludeAGNXray
    G = 2 + (flag in (1, 3)) + (flag in (1, 2))
    g.n_histories = n_hist
    g.x_e_box_prev, g.x_e_box, g.Tk_box, g.TS_box = _p(xe), _p(xe_out), _p(Tk), _p(TS)
    sf = np.zeros(1, np.float64)
    g.stored_fcoll = _p(sf)
    out = mxr.TsGlobals()
    red0 = np.zeros(5, np.float64)
    sm = np.zeros(G * max(tables.n, 1) * nreal, np.float32) if want_smoothed else None
    fn = lib().oracle_ts_step_f64 if f64 else lib().oracle_ts_step_f32
    rc = fn(ctypes.byref(params), ctypes.byref(tables), ctypes.byref(g), ctypes.byref(out),
                                  _p(red0), _p(sm) if sm is not None else None)
    if rc != 0:
        raise RuntimeError(f"oracle ts step failed: {rc}")
    res = dict(x_e_box_prev=xe, x_e_box=xe_out, Tk_box=Tk, TS_box=TS, globals=out.as_dict(), red0=red0)
    if sm is not None:
        res["smoothed"] = sm.reshape(G, tables.n, nreal)
    return res
