"""ctypes binding of libmxr.so, the MI355X reionization grid path (C ABI in include/mxr.h).

The host side a Python caller (tests, bench) uses.  Device memory is handed over as plain
pointers: torch tensors on ``cuda`` are used only as allocations (``data_ptr()``), numpy
arrays as host buffers.  There is no CPU fallback: if libmxr.so is missing or no gfx950 device
is present every call raises.

Parameter names follow the reference's run_params_t / physics_params_t / run_units_t
(src/meraxes.h:142-201, 262-398).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libmxr.so")
TABLES_DIR = os.path.join(PKG_DIR, "tables")

MXR_OK = 0
MXR_ERR_INVALID_ARG = -1
MXR_ERR_UNSUPPORTED = -2
MXR_ERR_HIP = -3
MXR_ERR_OOM = -4
MXR_ERR_TABLES = -5
MXR_ERR_NO_DEVICE = -6
MXR_ERR_FILTER = -7

MXR_MEM_HOST = 0
MXR_MEM_DEVICE = 1
MXR_PATH_FUSED = 0
MXR_PATH_PLAIN = 1
MXR_PATH_FUSED_DENSE = 2

_fp = ctypes.POINTER(ctypes.c_float)


class HiiParams(ctypes.Structure):
    _fields_ = [
        ("ReionGridDim", ctypes.c_int),
        ("BoxSize", ctypes.c_double),
        ("Hubble_h", ctypes.c_double),
        ("OmegaM", ctypes.c_double),
        ("OmegaLambda", ctypes.c_double),
        ("OmegaK", ctypes.c_double),
        ("BaryonFrac", ctypes.c_double),
        ("ReionDeltaRFactor", ctypes.c_double),
        ("ReionFilterType", ctypes.c_int),
        ("ReionRtoMFilterType", ctypes.c_int),
        ("ReionUVBFlag", ctypes.c_int),
        ("Flag_IncludeSpinTemp", ctypes.c_int),
        ("Flag_IncludeRecombinations", ctypes.c_int),
        ("Flag_EvolvingReionRBubbleMax", ctypes.c_int),
        ("Flag_BHFeedback", ctypes.c_int),
        ("ReionSfrTimescale", ctypes.c_double),
        ("ReionEfficiency", ctypes.c_double),
        ("ReionNionPhotPerBary", ctypes.c_double),
        ("ReionGammaHaloBias", ctypes.c_double),
        ("ReionAlphaUV", ctypes.c_double),
        ("ReionRBubbleMax", ctypes.c_double),
        ("ReionRBubbleMaxRecomb", ctypes.c_double),
        ("ReionRBubbleMin", ctypes.c_double),
        ("UnitLength_in_cm", ctypes.c_double),
        ("UnitMass_in_g", ctypes.c_double),
        ("UnitVelocity_in_cm_per_s", ctypes.c_double),
        ("redshift", ctypes.c_double),
        ("prev_redshift", ctypes.c_double),
        ("TK_recfast", ctypes.c_double),
        ("TablesForXHeatingDir", ctypes.c_char_p),
    ]


class HiiGrids(ctypes.Structure):
    _fields_ = [
        ("mem", ctypes.c_int),
        ("deltax", ctypes.c_void_p),
        ("stars", ctypes.c_void_p),
        ("weighted_sfr", ctypes.c_void_p),
        ("x_e_box", ctypes.c_void_p),
        ("Tk_box", ctypes.c_void_p),
        ("xH", ctypes.c_void_p),
        ("z_at_ionization", ctypes.c_void_p),
        ("J_21_at_ionization", ctypes.c_void_p),
        ("r_bubble", ctypes.c_void_p),
        ("temp_kinetic_all_gas", ctypes.c_void_p),
    ]


class HiiGlobals(ctypes.Structure):
    _fields_ = [
        ("volume_weighted_global_xH", ctypes.c_double),
        ("volume_weighted_global_Gamma12", ctypes.c_double),
        ("volume_weighted_global_r_bubble", ctypes.c_double),
        ("volume_weighted_global_weighted_sfr", ctypes.c_double),
        ("volume_weighted_global_effective_bhar", ctypes.c_double),
        ("volume_weighted_global_temp_kinetic_all_gas", ctypes.c_double),
        ("volume_weighted_global_N_rec", ctypes.c_double),
        ("volume_weighted_global_residual_xH", ctypes.c_double),
        ("volume_weighted_global_clumping_factor", ctypes.c_double),
        ("mass_weighted_global_xH", ctypes.c_double),
        ("mass_weighted_global_Gamma12", ctypes.c_double),
        ("mass_weighted_global_r_bubble", ctypes.c_double),
        ("mass_weighted_global_temp_kinetic_all_gas", ctypes.c_double),
        ("mass_weighted_global_N_rec", ctypes.c_double),
        ("mass_weighted_global_residual_xH", ctypes.c_double),
        ("mass_weighted_global_clumping_factor", ctypes.c_double),
        ("n_radii", ctypes.c_int),
        ("device_ms", ctypes.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class KernelStat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("launches", ctypes.c_int), ("total_ms", ctypes.c_double),
                ("alg_bytes", ctypes.c_double)]


class TbParams(ctypes.Structure):
    _fields_ = [
        ("ReionGridDim", ctypes.c_int),
        ("Hubble_h", ctypes.c_double),
        ("OmegaM", ctypes.c_double),
        ("BaryonFrac", ctypes.c_double),
        ("redshift", ctypes.c_double),
        ("Flag_IncludeSpinTemp", ctypes.c_int),
        ("Flag_IncludePecVelsFor21cm", ctypes.c_int),
    ]


class TbGrids(ctypes.Structure):
    _fields_ = [
        ("mem", ctypes.c_int),
        ("xH", ctypes.c_void_p),
        ("deltax", ctypes.c_void_p),
        ("TS_box", ctypes.c_void_p),
        ("delta_T", ctypes.c_void_p),
    ]


# ---------------------------------------------------------------------------------------
# Parameter sets.  Cosmology and units: Genesis L10_N512
# (input/params/simulations/Genesis_L10_N512.par:17-31); reionisation parameters from
# input/params/defaults.par with Flag_EvolvingReionRBubbleMax = 0 and R_max = 50 Mpc/h.

GENESIS = dict(Hubble_h=0.6751, OmegaM=0.3121, OmegaLambda=0.6879, OmegaK=0.0, BaryonFrac=0.15714,
               UnitLength_in_cm=3.08568e24, UnitMass_in_g=1.989e43, UnitVelocity_in_cm_per_s=1e5)


def hii_params(N: int, box: float, filter_type: int = 1, zeta: float = 1.0, redshift: float = 8.0,
               spin_temp: bool = False, tables_dir: str | None = None, **over) -> HiiParams:
    p = HiiParams()
    p.ReionGridDim = N
    p.BoxSize = box
    for k, v in GENESIS.items():
        setattr(p, k, v)
    p.ReionDeltaRFactor = 1.1
    p.ReionFilterType = filter_type
    p.ReionRtoMFilterType = 0
    p.ReionUVBFlag = 1
    p.Flag_IncludeSpinTemp = int(spin_temp)
    p.Flag_IncludeRecombinations = 0
    p.Flag_EvolvingReionRBubbleMax = 0
    p.Flag_BHFeedback = 0
    p.ReionSfrTimescale = 0.5
    p.ReionEfficiency = zeta
    p.ReionNionPhotPerBary = 4000.0
    p.ReionGammaHaloBias = 2.0
    p.ReionAlphaUV = 2.0
    p.ReionRBubbleMax = 50.0
    p.ReionRBubbleMaxRecomb = 50.0
    p.ReionRBubbleMin = 0.4068
    p.redshift = redshift
    p.prev_redshift = redshift
    p.TK_recfast = -1.0
    p.TablesForXHeatingDir = (tables_dir or TABLES_DIR).encode()
    for k, v in over.items():
        setattr(p, k, v)
    return p


def tb_params(N: int, redshift: float = 8.0, spin_temp: bool = False) -> TbParams:
    p = TbParams()
    p.ReionGridDim = N
    p.Hubble_h = GENESIS["Hubble_h"]
    p.OmegaM = GENESIS["OmegaM"]
    p.BaryonFrac = GENESIS["BaryonFrac"]
    p.redshift = redshift
    p.Flag_IncludeSpinTemp = int(spin_temp)
    p.Flag_IncludePecVelsFor21cm = 0
    return p


def padded_len(N: int) -> int:
    return N * N * 2 * (N // 2 + 1)


# ---------------------------------------------------------------------------------------
class MxrError(RuntimeError):
    def __init__(self, code: int, where: str, detail: str = ""):
        self.code = code
        super().__init__(f"{where}: {code} {detail}")


_lib = None


def lib():
    """Load libmxr.so (raises if it was not built: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MxrError(MXR_ERR_NO_DEVICE, "libmxr.so", f"not built at {LIB_PATH}")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    L.mxr_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int]
    L.mxr_create.restype = ctypes.c_int
    L.mxr_destroy.argtypes = [vp]
    L.mxr_destroy.restype = None
    L.mxr_strerror.argtypes = [ctypes.c_int]
    L.mxr_strerror.restype = ctypes.c_char_p
    L.mxr_last_error.argtypes = []
    L.mxr_last_error.restype = ctypes.c_char_p
    L.mxr_set_path.argtypes = [vp, ctypes.c_int]
    L.mxr_set_path.restype = ctypes.c_int
    L.mxr_find_hii_bubbles.argtypes = [vp, ctypes.POINTER(HiiParams), ctypes.POINTER(HiiGrids),
                                       ctypes.POINTER(HiiGlobals)]
    L.mxr_find_hii_bubbles.restype = ctypes.c_int
    L.mxr_filter.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                             ctypes.c_float, ctypes.c_int]
    L.mxr_filter.restype = ctypes.c_int
    L.mxr_RtoM.argtypes = [ctypes.POINTER(HiiParams), ctypes.c_double]
    L.mxr_RtoM.restype = ctypes.c_double
    L.mxr_hii_radii.argtypes = [ctypes.POINTER(HiiParams), ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    L.mxr_hii_radii.restype = ctypes.c_int
    L.mxr_brightness_temperature.argtypes = [vp, ctypes.POINTER(TbParams), ctypes.POINTER(TbGrids),
                                             ctypes.POINTER(ctypes.c_double)]
    L.mxr_brightness_temperature.restype = ctypes.c_int
    L.mxr_profile_enable.argtypes = [vp, ctypes.c_int]
    L.mxr_profile_enable.restype = ctypes.c_int
    L.mxr_profile_read.argtypes = [vp, ctypes.POINTER(KernelStat), ctypes.c_int]
    L.mxr_profile_read.restype = ctypes.c_int
    _lib = L
    return L


def check(code: int, where: str):
    if code != MXR_OK:
        L = lib()
        raise MxrError(code, where, f"{L.mxr_strerror(code).decode()} ({L.mxr_last_error().decode()})")


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()  # torch tensor


def _mem_of(arrays):
    devs = {isinstance(a, np.ndarray) or (hasattr(a, "is_cuda") and not a.is_cuda) for a in arrays if a is not None}
    if len(devs) != 1:
        raise ValueError("all grids must be on the same side (host numpy / device torch)")
    return MXR_MEM_HOST if devs.pop() else MXR_MEM_DEVICE


HII_IN = ("deltax", "stars", "weighted_sfr", "x_e_box", "Tk_box")
HII_OUT = ("xH", "z_at_ionization", "J_21_at_ionization", "r_bubble", "temp_kinetic_all_gas")


def hii_grids(**arrays) -> HiiGrids:
    g = HiiGrids()
    g.mem = _mem_of(list(arrays.values()))
    for name in HII_IN + HII_OUT:
        setattr(g, name, _ptr(arrays.get(name)))
    return g


def _sync_torch():
    """Device tensors handed over by torch may still be written on torch's stream; the library
    runs on its own stream, so wait for torch's work first."""
    import torch

    torch.cuda.current_stream().synchronize()


def radii(params: HiiParams):
    n = lib().mxr_hii_radii(ctypes.byref(params), None, 0)
    buf = (ctypes.c_double * n)()
    lib().mxr_hii_radii(ctypes.byref(params), buf, n)
    return list(buf)


class Context:
    """One mxr_ctx: a device, a grid dimension, persistent buffers and tuned kernels."""

    def __init__(self, grid_dim: int, device: int = 0, path: int = MXR_PATH_FUSED):
        L = lib()
        h = ctypes.c_void_p()
        check(L.mxr_create(ctypes.byref(h), device, grid_dim), "mxr_create")
        self._h = h
        self.N = grid_dim
        self.set_path(path)

    def set_path(self, path: int):
        check(lib().mxr_set_path(self._h, path), "mxr_set_path")

    def find_hii_bubbles(self, params: HiiParams, **arrays) -> HiiGlobals:
        g = hii_grids(**arrays)
        if g.mem == MXR_MEM_DEVICE:
            _sync_torch()
        out = HiiGlobals()
        check(lib().mxr_find_hii_bubbles(self._h, ctypes.byref(params), ctypes.byref(g), ctypes.byref(out)),
              "mxr_find_hii_bubbles")
        return out

    def filter(self, box, local_ix_start: int, slab_nx: int, grid_dim: int, box_size: float, R: float,
               filter_type: int):
        mem = _mem_of([box])
        check(lib().mxr_filter(self._h, _ptr(box), mem, local_ix_start, slab_nx, grid_dim, box_size, R,
                               filter_type), "mxr_filter")

    def brightness_temperature(self, params: TbParams, xH, deltax, delta_T, TS_box=None) -> float:
        g = TbGrids()
        g.mem = _mem_of([xH, deltax, delta_T, TS_box])
        if g.mem == MXR_MEM_DEVICE:
            _sync_torch()
        g.xH, g.deltax, g.TS_box, g.delta_T = _ptr(xH), _ptr(deltax), _ptr(TS_box), _ptr(delta_T)
        ave = ctypes.c_double()
        check(lib().mxr_brightness_temperature(self._h, ctypes.byref(params), ctypes.byref(g), ctypes.byref(ave)),
              "mxr_brightness_temperature")
        return ave.value

    def profile(self, on: bool = True):
        check(lib().mxr_profile_enable(self._h, int(on)), "mxr_profile_enable")

    def kernel_stats(self):
        buf = (KernelStat * 32)()
        n = lib().mxr_profile_read(self._h, buf, 32)
        if n < 0:
            check(n, "mxr_profile_read")
        return {buf[i].name.decode(): dict(launches=buf[i].launches, total_ms=buf[i].total_ms,
                                           alg_bytes=buf[i].alg_bytes) for i in range(n) if buf[i].launches}

    def close(self):
        if self._h:
            lib().mxr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class HiiState:
    """Output / persistent grids of a find_HII_bubbles call (reference layouts)."""
    xH: object
    z_at_ionization: object
    J_21_at_ionization: object
    r_bubble: object
    temp_kinetic_all_gas: object

    @staticmethod
    def host(N: int) -> "HiiState":
        n = N ** 3
        return HiiState(np.ones(n, np.float32), np.full(n, -1.0, np.float32), np.zeros(n, np.float32),
                        np.zeros(n, np.float32), np.zeros(n, np.float32))

    def arrays(self):
        return dict(xH=self.xH, z_at_ionization=self.z_at_ionization, J_21_at_ionization=self.J_21_at_ionization,
                    r_bubble=self.r_bubble, temp_kinetic_all_gas=self.temp_kinetic_all_gas)
