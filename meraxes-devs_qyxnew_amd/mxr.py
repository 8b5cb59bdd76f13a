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
LIB_PATH = os.environ.get("MXR_LIB") or os.path.join(PKG_DIR, "libmxr.so")  # MXR_LIB: another build
TABLES_DIR = os.path.join(PKG_DIR, "tables")

MXR_OK = 0
MXR_ERR_INVALID_ARG = -1
MXR_ERR_UNSUPPORTED = -2
MXR_ERR_HIP = -3
MXR_ERR_OOM = -4
MXR_ERR_TABLES = -5
MXR_ERR_NO_DEVICE = -6
MXR_ERR_FILTER = -7
MXR_ERR_NO_WISDOM = -8

MXR_AGN_XRAY_NONE, MXR_AGN_XRAY_BOTH, MXR_AGN_XRAY_HARD, MXR_AGN_XRAY_SOFT = 0, 1, 2, 3
MXR_MEM_HOST = 0
MXR_MEM_DEVICE = 1
MXR_PATH_FUSED = 0
MXR_PATH_PLAIN = 1
MXR_PATH_FUSED_DENSE = 2
MXR_STORAGE_F32 = 0
MXR_STORAGE_BF16 = 1
MXR_BATCH_GRAPH = 0
MXR_BATCH_STREAMS = 1
MXR_BATCH_SERIAL = 2

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
        ("ReionAlphaUVBH", ctypes.c_double),
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
        ("Flag_TemperatureDependentRec", ctypes.c_int),
        ("Y_He", ctypes.c_double),
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
        ("storage", ctypes.c_int),
        ("effective_bhm", ctypes.c_void_p),
        ("effective_bhar", ctypes.c_void_p),
        ("N_rec", ctypes.c_void_p),
        ("Gamma12", ctypes.c_void_p),
        ("residual_xH", ctypes.c_void_p),
        ("clumping_factor", ctypes.c_void_p),
        ("t_resp", ctypes.c_void_p),
    ]


class Tuning(ctypes.Structure):
    """mxr_tuning: launch shapes of the fused path (include/mxr.h)."""
    _fields_ = [("zdirect_kmax", ctypes.c_int), ("zdirect_rpb", ctypes.c_int), ("live_tiles", ctypes.c_int),
                ("col_tc_x", ctypes.c_int), ("col_tc_y", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


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


MXR_GRID_DENSITY, MXR_GRID_X_VELOCITY, MXR_GRID_Y_VELOCITY, MXR_GRID_Z_VELOCITY = 0, 1, 2, 3


class GridFileInfo(ctypes.Structure):
    _fields_ = [("n_cell", ctypes.c_int * 3), ("box_size", ctypes.c_double * 3), ("n_grids", ctypes.c_int),
                ("ma_scheme", ctypes.c_int)]


class GridParams(ctypes.Structure):
    _fields_ = [("BoxSize", ctypes.c_double), ("Hubble_h", ctypes.c_double), ("NPart", ctypes.c_longlong),
                ("PartMass", ctypes.c_double), ("TreesID", ctypes.c_int)]


def grid_params(box: float, npart: int = 2160 ** 3, part_mass: float = 1.0, hubble_h: float | None = None,
                trees_id: int = 1) -> GridParams:
    p = GridParams()
    p.BoxSize = box
    p.Hubble_h = GENESIS["Hubble_h"] if hubble_h is None else hubble_h
    p.NPart = npart
    p.PartMass = part_mass
    p.TreesID = trees_id
    return p


def grid_file_info(fname: str) -> GridFileInfo:
    info = GridFileInfo()
    check(lib().mxr_grid_file_read_info(fname.encode(), ctypes.byref(info)), "mxr_grid_file_read_info")
    return info


class BaryonParams(ctypes.Structure):
    _fields_ = [
        ("ReionGridDim", ctypes.c_int),
        ("BoxSize", ctypes.c_double),
        ("Flag_IncludeSpinTemp", ctypes.c_int),
        ("Flag_BHFeedback", ctypes.c_int),
        ("Flag_InstantaneousSFR", ctypes.c_int),
        ("Flag_IncludeAGNXray", ctypes.c_int),
        ("BlackHoleMassLimitReion", ctypes.c_double),
        ("sfr_timescale", ctypes.c_double),
        ("NstoreSnapshots_Heating", ctypes.c_int),
        ("Flag_XrayMetallicity", ctypes.c_int),
        ("LXrayGal", ctypes.c_double),
        ("XrayLogLNorm", ctypes.c_double),
        ("XrayMetallicitySlope", ctypes.c_double),
        ("XrayMetallicitySolar", ctypes.c_double),
        ("XrayOxygenAbundanceFactor", ctypes.c_double),
        ("NuXrayThreshold", ctypes.c_double),
        ("NuXraySoftCut", ctypes.c_double),
        ("SpecIndexXrayGal", ctypes.c_double),
        ("UnitMass_in_g", ctypes.c_double),
        ("UnitTime_in_s", ctypes.c_double),
    ]


GALAXY_FIELDS = ("Pos", "Type", "FescWeightedGSM", "FescWeightedSfr", "Sfr", "GrossStellarMass", "ColdGas",
                 "MetalsColdGas", "BlackHoleMass", "EffectiveBHM", "EffectiveBHAR", "BHXrayEmissivity_hard",
                 "BHXrayEmissivity_soft")


class Galaxies(ctypes.Structure):
    _fields_ = [("mem", ctypes.c_int), ("n", ctypes.c_long)] + [(f, ctypes.c_void_p) for f in GALAXY_FIELDS]


BARYON_GRIDS = ("stars", "weighted_sfr", "effective_bhm", "effective_bhar", "sfr", "sfr_histories",
                "xray_luminosity", "xray_luminosity_histories", "BHXrayEmissivity_hard", "bh_xray_histories_hard",
                "BHXrayEmissivity_soft", "bh_xray_histories_soft")


class BaryonGrids(ctypes.Structure):
    _fields_ = [("mem", ctypes.c_int)] + [(f, ctypes.c_void_p) for f in BARYON_GRIDS] + [("history_head", ctypes.c_void_p)]


class TbParams(ctypes.Structure):
    _fields_ = [
        ("ReionGridDim", ctypes.c_int),
        ("BoxSize", ctypes.c_double),
        ("Hubble_h", ctypes.c_double),
        ("OmegaM", ctypes.c_double),
        ("OmegaLambda", ctypes.c_double),
        ("OmegaR", ctypes.c_double),
        ("BaryonFrac", ctypes.c_double),
        ("redshift", ctypes.c_double),
        ("Flag_IncludeSpinTemp", ctypes.c_int),
        ("Flag_IncludePecVelsFor21cm", ctypes.c_int),
        ("TreesID", ctypes.c_int),
    ]


class TbGrids(ctypes.Structure):
    _fields_ = [
        ("mem", ctypes.c_int),
        ("xH", ctypes.c_void_p),
        ("deltax", ctypes.c_void_p),
        ("TS_box", ctypes.c_void_p),
        ("vel", ctypes.c_void_p),
        ("delta_T", ctypes.c_void_p),
    ]


MXR_TS_MAX_STEPS = 64
MXR_TS_MAX_OPS = 16
MXR_X_INT_NXHII = 14
TS_OP_NAMES = {0: "clear", 1: "scale", 2: "load", 3: "add"}


class TsParams(ctypes.Structure):
    _fields_ = [
        ("ReionGridDim", ctypes.c_int),
        ("BoxSize", ctypes.c_double),
        ("Hubble_h", ctypes.c_double),
        ("OmegaM", ctypes.c_double),
        ("OmegaLambda", ctypes.c_double),
        ("OmegaK", ctypes.c_double),
        ("OmegaR", ctypes.c_double),
        ("wLambda", ctypes.c_double),
        ("BaryonFrac", ctypes.c_double),
        ("Y_He", ctypes.c_double),
        ("ReionRtoMFilterType", ctypes.c_int),
        ("TsNumFilterSteps", ctypes.c_int),
        ("TsHeatingFilterType", ctypes.c_int),
        ("ReionMaxHeatingRedshift", ctypes.c_double),
        ("ReionEfficiency", ctypes.c_double),
        ("ReionNionPhotPerBary", ctypes.c_double),
        ("NuXrayThreshold", ctypes.c_double),
        ("NuXraySoftCut", ctypes.c_double),
        ("NuXrayMax", ctypes.c_double),
        ("SpecIndexXrayGal", ctypes.c_double),
        ("Flag_IncludeAGNXray", ctypes.c_int),
        ("SpecIndexXrayAGNSoft", ctypes.c_double),
        ("SpecIndexXrayAGNHard", ctypes.c_double),
        ("UnitLength_in_cm", ctypes.c_double),
        ("UnitMass_in_g", ctypes.c_double),
        ("UnitVelocity_in_cm_per_s", ctypes.c_double),
        ("snapshot", ctypes.c_int),
        ("ZZ", ctypes.POINTER(ctypes.c_double)),
        ("TablesForXHeatingDir", ctypes.c_char_p),
    ]


class TsGrids(ctypes.Structure):
    _fields_ = [
        ("mem", ctypes.c_int),
        ("deltax", ctypes.c_void_p),
        ("stars", ctypes.c_void_p),
        ("sfr_histories", ctypes.c_void_p),
        ("xray_luminosity_histories", ctypes.c_void_p),
        ("bh_xray_histories_hard", ctypes.c_void_p),
        ("bh_xray_histories_soft", ctypes.c_void_p),
        ("n_histories", ctypes.c_int),
        ("x_e_box_prev", ctypes.c_void_p),
        ("x_e_box", ctypes.c_void_p),
        ("Tk_box", ctypes.c_void_p),
        ("TS_box", ctypes.c_void_p),
        ("stored_fcoll", ctypes.c_void_p),
        ("history_head", ctypes.c_int),
    ]


_S, _K, _X = MXR_TS_MAX_STEPS, MXR_TS_MAX_OPS, MXR_X_INT_NXHII


class TsTables(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int),
        ("homogeneous", ctypes.c_int),
        ("NO_LIGHT", ctypes.c_int),
        ("zp", ctypes.c_double),
        ("dzp", ctypes.c_double),
        ("dt_dzp", ctypes.c_double),
        ("growth_factor_zp", ctypes.c_double),
        ("dgrowth_factor_dzp", ctypes.c_double),
        ("const_zp_prefactor_GAL", ctypes.c_double),
        ("collapse_fraction", ctypes.c_double),
        ("x_e_ave", ctypes.c_double),
        ("filling_factor_of_HI_zp", ctypes.c_double),
        ("R", ctypes.c_double * _S),
        ("zpp_edge", ctypes.c_double * _S),
        ("zpp", ctypes.c_double * _S),
        ("nu_tau_one", ctypes.c_double * _S),
        ("sum_lyn", ctypes.c_double * _S),
        ("freq_int_heat", (ctypes.c_double * _S) * _X),
        ("freq_int_ion", (ctypes.c_double * _S) * _X),
        ("freq_int_lya", (ctypes.c_double * _S) * _X),
        ("const_zp_prefactor_AGN_soft", ctypes.c_double),
        ("const_zp_prefactor_AGN_hard", ctypes.c_double),
        ("freq_int_heat_AGN_soft", (ctypes.c_double * _S) * _X),
        ("freq_int_ion_AGN_soft", (ctypes.c_double * _S) * _X),
        ("freq_int_lya_AGN_soft", (ctypes.c_double * _S) * _X),
        ("freq_int_heat_AGN_hard", (ctypes.c_double * _S) * _X),
        ("freq_int_ion_AGN_hard", (ctypes.c_double * _S) * _X),
        ("freq_int_lya_AGN_hard", (ctypes.c_double * _S) * _X),
        ("x_int_XHII", ctypes.c_float * _X),
        ("n_ops", ctypes.c_int * _S),
        ("op_kind", (ctypes.c_int * _K) * _S),
        ("op_hist", (ctypes.c_int * _K) * _S),
        ("op_weight", (ctypes.c_double * _K) * _S),
        ("hist_max", ctypes.c_int),
    ]

    def arr(self, name):
        """numpy view of a field, trimmed to the n shells."""
        a = np.ctypeslib.as_array(getattr(self, name))
        if name in ("R", "zpp_edge", "zpp", "nu_tau_one", "sum_lyn"):
            return a[: self.n].copy()
        if name.startswith("freq_int"):
            return a[:, : self.n].copy()
        return a.copy()

    def ops(self):
        """[[(kind_name, hist, weight), ...] per shell]"""
        return [[(TS_OP_NAMES[self.op_kind[r][k]], self.op_hist[r][k], self.op_weight[r][k])
                 for k in range(self.n_ops[r])] for r in range(self.n)]


class TsGlobals(ctypes.Structure):
    _fields_ = [
        ("volume_ave_J_alpha", ctypes.c_double),
        ("volume_ave_xalpha", ctypes.c_double),
        ("volume_ave_Xheat", ctypes.c_double),
        ("volume_ave_Xion", ctypes.c_double),
        ("volume_ave_TS", ctypes.c_double),
        ("volume_ave_TK", ctypes.c_double),
        ("volume_ave_xe", ctypes.c_double),
        ("XrayEmissivity_HMXB", ctypes.c_double),
        ("XrayEmissivity_hard", ctypes.c_double),
        ("XrayEmissivity_soft", ctypes.c_double),
        ("volume_ave_Xheat_AGN_soft", ctypes.c_double),
        ("volume_ave_Xheat_AGN_hard", ctypes.c_double),
        ("device_ms", ctypes.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


# ---------------------------------------------------------------------------------------
# Parameter sets.  Cosmology and units: Genesis L10_N512
# (input/params/simulations/Genesis_L10_N512.par:17-31); reionisation parameters from
# input/params/defaults.par with Flag_EvolvingReionRBubbleMax = 0 and R_max = 50 Mpc/h.

GENESIS = dict(Hubble_h=0.6751, OmegaM=0.3121, OmegaLambda=0.6879, OmegaK=0.0, BaryonFrac=0.15714,
               UnitLength_in_cm=3.08568e24, UnitMass_in_g=1.989e43, UnitVelocity_in_cm_per_s=1e5)


def hii_params(N: int, box: float, filter_type: int = 1, zeta: float = 1.0, redshift: float = 8.0,
               spin_temp: bool = False, tables_dir: str | None = None, bh: bool = False, **over) -> HiiParams:
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
    p.Flag_BHFeedback = int(bh)
    p.ReionSfrTimescale = 0.5
    p.ReionEfficiency = zeta
    p.ReionNionPhotPerBary = 4000.0
    p.ReionGammaHaloBias = 2.0
    p.ReionAlphaUV = 2.0
    p.ReionAlphaUVBH = 2.0  # input/params/defaults.par:250
    p.ReionRBubbleMax = 50.0
    p.ReionRBubbleMaxRecomb = 50.0
    p.ReionRBubbleMin = 0.4068
    p.redshift = redshift
    p.prev_redshift = redshift
    p.TK_recfast = -1.0
    p.TablesForXHeatingDir = (tables_dir or TABLES_DIR).encode()
    p.Flag_TemperatureDependentRec = 0
    p.Y_He = 0.24  # input/params/defaults.par
    for k, v in over.items():
        setattr(p, k, v)
    return p


MXR_TREES_VELOCIRAPTOR = 0
MXR_TREES_GBPTREES = 1
MXR_TREES_VELOCIRAPTOR_AUG = 2


def tb_params(N: int, redshift: float = 8.0, spin_temp: bool = False, box: float = 100.0, pecvel: int = 0,
              trees_id: int = MXR_TREES_VELOCIRAPTOR) -> TbParams:
    p = TbParams()
    p.ReionGridDim = N
    p.BoxSize = box
    p.Hubble_h = GENESIS["Hubble_h"]
    p.OmegaM = GENESIS["OmegaM"]
    p.OmegaLambda = GENESIS["OmegaLambda"]
    p.OmegaR = 0.0
    p.BaryonFrac = GENESIS["BaryonFrac"]
    p.redshift = redshift
    p.Flag_IncludeSpinTemp = int(spin_temp)
    p.Flag_IncludePecVelsFor21cm = pecvel
    p.TreesID = trees_id
    return p


def baryon_params(N: int, box: float, spin_temp: bool = False, bh: bool = False, agn_xray: int = 0,
                  instantaneous_sfr: bool = True, xray_metallicity: bool = False, nstore: int = 1,
                  sfr_timescale: float = 0.05, **over) -> BaryonParams:
    """construct_baryon_grids parameters with input/params/defaults.par values (LXrayGal 3.16e40,
    Kyritsis et al. X-ray metallicity terms, NuXray 500/2000 eV, alpha 1, Genesis units)."""
    p = BaryonParams()
    p.ReionGridDim = N
    p.BoxSize = box
    p.Flag_IncludeSpinTemp = int(spin_temp)
    p.Flag_BHFeedback = int(bh)
    p.Flag_InstantaneousSFR = int(instantaneous_sfr)
    p.Flag_IncludeAGNXray = agn_xray
    p.BlackHoleMassLimitReion = -1.0
    p.sfr_timescale = sfr_timescale
    p.NstoreSnapshots_Heating = nstore
    p.Flag_XrayMetallicity = int(xray_metallicity)
    p.LXrayGal = 3.16e40
    p.XrayLogLNorm = 40.68
    p.XrayMetallicitySlope = -1.22
    p.XrayMetallicitySolar = 0.0134
    p.XrayOxygenAbundanceFactor = 1.0
    p.NuXrayThreshold = 500.0
    p.NuXraySoftCut = 2000.0
    p.SpecIndexXrayGal = 1.0
    p.UnitMass_in_g = GENESIS["UnitMass_in_g"]
    p.UnitTime_in_s = GENESIS["UnitLength_in_cm"] / GENESIS["UnitVelocity_in_cm_per_s"]
    for k, v in over.items():
        setattr(p, k, v)
    return p


# Spin-temperature defaults: input/params/defaults.par:51-68, 245, 274-282; OmegaR 0 and
# wLambda -1 as in input/params/simulations/Genesis_*.par.
TS_DEFAULTS = dict(OmegaR=0.0, wLambda=-1.0, Y_He=0.24, ReionRtoMFilterType=0, TsNumFilterSteps=40,
                   TsHeatingFilterType=1, ReionMaxHeatingRedshift=20.0, ReionEfficiency=1.0,
                   ReionNionPhotPerBary=4000.0, NuXrayThreshold=500.0, NuXraySoftCut=2000.0, NuXrayMax=10000.0,
                   SpecIndexXrayGal=1.0, Flag_IncludeAGNXray=0, SpecIndexXrayAGNSoft=2.2, SpecIndexXrayAGNHard=1.7)


def snapshot_redshifts(z_hi: float = 20.0, z_lo: float = 8.0, n: int = 41) -> np.ndarray:
    """Snapshot redshifts linear in z, z_hi first (snapshot 0) down to z_lo."""
    return np.linspace(z_hi, z_lo, n).astype(np.float64)


def ts_params(N: int, box: float, ZZ: np.ndarray, snapshot: int, tables_dir: str | None = None, **over) -> TsParams:
    p = TsParams()
    p.ReionGridDim = N
    p.BoxSize = box
    for k in ("Hubble_h", "OmegaM", "OmegaLambda", "OmegaK", "BaryonFrac", "UnitLength_in_cm", "UnitMass_in_g",
              "UnitVelocity_in_cm_per_s"):
        setattr(p, k, GENESIS[k])
    for k, v in TS_DEFAULTS.items():
        setattr(p, k, v)
    zz = np.ascontiguousarray(ZZ, dtype=np.float64)
    p._zz = zz  # keeps the array alive with the struct
    p.ZZ = zz.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    p.snapshot = snapshot
    p.TablesForXHeatingDir = (tables_dir or TABLES_DIR).encode()
    for k, v in over.items():
        setattr(p, k, v)
    return p


def padded_len(N: int) -> int:
    return N * N * 2 * (N // 2 + 1)


# ---------------------------------------------------------------------------------------
class MxrError(RuntimeError):
    def __init__(self, code: int, where: str, detail: str = ""):
        self.code = code
        super().__init__(f"{where}: {code} {detail}")


_lib = None


def _bind_torch_runtime():
    """Import torch (when installed) before libmxr loads.

    torch ships its own HIP and HSA runtime libraries under the same sonames as the system's
    ones that libmxr links.  Loaded after torch, libmxr binds to the copies already in the
    process, so the process holds one runtime and device pointers pass freely between torch
    and libmxr.  Loaded the other way round, the process would hold two runtimes, and only the
    first to initialise would see the GPU.  Without torch this does nothing."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """Load libmxr.so (raises if it was not built: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MxrError(MXR_ERR_NO_DEVICE, "libmxr.so", f"not built at {LIB_PATH}")
    _bind_torch_runtime()
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
    L.mxr_get_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
    L.mxr_get_tuning.restype = ctypes.c_int
    L.mxr_set_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
    L.mxr_set_tuning.restype = ctypes.c_int
    L.mxr_find_hii_bubbles.argtypes = [vp, ctypes.POINTER(HiiParams), ctypes.POINTER(HiiGrids),
                                       ctypes.POINTER(HiiGlobals)]
    L.mxr_find_hii_bubbles.restype = ctypes.c_int
    _dp = ctypes.POINTER(ctypes.c_double)
    L.mxr_recombination_tables.argtypes = [vp, ctypes.POINTER(HiiParams), _dp, _dp, _dp, _dp]
    L.mxr_recombination_tables.restype = ctypes.c_int
    L.mxr_filter.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                             ctypes.c_float, ctypes.c_int]
    L.mxr_filter.restype = ctypes.c_int
    L.mxr_RtoM.argtypes = [ctypes.POINTER(HiiParams), ctypes.c_double]
    L.mxr_RtoM.restype = ctypes.c_double
    L.mxr_hii_radii.argtypes = [ctypes.POINTER(HiiParams), ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    L.mxr_hii_radii.restype = ctypes.c_int
    L.mxr_velocity_gradient.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]
    L.mxr_velocity_gradient.restype = ctypes.c_int
    L.mxr_brightness_temperature.argtypes = [vp, ctypes.POINTER(TbParams), ctypes.POINTER(TbGrids),
                                             ctypes.POINTER(ctypes.c_double)]
    L.mxr_brightness_temperature.restype = ctypes.c_int
    L.mxr_brightness_temperature_noalloc.argtypes = [ctypes.POINTER(TbParams), ctypes.POINTER(TbGrids), vp, vp]
    L.mxr_brightness_temperature_noalloc.restype = ctypes.c_int
    L.mxr_supported_dims.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    L.mxr_supported_dims.restype = ctypes.c_int
    L.mxr_ps_length.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
    L.mxr_ps_length.restype = ctypes.c_int
    fpp = ctypes.POINTER(ctypes.c_float)
    L.mxr_power_spectrum.argtypes = [vp, vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, fpp, fpp,
                                     fpp]
    L.mxr_power_spectrum.restype = ctypes.c_int
    L.mxr_ps_counts.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    L.mxr_ps_counts.restype = ctypes.c_int
    L.mxr_grid_file_read_info.argtypes = [ctypes.c_char_p, ctypes.POINTER(GridFileInfo)]
    L.mxr_grid_file_read_info.restype = ctypes.c_int
    L.mxr_prepare_grid.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                   ctypes.POINTER(GridParams), vp, ctypes.c_int]
    L.mxr_prepare_grid.restype = ctypes.c_int
    L.mxr_read_grid_gbptrees.argtypes = [vp, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(GridParams), vp, ctypes.c_int]
    L.mxr_read_grid_gbptrees.restype = ctypes.c_int
    L.mxr_construct_baryon_grids.argtypes = [vp, ctypes.POINTER(BaryonParams), ctypes.POINTER(Galaxies),
                                             ctypes.POINTER(BaryonGrids), ctypes.POINTER(ctypes.c_long)]
    L.mxr_construct_baryon_grids.restype = ctypes.c_int
    L.mxr_compute_ts.argtypes = [vp, ctypes.POINTER(TsParams), ctypes.POINTER(TsGrids), ctypes.POINTER(TsGlobals),
                                 ctypes.POINTER(TsTables)]
    L.mxr_compute_ts.restype = ctypes.c_int
    L.mxr_ts_build_tables.argtypes = [ctypes.POINTER(TsParams), ctypes.c_double, ctypes.c_double,
                                      ctypes.POINTER(ctypes.c_double), ctypes.POINTER(TsTables)]
    L.mxr_ts_build_tables.restype = ctypes.c_int
    L.mxr_ts_history_depth.argtypes = [ctypes.POINTER(TsParams), ctypes.c_int]
    L.mxr_ts_history_depth.restype = ctypes.c_int
    L.mxr_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.mxr_batch_create.restype = ctypes.c_int
    L.mxr_batch_destroy.argtypes = [vp]
    L.mxr_batch_destroy.restype = None
    L.mxr_find_hii_bubbles_batched.argtypes = [vp, ctypes.POINTER(HiiParams), ctypes.POINTER(HiiGrids),
                                               ctypes.POINTER(HiiGlobals), ctypes.c_int]
    L.mxr_find_hii_bubbles_batched.restype = ctypes.c_int
    cp = ctypes.c_char_p
    L.mxr_autotune.argtypes = [vp, ctypes.POINTER(HiiParams), ctypes.POINTER(HiiGrids), cp, ctypes.POINTER(Tuning)]
    L.mxr_autotune.restype = ctypes.c_int
    L.mxr_load_wisdom.argtypes = [vp, ctypes.POINTER(HiiParams), cp]
    L.mxr_load_wisdom.restype = ctypes.c_int
    L.mxr_wisdom_lookup.argtypes = [cp, cp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(Tuning)]
    L.mxr_wisdom_lookup.restype = ctypes.c_int
    L.mxr_wisdom_store.argtypes = [cp, cp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(Tuning), ctypes.c_double]
    L.mxr_wisdom_store.restype = ctypes.c_int
    L.mxr_batch_autotune.argtypes = [vp, ctypes.POINTER(HiiParams), ctypes.POINTER(HiiGrids), cp]
    L.mxr_batch_autotune.restype = ctypes.c_int
    L.mxr_batch_load_wisdom.argtypes = [vp, ctypes.POINTER(HiiParams), cp]
    L.mxr_batch_load_wisdom.restype = ctypes.c_int
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


HII_IN = ("deltax", "stars", "weighted_sfr", "x_e_box", "Tk_box", "effective_bhm", "effective_bhar")
HII_OUT = ("xH", "z_at_ionization", "J_21_at_ionization", "r_bubble", "temp_kinetic_all_gas")
HII_REC = ("N_rec", "Gamma12", "residual_xH", "clumping_factor", "t_resp")  # Flag_IncludeRecombinations


def _is_bf16(a):
    return a is not None and not isinstance(a, np.ndarray) and str(getattr(a, "dtype", "")) == "torch.bfloat16"


def hii_grids(**arrays) -> HiiGrids:
    """Grid pointers; stars and weighted_sfr given as torch.bfloat16 (or numpy uint16 bit
    patterns with storage=MXR_STORAGE_BF16) select the BF16-storage variant."""
    storage = arrays.pop("storage", None)
    g = HiiGrids()
    g.mem = _mem_of(list(arrays.values()))
    for name in HII_IN + HII_OUT + HII_REC:
        setattr(g, name, _ptr(arrays.get(name)))
    if storage is None:
        storage = MXR_STORAGE_BF16 if _is_bf16(arrays.get("stars")) else MXR_STORAGE_F32
    g.storage = storage
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


def ts_history_depth(params: TsParams, last_snapshot: int) -> int:
    n = lib().mxr_ts_history_depth(ctypes.byref(params), last_snapshot)
    if n < 0:
        check(n, "mxr_ts_history_depth")
    return n


def ts_build_tables(params: TsParams, collapse_fraction: float, x_e_ave: float, stored_fcoll: np.ndarray) -> TsTables:
    """Host-only tables of one spin-temperature step (no GPU needed)."""
    sf = np.ascontiguousarray(stored_fcoll, dtype=np.float64)
    t = TsTables()
    check(lib().mxr_ts_build_tables(ctypes.byref(params), collapse_fraction, x_e_ave,
                                    sf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(t)),
          "mxr_ts_build_tables")
    return t


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

    def tuning(self) -> "Tuning":
        t = Tuning()
        check(lib().mxr_get_tuning(self._h, ctypes.byref(t)), "mxr_get_tuning")
        return t

    def set_tuning(self, **fields) -> "Tuning":
        """Change the named launch-shape fields (mxr_tuning); returns the previous setting."""
        old = self.tuning()
        t = self.tuning()
        for k, v in fields.items():
            if k not in dict(Tuning._fields_):
                raise KeyError(k)
            setattr(t, k, v)
        check(lib().mxr_set_tuning(self._h, ctypes.byref(t)), "mxr_set_tuning")
        return old

    def restore_tuning(self, t: "Tuning"):
        check(lib().mxr_set_tuning(self._h, ctypes.byref(t)), "mxr_set_tuning")

    def autotune(self, params: HiiParams, wisdom: str | None = None, **arrays) -> "Tuning":
        """mxr_autotune on these input grids (state grids are not needed or touched)."""
        g = hii_grids(**_with_dummy_state(arrays))
        if g.mem == MXR_MEM_DEVICE:
            _sync_torch()
        t = Tuning()
        check(lib().mxr_autotune(self._h, ctypes.byref(params), ctypes.byref(g), _path(wisdom), ctypes.byref(t)),
              "mxr_autotune")
        return t

    def load_wisdom(self, params: HiiParams, wisdom: str) -> int:
        """mxr_load_wisdom; returns MXR_OK or MXR_ERR_NO_WISDOM (other errors raise)."""
        rc = lib().mxr_load_wisdom(self._h, ctypes.byref(params), _path(wisdom))
        if rc not in (MXR_OK, MXR_ERR_NO_WISDOM):
            check(rc, "mxr_load_wisdom")
        return rc

    def find_hii_bubbles(self, params: HiiParams, **arrays) -> HiiGlobals:
        g = hii_grids(**arrays)
        if g.mem == MXR_MEM_DEVICE:
            _sync_torch()
        out = HiiGlobals()
        check(lib().mxr_find_hii_bubbles(self._h, ctypes.byref(params), ctypes.byref(g), ctypes.byref(out)),
              "mxr_find_hii_bubbles")
        return out

    def recombination_tables(self, params: HiiParams):
        """mxr_recombination_tables: (lnGamma[150], ln_rate, clumping, ln_residual), the tables as
        float64 arrays of shape (200, 100, 150) (z_ct, t_ct, ln Gamma_12 node)."""
        x = np.zeros(150, np.float64)
        tabs = [np.zeros((200, 100, 150), np.float64) for _ in range(3)]
        dp = ctypes.POINTER(ctypes.c_double)
        check(lib().mxr_recombination_tables(self._h, ctypes.byref(params), x.ctypes.data_as(dp),
                                             *[t.ctypes.data_as(dp) for t in tabs]), "mxr_recombination_tables")
        return (x, *tabs)

    def filter(self, box, local_ix_start: int, slab_nx: int, grid_dim: int, box_size: float, R: float,
               filter_type: int):
        mem = _mem_of([box])
        check(lib().mxr_filter(self._h, _ptr(box), mem, local_ix_start, slab_nx, grid_dim, box_size, R,
                               filter_type), "mxr_filter")

    def brightness_temperature(self, params: TbParams, xH, deltax, delta_T, TS_box=None, vel=None) -> float:
        """mxr_brightness_temperature; `vel` (padded, with PecVels > 0) is converted in place."""
        g = TbGrids()
        g.mem = _mem_of([xH, deltax, delta_T, TS_box, vel])
        if g.mem == MXR_MEM_DEVICE:
            _sync_torch()
        g.xH, g.deltax, g.TS_box, g.delta_T = _ptr(xH), _ptr(deltax), _ptr(TS_box), _ptr(delta_T)
        g.vel = _ptr(vel)
        ave = ctypes.c_double()
        check(lib().mxr_brightness_temperature(self._h, ctypes.byref(params), ctypes.byref(g), ctypes.byref(ave)),
              "mxr_brightness_temperature")
        return ave.value

    def velocity_gradient(self, box, slab_nx: int, grid_dim: int, box_size: float):
        """mxr_velocity_gradient on a complex64 half spectrum (numpy or torch, in place)."""
        mem = _mem_of([box])
        if mem == MXR_MEM_DEVICE:
            _sync_torch()
        check(lib().mxr_velocity_gradient(self._h, _ptr(box), mem, slab_nx, grid_dim, box_size),
              "mxr_velocity_gradient")

    def power_spectrum(self, delta_T, box_size: float, hubble_h: float):
        """mxr_power_spectrum of a real grid (numpy or torch).  Returns (PS_k, PS_data, PS_error) as
        float32 arrays of ps_length() bins."""
        mem = _mem_of([delta_T])
        if mem == MXR_MEM_DEVICE:
            _sync_torch()
        nb = ps_length(self.N, box_size, hubble_h)
        out = [np.zeros(nb, np.float32) for _ in range(3)]
        fpp = ctypes.POINTER(ctypes.c_float)
        check(lib().mxr_power_spectrum(self._h, _ptr(delta_T), mem, box_size, hubble_h, nb,
                                       *[o.ctypes.data_as(fpp) for o in out]), "mxr_power_spectrum")
        return tuple(out)

    def ps_counts(self, box_size: float, hubble_h: float) -> np.ndarray:
        """mxr_ps_counts: modes per power-spectrum bin (float64)."""
        nb = ps_length(self.N, box_size, hubble_h)
        out = np.zeros(nb, np.float64)
        check(lib().mxr_ps_counts(self._h, box_size, hubble_h, nb, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
              "mxr_ps_counts")
        return out

    def prepare_grid(self, hi, n_hi: int, prop: int, file_box: float, params: GridParams, slab):
        """mxr_prepare_grid: hi is the n_hi^3 real-layout grid, slab the padded N^3 output (numpy or
        torch each, written in place)."""
        mi, mo = _mem_of([hi]), _mem_of([slab])
        if MXR_MEM_DEVICE in (mi, mo):
            _sync_torch()
        check(lib().mxr_prepare_grid(self._h, _ptr(hi), n_hi, mi, prop, file_box, ctypes.byref(params), _ptr(slab), mo),
              "mxr_prepare_grid")

    def read_grid_gbptrees(self, fname: str, prop: int, params: GridParams, slab):
        """mxr_read_grid_gbptrees into the padded N^3 slab (numpy or torch)."""
        mo = _mem_of([slab])
        if mo == MXR_MEM_DEVICE:
            _sync_torch()
        check(lib().mxr_read_grid_gbptrees(self._h, fname.encode(), prop, ctypes.byref(params), _ptr(slab), mo),
              "mxr_read_grid_gbptrees")

    def construct_baryon_grids(self, params: BaryonParams, galaxies: dict, grids: dict, history_head=None) -> int:
        """mxr_construct_baryon_grids.  galaxies maps GALAXY_FIELDS names to arrays (numpy or torch,
        Pos as (n, 3) float32, Type int32, the rest float64); grids maps BARYON_GRIDS names to
        padded float32 arrays (histories NstoreSnapshots_Heating grids long), written in place.
        history_head: None for the reference's shifted histories, or a one-element int32 numpy
        array holding the ring's head, updated in place.
        Returns the number of live galaxies below BlackHoleMassLimitReion."""
        g = Galaxies()
        g.mem = _mem_of([v for v in galaxies.values() if v is not None])
        g.n = len(galaxies["Type"])
        for f in GALAXY_FIELDS:
            setattr(g, f, _ptr(galaxies.get(f)))
        o = BaryonGrids()
        o.mem = _mem_of([v for v in grids.values() if v is not None])
        for f in BARYON_GRIDS:
            setattr(o, f, _ptr(grids.get(f)))
        if history_head is not None:
            assert isinstance(history_head, np.ndarray) and history_head.dtype == np.int32 and history_head.size == 1
            o.history_head = history_head.ctypes.data
        if MXR_MEM_DEVICE in (g.mem, o.mem):
            _sync_torch()
        below = ctypes.c_long()
        check(lib().mxr_construct_baryon_grids(self._h, ctypes.byref(params), ctypes.byref(g), ctypes.byref(o),
                                               ctypes.byref(below)), "mxr_construct_baryon_grids")
        return below.value

    def compute_ts(self, params: TsParams, deltax, stars, sfr_histories, xray_luminosity_histories, n_histories: int,
                   x_e_box_prev, x_e_box, Tk_box, TS_box, stored_fcoll: np.ndarray, want_tables: bool = True,
                   history_head: int = 0, bh_xray_histories_soft=None, bh_xray_histories_hard=None):
        """One ComputeTs step.  Returns (TsGlobals, TsTables or None); stored_fcoll (host float64, one
        entry per snapshot) is updated in place.  history_head: ring slot of the newest history grid.
        The AGN histories (same layout as the others) are read as params.Flag_IncludeAGNXray asks."""
        g = TsGrids()
        bh = [a for a in (bh_xray_histories_soft, bh_xray_histories_hard) if a is not None]
        g.mem = _mem_of([deltax, stars, sfr_histories, xray_luminosity_histories, x_e_box_prev, x_e_box, Tk_box,
                         TS_box] + bh)
        if g.mem == MXR_MEM_DEVICE:
            _sync_torch()
        g.deltax, g.stars = _ptr(deltax), _ptr(stars)
        g.sfr_histories, g.xray_luminosity_histories = _ptr(sfr_histories), _ptr(xray_luminosity_histories)
        g.bh_xray_histories_soft = _ptr(bh_xray_histories_soft) if bh_xray_histories_soft is not None else None
        g.bh_xray_histories_hard = _ptr(bh_xray_histories_hard) if bh_xray_histories_hard is not None else None
        g.n_histories = n_histories
        g.x_e_box_prev, g.x_e_box, g.Tk_box, g.TS_box = _ptr(x_e_box_prev), _ptr(x_e_box), _ptr(Tk_box), _ptr(TS_box)
        assert isinstance(stored_fcoll, np.ndarray) and stored_fcoll.dtype == np.float64
        g.stored_fcoll = stored_fcoll.ctypes.data
        g.history_head = history_head
        out = TsGlobals()
        tab = TsTables() if want_tables else None
        check(lib().mxr_compute_ts(self._h, ctypes.byref(params), ctypes.byref(g), ctypes.byref(out),
                                   ctypes.byref(tab) if tab is not None else None), "mxr_compute_ts")
        return out, tab

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


def _path(p):
    return None if p is None else os.fsencode(p)


def _with_dummy_state(arrays: dict) -> dict:
    """Input grids plus placeholder state pointers for calls that never touch the state (the
    autotuner runs on its own scratch state)."""
    a = dict(arrays)
    ref = a["deltax"]
    for k in ("xH", "z_at_ionization", "J_21_at_ionization", "r_bubble", "temp_kinetic_all_gas"):
        a.setdefault(k, ref)
    return a


def brightness_temperature_noalloc(params: TbParams, xH, deltax, delta_T, ave, TS_box=None, stream=None):
    """mxr_brightness_temperature_noalloc on device tensors: delta_T and the one-element FP64
    tensor `ave` are written on `stream` (a hipStream_t value; None for the null stream),
    without synchronising."""
    g = TbGrids()
    g.mem = _mem_of([xH, deltax, delta_T, ave, TS_box])  # host memory is refused by the library
    g.xH, g.deltax, g.TS_box, g.delta_T = _ptr(xH), _ptr(deltax), _ptr(TS_box), _ptr(delta_T)
    g.vel = None
    check(lib().mxr_brightness_temperature_noalloc(ctypes.byref(params), ctypes.byref(g), _ptr(ave), stream),
          "mxr_brightness_temperature_noalloc")


def supported_dims() -> list:
    """mxr_supported_dims: the grid dimensions this build compiles."""
    n = lib().mxr_supported_dims(None, 0)
    out = (ctypes.c_int * n)()
    lib().mxr_supported_dims(out, n)
    return list(out)


def ps_length(grid_dim: int, box_size: float, hubble_h: float) -> int:
    """mxr_ps_length: the number of k bins of the power spectrum."""
    n = lib().mxr_ps_length(grid_dim, box_size, hubble_h)
    if n < 0:
        check(n, "mxr_ps_length")
    return n


def wisdom_lookup(path: str, arch: str, cus: int, grid_dim: int, filter_type: int, spin: int = 0):
    """The tuning a wisdom file holds for this key, or None."""
    t = Tuning()
    rc = lib().mxr_wisdom_lookup(_path(path), arch.encode(), cus, grid_dim, filter_type, spin, ctypes.byref(t))
    if rc == MXR_ERR_NO_WISDOM:
        return None
    check(rc, "mxr_wisdom_lookup")
    return t


def wisdom_store(path: str, arch: str, cus: int, grid_dim: int, filter_type: int, spin: int, tuning: "Tuning",
                 ms: float = 0.0):
    check(lib().mxr_wisdom_store(_path(path), arch.encode(), cus, grid_dim, filter_type, spin, ctypes.byref(tuning),
                                 ms), "mxr_wisdom_store")


class Batch:
    """A batch of independent find_HII_bubbles realisations (grid dimensions may differ), run as
    one captured hipGraph (MXR_BATCH_GRAPH), on one stream each (MXR_BATCH_STREAMS) or one at a
    time (MXR_BATCH_SERIAL).  Grids must be device (torch cuda) tensors."""

    def __init__(self, grid_dims, device: int = 0):
        L = lib()
        h = ctypes.c_void_p()
        dims = (ctypes.c_int * len(grid_dims))(*grid_dims)
        check(L.mxr_batch_create(ctypes.byref(h), device, len(grid_dims), dims), "mxr_batch_create")
        self._h = h
        self.dims = list(grid_dims)

    def find_hii_bubbles(self, params_list, arrays_list, mode: int = MXR_BATCH_GRAPH):
        n = len(self.dims)
        assert len(params_list) == n and len(arrays_list) == n
        P = (HiiParams * n)(*params_list)
        G = (HiiGrids * n)(*[hii_grids(**a) for a in arrays_list])
        out = (HiiGlobals * n)()
        _sync_torch()
        check(lib().mxr_find_hii_bubbles_batched(self._h, P, G, out, mode), "mxr_find_hii_bubbles_batched")
        return list(out)

    def autotune(self, params_list, arrays_list, wisdom: str | None = None):
        n = len(self.dims)
        P = (HiiParams * n)(*params_list)
        G = (HiiGrids * n)(*[hii_grids(**_with_dummy_state(a)) for a in arrays_list])
        _sync_torch()
        check(lib().mxr_batch_autotune(self._h, P, G, _path(wisdom)), "mxr_batch_autotune")

    def load_wisdom(self, params_list, wisdom: str) -> int:
        n = len(self.dims)
        P = (HiiParams * n)(*params_list)
        rc = lib().mxr_batch_load_wisdom(self._h, P, _path(wisdom))
        if rc not in (MXR_OK, MXR_ERR_NO_WISDOM):
            check(rc, "mxr_batch_load_wisdom")
        return rc

    def close(self):
        if self._h:
            lib().mxr_batch_destroy(self._h)
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
    # Flag_IncludeRecombinations only: N_rec (padded, persists, starts 0), Gamma12 (persists,
    # starts 0), residual_xH, clumping_factor, t_resp
    N_rec: object = None
    Gamma12: object = None
    residual_xH: object = None
    clumping_factor: object = None
    t_resp: object = None

    @staticmethod
    def host(N: int, recombinations: bool = False) -> "HiiState":
        n = N ** 3
        st = HiiState(np.ones(n, np.float32), np.full(n, -1.0, np.float32), np.zeros(n, np.float32),
                      np.zeros(n, np.float32), np.zeros(n, np.float32))
        if recombinations:
            st.N_rec = np.zeros(N * N * 2 * (N // 2 + 1), np.float32)
            st.Gamma12, st.residual_xH, st.clumping_factor, st.t_resp = (np.zeros(n, np.float32) for _ in range(4))
        return st

    def arrays(self):
        d = dict(xH=self.xH, z_at_ionization=self.z_at_ionization, J_21_at_ionization=self.J_21_at_ionization,
                 r_bubble=self.r_bubble, temp_kinetic_all_gas=self.temp_kinetic_all_gas)
        if self.N_rec is not None:
            d.update(N_rec=self.N_rec, Gamma12=self.Gamma12, residual_xH=self.residual_xH,
                     clumping_factor=self.clumping_factor, t_resp=self.t_resp)
        return d
