// mxr_ts.hpp -- shared definitions of the spin-temperature step (host tables <-> kernels).
// Not part of the public ABI (include/mxr.h is).
#pragma once

#include "mxr_internal.hpp"

namespace mxr {

constexpr int TS_MAXF = MXR_TS_MAX_STEPS;  // shells
constexpr int TS_MAX_OPS = MXR_TS_MAX_OPS; // source-grid operations before one shell's forward FFT
constexpr int XNH = MXR_X_INT_NXHII;       // x_e bins of the secondary-electron tables
constexpr int XNE = 258;                   // energies of those tables
constexpr int KAP_N = 27;                  // kappa_10 knots (XRayHeatingFunctions.c:1736-1789)

// constants restated from src/meraxes.h:21-45 and src/core/reionization.h:10-47,
// XRayHeatingFunctions.h:20-33
constexpr double SPEED_OF_LIGHT = 2.9979e10;
constexpr double BOLTZMANN = 1.3806e-16;
constexpr double XRAY_LUMINOSITY_UNIT = 1.0e40;
constexpr double SOLAR_LUM = 3.826e33;  // erg/s (src/meraxes.h:23)
constexpr double NU_over_EV = 1.60217646e-12 / PLANCK;
constexpr double NUIONIZATION = 13.60 * NU_over_EV;
constexpr double HeI_NUIONIZATION = 24.59 * NU_over_EV;
constexpr double HeII_NUIONIZATION = NUIONIZATION * 4;
constexpr double NU_LA = 10.2 * NU_over_EV;
constexpr double TINY = 1e-30;
constexpr double FRACT_FLOAT_ERR = 1e-7;
constexpr float MAX_TK = 5e4f;
constexpr float MIN_TK = 0.1f;
constexpr double CLUMPING_FACTOR = 2.0;
constexpr double T21 = 0.0628;
constexpr double A10_HYPERFINE = 2.85e-15;
constexpr float R_XLy_MAX = 500.0f;  // ComputeTs.h:6
constexpr int NSPEC_MAX = 23;
constexpr int TS_MAX_ITER = 1000;           // bound of the T_S fixed point (converges in a few)

// one operation of the reference's SFR/X-ray source grids (ComputeTs.c:366-478 with
// load_reion_sfr_grids, reionization.c:2750-2797), applied element-wise in float
enum { TS_OP_CLEAR = MXR_TS_OP_CLEAR, TS_OP_SCALE = MXR_TS_OP_SCALE, TS_OP_LOAD = MXR_TS_OP_LOAD,
       TS_OP_ADD = MXR_TS_OP_ADD };
struct TsOp {
  int kind;
  int hist;   // history index (LOAD, ADD)
  float wf;   // float weight (LOAD, ADD)
  double wd;  // double weight (SCALE)
};

// cosmology and number densities with the reference's float-argument promotions
struct TsCosmo {
  double OmegaM, OmegaL, OmegaR, OmegaK, wL, h;
  double No, He_No, f_H, f_He, N_b0;
  void init(const mxr_ts_params* p);
  double dtdz(float z) const;    // XRayHeatingFunctions.c:426-440
  double hubble(float z) const;  // :443-453
  double drdz(float z) const;    // :519-522
  double dicke(double z) const;  // :1595-1629 (returns -1 where the reference has no growth function)
  double ddicke_dz(double z) const;
};

// shell geometry, history plan and the scalars that precede the GPU work
struct TsPlan {
  int F = 0;
  int agn_soft = 0, agn_hard = 0;  // AGN bands the flag asks for (ComputeTs.c:159-162)
  int homogeneous = 0;
  double zp = 0, dzp = 0, dt_dzp = 0;
  double R[TS_MAXF], zpp_edge[TS_MAXF], zpp[TS_MAXF];
  int nops[TS_MAXF];
  TsOp ops[TS_MAXF][TS_MAX_OPS];
  int hist_max = -1;      // largest history index read
  double weight_end = 1;  // the reference's `weight` after the shell loop (read by the kink fix)
};

// the per-call table block the kernels read (uniform across a wavefront)
struct TsDev {
  int F, no_light;
  float zpf;           // (float)zp, the evolve step's redshift
  float cT_ad;         // homogeneous branch: cT_approx(z)
  float xe_h, TK_h;    // homogeneous branch: xion_RECFAST(z), T_RECFAST(z) as float
  double zp;           // double zp (MIN_TK reset)
  double dzp, dt_dzp, growth, dgrowth, prefac_gal, conv;
  double pv, UMUT, ULm3;          // pixel volume (Mpc^3), UnitMass/UnitTime, UnitLength^-3
  double No, N_b0, f_H, f_He, h, HUBBLEh;
  double P2, nb3;                  // pow(1 + zp_f, 2), pow(1 + zp_f, 3) with a float 1 + zp
  double Trad, Trad4, opz3, hub_z, hub_ratio;  // at z = (double)zp_f
  double K1, R03;                  // R_ct = 0 collapse fraction: RtoM(R) h h, R^3
  // per-cell tail constants: 1/Trad, 1.342881e-7 No (1+z)^3 / H(z), 1.66e11 / (1+z),
  // -1.51e-4 Trad^4 / (H(z)/H0 h (1+z)), 3 / (1+z), dt/dz 2 / (3 k_B)
  double iTrad, c_tgp, c_xa, c_comp, c_adia, c_heat;
  // shell sums with the unit constants folded in: for k = heat, ion, lya the rate at x_e in
  // bin m is sum_r L_X,r (fold[2k][m][r] dx + fold[2k+1][m][r]), dx = x_e - x[m], with
  // L_X,r the smoothed X-ray field (1e40 erg/s per cell); the stellar Lyman-alpha sum is
  // sum_r SFR_r bs[r] (SFR in internal units per cell)
  double fold[6][XNH][TS_MAXF];
  double bs[TS_MAXF];
  double cx;                       // 1e40 UnitLength^-3 / V_pix: X-ray field -> emissivity density
  // AGN X-ray bands (ComputeTs.c:835-899, 1157-1340): nagn of them, band b of kind agn_kind[b]
  // (0 soft, 1 hard).  fold_agn[b] is laid out as fold with the band's (1 + z'')^-alpha, its
  // unit constant cagn and its prefactor const_zp_prefactor_AGN_* folded in, so the heat,
  // ionisation and Lyman-alpha rates of band b are sum_r B_b,r (fold_agn[b][2k][m][r] dx +
  // fold_agn[b][2k+1][m][r]) with B_b,r the smoothed emissivity field (1e10 L_sun per cell)
  int nagn;
  int agn_kind[2];
  double cagn;                     // 1e10 SOLAR_LUM / V_pix UnitLength^-3
  double fold_agn[2][6][XNH][TS_MAXF];
  float xint[XNH];
  // kappa splines in log-log (x, y, c = S''/2): kappa_10 (27), e-H (20), p-H (17)
  int kn[3];
  double kx[3][KAP_N], ky[3][KAP_N], kc[3][KAP_N];
  double k10_E1, k10_E2;           // exp(kap[26]), exp(tkin[26]) for the T^0.381 extrapolation
};

// host: plan and tables (mxr_ts_host.hip)
int ts_history_depth(const mxr_ts_params* p, int last_snapshot);
int ts_plan(const mxr_ts_params* p, TsPlan* plan);
// after the R_ct = 0 reductions: fill the device block (and the public tables when asked)
int ts_tables(const mxr_ts_params* p, const TsPlan& plan, double collapse_fraction, double x_e_ave,
              const double* stored_fcoll, TsDev* dev, mxr_ts_tables* out);
// homogeneous branch scalars
int ts_homogeneous(const mxr_ts_params* p, TsDev* dev);
double ts_rtom(const mxr_ts_params* p, double R);

}  // namespace mxr
