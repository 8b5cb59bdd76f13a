/*
 * ts.c -- oracle restatement of the spin-temperature step (TEST INFRASTRUCTURE ONLY; see
 * oracle.h).  It consumes the grid-independent tables (shell geometry, history plan,
 * frequency tables, sum_lyn, prefactors) that the library reports for the same call -- those
 * are checked separately against oracle/ts_tables.py -- and restates everything that runs
 * over the grid in the reference's own loop structure and C types:
 *
 *   src/core/ComputeTs.c:252-331    homogeneous branch
 *   src/core/ComputeTs.c:339-746    shell loop: source grids, r2c, /N^3 (float), filter, c2r,
 *                                   clamp, smoothed fields, R_ct = 0 sums
 *   src/core/ComputeTs.c:1171-1505  per-cell tables at x_e, evolveInt, Euler step, get_Ts,
 *                                   volume averages
 *   src/core/reionization.c:2750-2797  load_reion_sfr_grids (new load / accumulate, float)
 *   src/core/reionization.c:2799-2870  load_reion_bh_grids (the AGN grids: loads and
 *                                   accumulations only; ComputeTs.c clears and rescales sfr
 *                                   and xray_luminosity alone)
 *   src/core/XRayHeatingFunctions.c:134-252 (T/xion_RECFAST), 426-453 (dtdz, hubble),
 *     1308-1561 (evolveInt), 1564-1595 (dT_comp, alpha_A), 1639-2008 (get_Ts, xcoll, kappa_10*,
 *     xalpha_tilde, taugp, Salpha_tilde, Tc_eff)
 *
 * Splines are the textbook natural cubic spline (oracle_spline_natural), the interpolant GSL's
 * cspline builds.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define HUBBLE 3.2407789e-18
#define GRAVITY 6.672e-8
#define PROTONMASS 1.6726e-24
#define SOLAR_MASS 1.989e33
#define SPEED_OF_LIGHT 2.9979e10
#define BOLTZMANN 1.3806e-16
#define TCMB 2.728
#define XRAY_LUMINOSITY_UNIT 1.0e40
#define SOLAR_LUM 3.826e33
#define FRACT_FLOAT_ERR (double)(1e-7)
#define MAX_TK (float)5e4
#define MIN_TK (float)0.1
#define CLUMPING_FACTOR (double)(2)
#define T21 (double)(0.0628)
#define A10_HYPERFINE (double)(2.85e-15)
#define L_FACTOR 0.620350491
#define NXHII MXR_X_INT_NXHII
#define RECFAST_NPTS 501

typedef struct {
  double OMm, OMl, OMr, h;
  double No, N_b0, f_H, f_He;
  /* splines (log-log) */
  double k10x[27], k10y[27], kex[20], key[20], kpx[17], kpy[17];
  double zt[RECFAST_NPTS], xion[RECFAST_NPTS], TKr[RECFAST_NPTS];
} ts_env;

static double dtdz(const ts_env* e, float z)
{
  double x = sqrt(e->OMl / e->OMm) * pow(1 + z, -3.0 / 2.0);
  double dxdz = sqrt(e->OMl / e->OMm) * pow(1 + z, -5.0 / 2.0) * (-3.0 / 2.0);
  double const1 = 2 * sqrt(1 + e->OMm / e->OMl) / (3.0 * HUBBLE * e->h);
  double numer = dxdz * (1 + x * pow(pow(x, 2) + 1, -0.5));
  double denom = x + sqrt(pow(x, 2) + 1);
  return (const1 * numer / denom);
}

static double hubble(const ts_env* e, float z)
{
  return HUBBLE * e->h * sqrt(e->OMm * pow(1 + z, 3) + e->OMr * pow(1 + z, 4) + e->OMl);
}

static int env_init(ts_env* e, const mxr_ts_params* p)
{
  static const double tk[27] = {1.0,  2.0,  4.0,  6.0,   8.0,   10.0,   15.0,   20.0,   25.0,
                                30.0, 40.0, 50.0, 60.0,  70.0,  80.0,   90.0,   100.0,  200.0,
                                300.0, 501.0, 701.0, 1000.0, 2000.0, 3000.0, 5000.0, 7000.0, 10000.0};
  static const double kp[27] = {1.38e-13, 1.43e-13, 2.71e-13, 6.60e-13, 1.47e-12, 2.88e-12, 9.10e-12,
                                1.78e-11, 2.73e-11, 3.67e-11, 5.38e-11, 6.86e-11, 8.14e-11, 9.25e-11,
                                1.02e-10, 1.11e-10, 1.19e-10, 1.75e-10, 2.09e-10, 2.565e-10, 2.91e-10,
                                3.31e-10, 4.27e-10, 4.97e-10, 6.03e-10, 6.87e-10, 7.87e-10};
  e->OMm = p->OmegaM;
  e->OMl = p->OmegaLambda;
  e->OMr = p->OmegaR;
  e->h = p->Hubble_h;
  double rho = (3.0 * HUBBLE * HUBBLE * e->h * e->h / (8.0 * M_PI * GRAVITY));
  double OMb = p->BaryonFrac * p->OmegaM;
  e->No = rho * OMb * (1 - p->Y_He) / PROTONMASS;
  double He_No = rho * OMb * p->Y_He / (4.0 * PROTONMASS);
  e->f_H = e->No / (e->No + He_No);
  e->f_He = He_No / (e->No + He_No);
  e->N_b0 = e->No + He_No;
  for (int i = 0; i < 27; i++) {
    e->k10x[i] = log(tk[i]);
    e->k10y[i] = log(kp[i]);
  }
  char fn[1024];
  const char* names[2] = {"kappa_eH_table.dat", "kappa_pH_table.dat"};
  int ns[2] = {20, 17};
  double* xs[2] = {e->kex, e->kpx};
  double* ys[2] = {e->key, e->kpy};
  for (int t = 0; t < 2; t++) {
    snprintf(fn, sizeof(fn), "%s/%s", p->TablesForXHeatingDir, names[t]);
    FILE* F = fopen(fn, "r");
    if (!F)
      return -1;
    for (int i = 0; i < ns[t]; i++) {
      float a, b;
      if (fscanf(F, "%f %e", &a, &b) != 2) {
        fclose(F);
        return -1;
      }
      xs[t][i] = log((double)a);
      ys[t][i] = log((double)b);
    }
    fclose(F);
  }
  snprintf(fn, sizeof(fn), "%s/recfast_LCDM.dat", p->TablesForXHeatingDir);
  FILE* F = fopen(fn, "r");
  if (!F)
    return -1;
  for (int i = RECFAST_NPTS - 1; i >= 0; i--) {
    float z, x, trash, T;
    if (fscanf(F, "%f %E %E %E", &z, &x, &trash, &T) != 4) {
      fclose(F);
      return -1;
    }
    e->zt[i] = z;
    e->xion[i] = x;
    e->TKr[i] = T;
  }
  fclose(F);
  return 0;
}

static double kappa_10(const ts_env* e, double TK)
{
  double ans;
  if (log(TK) < e->k10x[0])
    ans = e->k10y[0];
  else if (log(TK) > e->k10x[26])
    ans = log(exp(e->k10y[26]) * pow(TK / exp(e->k10x[26]), 0.381));
  else
    ans = oracle_spline_natural(e->k10x, e->k10y, 27, log(TK));
  return exp(ans);
}

static double kappa_lin(const double* X, const double* Y, int n, double T)
{
  double ans;
  T = log(T);
  if (T < X[0])
    ans = Y[0];
  else if (T > X[n - 1])
    ans = Y[n - 1] + ((Y[n - 1] - Y[n - 2]) / (X[n - 1] - X[n - 2]) * (T - X[n - 1]));
  else
    ans = oracle_spline_natural(X, Y, n, T);
  return exp(ans);
}

static double xcoll(const ts_env* e, double z, double TK, double delta, double xe)
{
  double Trad = TCMB * (1.0 + z);
  double nH = (1.0 - xe) * e->No * pow(1.0 + z, 3.0) * (1.0 + delta);
  double xHI = T21 / Trad * nH * kappa_10(e, TK) / A10_HYPERFINE;
  double ne = xe * e->N_b0 * pow(1.0 + z, 3.0) * (1.0 + delta);
  double xel = T21 / Trad * ne * kappa_lin(e->kex, e->key, 20, TK) / A10_HYPERFINE;
  double np = xe * e->No * pow(1.0 + z, 3.0) * (1.0 + delta);
  double xp = T21 / Trad * np * kappa_lin(e->kpx, e->kpy, 17, TK) / A10_HYPERFINE;
  return xHI + xel + xp;
}

static double Salpha_tilde(double TK, double TS, double tauGP)
{
  double xi = pow(1.0e-7 * tauGP / TK / TK, 1.0 / 3.0);
  double ans = 1.0 - 0.0631789 / TK + 0.115995 / TK / TK - 0.401403 / TS / TK;
  ans += 0.336463 / TS / TK / TK;
  ans /= 1.0 + 2.98394 * xi + 1.53583 * xi * xi + 3.85289 * xi * xi * xi;
  return ans;
}

static float get_Ts(const ts_env* e, float z, float delta, float TK, float xe, float Jalpha, float* curr_xalpha)
{
  double Trad = TCMB * (1.0 + z);
  double xc = xcoll(e, z, TK, delta, xe);
  double TS, TSold, TSinv, xa = 0;
  if (Jalpha > 1.0e-20) {
    TS = Trad;
    TSold = 0.0;
    for (int it = 0; it < 1000 && fabs(TS - TSold) / TS > 1.0e-3; it++) {
      TSold = TS;
      double tgp = 1.342881e-7 / hubble(e, (float)z) * e->No * pow(1 + (double)z, 3) * (1.0 + delta) * (1.0 - xe);
      double S = Salpha_tilde(TK, TS, tgp);
      xa = 1.66e11 / (1.0 + z) * S * Jalpha;
      double Tc = 1.0 / TK + 0.405535 / TK * (1.0 / TS - 1.0 / TK);
      Tc = 1.0 / Tc;
      TSinv = (1.0 / Trad + xa / Tc + xc / TK) / (1.0 + xa + xc);
      TS = 1.0 / TSinv;
    }
    *curr_xalpha = (float)xa;
  } else {
    TSinv = (1.0 / Trad + xc / TK) / (1.0 + xc);
    TS = 1.0 / TSinv;
    *curr_xalpha = 0;
  }
  if (TS < 0.)
    TS = fabs(TS);
  return (float)TS;
}

static double alpha_A(double T)
{
  double logT = log(T / (double)1.1604505e4);
  return exp(-28.6130338 - 0.72411256 * logT - 2.02604473e-2 * pow(logT, 2) - 2.38086188e-3 * pow(logT, 3) -
             3.21260521e-4 * pow(logT, 4) - 1.42150291e-5 * pow(logT, 5) + 4.98910892e-6 * pow(logT, 6) +
             5.75561414e-7 * pow(logT, 7) - 1.85676704e-8 * pow(logT, 8) - 3.07113524e-9 * pow(logT, 9));
}

static double dT_comp(const ts_env* e, double z, double TK, double xe)
{
  double Trad = TCMB * (1.0 + z);
  double ans = (-1.51e-4) * (xe / (1.0 + xe + e->f_He)) / (hubble(e, (float)z) / (HUBBLE * e->h)) / e->h *
               pow(Trad, 4.0) / (1.0 + z);
  ans *= Trad - TK;
  return ans;
}

/* evolveInt without mini-haloes or Lyman-Werner (XRayHeatingFunctions.c:1308-1561): the
 * galaxy terms and the AGN soft / hard bands (XAGN_* zero and freq_int_*_AGN_* zero for a
 * band the flag leaves out, as ComputeTs.c:1189-1191, 1260-1300 pass them) */
static void evolve_int(const ts_env* e, const mxr_ts_tables* tb, const mxr_ts_params* p, float zp,
                       float curr_delNL0, const double* SFR, const double* XR, const double* XAGN_soft,
                       const double* XAGN_hard, const double* fheat, const double* fion, const double* flya,
                       const double* fheat_s, const double* fion_s, const double* flya_s, const double* fheat_h,
                       const double* fion_h, const double* flya_h, int NO_LIGHT, const double* y, double* deriv)
{
  const int F = tb->n;
  const double Conversion_factor = (SPEED_OF_LIGHT / (4. * M_PI)) / (PROTONMASS / SOLAR_MASS);
  double x_e = y[0], T = y[1];
  double n_b = e->N_b0 * pow(1 + zp, 3) * (1 + curr_delNL0);
  double dxheat = 0, dxion = 0, dxlya = 0, dstarlya = 0;
  double heat_s = 0, ion_s = 0, lya_s = 0, heat_h = 0, ion_h = 0, lya_h = 0;
  deriv[5] = 0.0;
  deriv[6] = 0.0;
  if (!NO_LIGHT) {
    for (int ct = 0; ct < F; ct++) {
      double zpp, dzpp;
      if (ct == 0) {
        zpp = (tb->zpp_edge[0] + zp) * 0.5;
        dzpp = zp - tb->zpp_edge[0];
      } else {
        zpp = (tb->zpp_edge[ct] + tb->zpp_edge[ct - 1]) * 0.5;
        dzpp = tb->zpp_edge[ct - 1] - tb->zpp_edge[ct];
      }
      double dt_dzpp = dtdz(e, zpp);
      double integrand = XR[ct] * pow(1 + zpp, -p->SpecIndexXrayGal);
      dxheat += dt_dzpp * dzpp * integrand * fheat[ct];
      dxion += dt_dzpp * dzpp * integrand * fion[ct];
      dxlya += dt_dzpp * dzpp * integrand * flya[ct];
      dstarlya += SFR[ct] * pow(1 + zp, 2) * (1 + zpp) * tb->sum_lyn[ct] * dt_dzpp * dzpp;
      double integrand_s = XAGN_soft[ct] * pow(1 + zpp, -p->SpecIndexXrayAGNSoft);
      heat_s += dt_dzpp * dzpp * integrand_s * fheat_s[ct];
      ion_s += dt_dzpp * dzpp * integrand_s * fion_s[ct];
      lya_s += dt_dzpp * dzpp * integrand_s * flya_s[ct];
      double integrand_h = XAGN_hard[ct] * pow(1 + zpp, -p->SpecIndexXrayAGNHard);
      heat_h += dt_dzpp * dzpp * integrand_h * fheat_h[ct];
      ion_h += dt_dzpp * dzpp * integrand_h * fion_h[ct];
      lya_h += dt_dzpp * dzpp * integrand_h * flya_h[ct];
    }
    dxheat *= tb->const_zp_prefactor_GAL;
    dxion *= tb->const_zp_prefactor_GAL;
    dxlya *= tb->const_zp_prefactor_GAL * n_b;
    dstarlya *= Conversion_factor;
    heat_s *= tb->const_zp_prefactor_AGN_soft;
    ion_s *= tb->const_zp_prefactor_AGN_soft;
    lya_s *= tb->const_zp_prefactor_AGN_soft * n_b;
    heat_h *= tb->const_zp_prefactor_AGN_hard;
    ion_h *= tb->const_zp_prefactor_AGN_hard;
    lya_h *= tb->const_zp_prefactor_AGN_hard * n_b;
  }
  double sink = alpha_A(T) * CLUMPING_FACTOR * x_e * x_e * e->f_H * n_b;
  double dxe_dzp = tb->dt_dzp * (dxion + ion_s + ion_h - sink);
  deriv[0] = dxe_dzp;
  double dadia = 3 / (1.0 + zp);
  if (fabs(curr_delNL0) > FRACT_FLOAT_ERR)
    dadia += tb->dgrowth_factor_dzp / (1.0 / curr_delNL0 + tb->growth_factor_zp);
  dadia *= (2.0 / 3.0) * T;
  double dspec = -dxe_dzp * T / (1 + x_e);
  double dcomp = dT_comp(e, zp, T, x_e);
  deriv[5] = heat_s * tb->dt_dzp * 2.0 / 3.0 / BOLTZMANN / (1.0 + x_e);
  deriv[6] = heat_h * tb->dt_dzp * 2.0 / 3.0 / BOLTZMANN / (1.0 + x_e);
  double dxheat_dzp = dxheat * tb->dt_dzp * 2.0 / 3.0 / BOLTZMANN / (1.0 + x_e);
  dxheat_dzp += deriv[5] + deriv[6];
  deriv[1] = dxheat_dzp + dcomp + dspec + dadia;
  deriv[2] = dxlya + lya_s + lya_h + dstarlya;
  deriv[3] = dxheat_dzp;
  deriv[4] = tb->dt_dzp * (dxion + ion_s + ion_h);
}

static float cT_approx(float z) { return 0.58 - 0.006 * (z - 10.0); }

/* The shells of one call (ComputeTs.c:366-746) in FP32 (fftwf, float filter and clamp, as
 * the reference) or FP64 (double source grids, FFTs, filter and clamp).  sm receives the
 * smoothed shells [G][F][N^3] of the G source grids in hist: SFR, X-ray, then the AGN bands
 * (soft before hard) the call includes.  Every grid but the X-ray one is clamped at 0; the
 * AGN grids take the plan's loads and accumulations only. */
#define SMOOTH_BODY(REAL, R2C, C2R, FILTER, RCAST, FMAX)                                                              \
  do {                                                                                                                 \
    REAL* src[4] = {NULL, NULL, NULL, NULL};                                                                           \
    REAL* spec[4] = {NULL, NULL, NULL, NULL};                                                                          \
    int ok = 1;                                                                                                        \
    for (int k = 0; k < G; k++) {                                                                                      \
      src[k] = malloc(sizeof(REAL) * npad);                                                                            \
      spec[k] = malloc(sizeof(REAL) * 2 * nc);                                                                         \
      ok = ok && src[k] && spec[k];                                                                                    \
    }                                                                                                                  \
    REAL* work = malloc(sizeof(REAL) * 2 * nc);                                                                        \
    if (!ok || !work) {                                                                                                \
      rc = -3;                                                                                                         \
    } else {                                                                                                           \
      for (int k = 0; k < G; k++)                                                                                      \
        for (long i = 0; i < npad; i++)                                                                                \
          src[k][i] = hist[k][i];                                                                                      \
      for (int r = 0; r < F; r++) {                                                                                    \
        oracle_progress("shell", r, F);                                                                                \
        for (int q = 0; q < tb->n_ops[r]; q++) {                                                                       \
          const int kind = tb->op_kind[r][q], hi = tb->op_hist[r][q];                                                  \
          const double w = tb->op_weight[r][q];                                                                        \
          for (int k = 0; k < G; k++) {                                                                                \
            const float* h = hist[k] + (long)hi * npad;                                                                \
            if (k >= 2 && (kind == MXR_TS_OP_CLEAR || kind == MXR_TS_OP_SCALE))                                        \
              continue;                                                                                                \
            for (long i = 0; i < npad; i++) {                                                                          \
              if (kind == MXR_TS_OP_CLEAR)                                                                             \
                src[k][i] = 0;                                                                                         \
              else if (kind == MXR_TS_OP_SCALE)                                                                        \
                src[k][i] *= w;                                                                                        \
              else if (kind == MXR_TS_OP_LOAD)                                                                         \
                src[k][i] = h[i] * (RCAST)w;                                                                           \
              else                                                                                                     \
                src[k][i] += h[i] * (RCAST)w;                                                                          \
            }                                                                                                          \
          }                                                                                                            \
        }                                                                                                              \
        for (int k = 0; k < G; k++) {                                                                                  \
          if (r == 0 || tb->n_ops[r] > 0) {                                                                            \
            R2C(N, src[k], spec[k]);                                                                                   \
            for (long i = 0; i < 2 * nc; i++)                                                                          \
              spec[k][i] /= (RCAST)total_n_cells;                                                                      \
          }                                                                                                            \
          memcpy(work, spec[k], sizeof(REAL) * 2 * nc);                                                                \
          if (r > 0)                                                                                                   \
            FILTER(work, 0, N, N, p->BoxSize, (RCAST)tb->R[r], p->TsHeatingFilterType);                                \
          C2R(N, work);                                                                                                \
          double* dst = sm + ((size_t)k * F + r) * nreal;                                                              \
          for (long ir = 0; ir < nreal; ir++) {                                                                        \
            long ip = (ir / N) * 2 * H + ir % N;                                                                       \
            dst[ir] = k != 1 ? FMAX(work[ip], 0.0) : work[ip];                                                         \
          }                                                                                                            \
        }                                                                                                              \
      }                                                                                                                \
    }                                                                                                                  \
    for (int k = 0; k < 4; k++) {                                                                                      \
      free(src[k]);                                                                                                    \
      free(spec[k]);                                                                                                   \
    }                                                                                                                  \
    free(work);                                                                                                        \
  } while (0)

static int smooth(const mxr_ts_params* p, const mxr_ts_tables* tb, const float* const* hist, int G, double* sm,
                  int f64)
{
  const int N = p->ReionGridDim, H = N / 2 + 1, F = tb->n;
  const long nreal = (long)N * N * N, nc = (long)N * N * H, npad = 2 * nc;
  const double total_n_cells = pow((double)N, 3);
  int rc = 0;
  if (f64)
    SMOOTH_BODY(double, oracle_r2c3d_f64, oracle_c2r3d_f64, oracle_filter_f64, double, fmax);
  else
    SMOOTH_BODY(float, oracle_r2c3d_f32, oracle_c2r3d_f32, oracle_filter_f32, float, fmaxf);
  return rc;
}

static int ts_step(const mxr_ts_params* p, const mxr_ts_tables* tb, const mxr_ts_grids* g, mxr_ts_globals* out,
                   double* red0, float* smoothed, int f64)
{
  if (!p || !tb || !g || !out || g->mem != MXR_MEM_HOST)
    return -1;
  ts_env* e = (ts_env*)malloc(sizeof(ts_env));
  if (!e || env_init(e, p)) {
    free(e);
    return -2;
  }
  const int N = p->ReionGridDim, H = N / 2 + 1;
  const long nreal = (long)N * N * N, nc = (long)N * N * H, npad = 2 * nc;
  const double total_n_cells = pow((double)N, 3);
  const double box_size = p->BoxSize / p->Hubble_h;
  const double pixel_volume = pow(box_size / (double)N, 3);
  const double zp = p->ZZ[p->snapshot];
  const double UT = p->UnitLength_in_cm / p->UnitVelocity_in_cm_per_s;
  float* xe = g->x_e_box_prev;
  float* Tk = g->Tk_box;
  float* TS = g->TS_box;
  double Ave_Ts = 0, Ave_Tk = 0, Ave_xe = 0;

  if (tb->homogeneous) {
    float zf = (float)zp;
    float xeh = (float)oracle_spline_natural(e->zt, e->xion, RECFAST_NPTS, zf);
    float TKh = (float)oracle_spline_natural(e->zt, e->TKr, RECFAST_NPTS, zf);
    float cT_ad = cT_approx(zf);
    for (long ir = 0; ir < nreal; ir++) {
      long ip = (ir / N) * 2 * H + ir % N;
      float xa;
      xe[ip] = xeh;
      Tk[ir] = TKh * (1.0 + cT_ad * g->deltax[ip]);
      TS[ir] = get_Ts(e, zf, g->deltax[ip], Tk[ir], xe[ip], 0, &xa);
    }
  } else {
    const int F = tb->n;
    /* ComputeTs.c:159-162: which AGN bands the flag includes */
    const int agn_soft = p->Flag_IncludeAGNXray == 1 || p->Flag_IncludeAGNXray == 3;
    const int agn_hard = p->Flag_IncludeAGNXray == 1 || p->Flag_IncludeAGNXray == 2;
    const float* hist[4] = {g->sfr_histories, g->xray_luminosity_histories, NULL, NULL};
    int G = 2, ks = -1, kh = -1;
    if (agn_soft) {
      ks = G;
      hist[G++] = g->bh_xray_histories_soft;
    }
    if (agn_hard) {
      kh = G;
      hist[G++] = g->bh_xray_histories_hard;
    }
    for (int k = 0; k < G; k++)
      if (!hist[k]) {
        free(e);
        return -1;
      }
    double* sm = malloc(sizeof(double) * G * (size_t)F * nreal);
    if (!sm || smooth(p, tb, hist, G, sm, f64)) {
      free(sm);
      free(e);
      return -3;
    }
    /* SMOOTHED_AGN_* of a shell and cell (ComputeTs.c:617-618, 626-627, 722-731) */
#define AGN_FIELD(k, r, ir) \
  ((k) < 0 ? 0.0 : sm[((size_t)(k) * F + (r)) * nreal + (ir)] * 1e10 * SOLAR_LUM / pixel_volume * pow(p->UnitLength_in_cm, -3.0))
    {
      double collapse_fraction = 0, x_e_ave = 0, hmxb = 0, agn_soft_ave = 0, agn_hard_ave = 0;
      mxr_hii_params hp;
      memset(&hp, 0, sizeof(hp));
      hp.OmegaM = p->OmegaM;
      hp.ReionRtoMFilterType = p->ReionRtoMFilterType;
      hp.UnitLength_in_cm = p->UnitLength_in_cm;
      hp.UnitMass_in_g = p->UnitMass_in_g;
      hp.UnitVelocity_in_cm_per_s = p->UnitVelocity_in_cm_per_s;
      const double R = tb->R[0];
      const double M = oracle_RtoM(&hp, R);
      for (long ir = 0; ir < nreal; ir++) {
        long ip = (ir / N) * 2 * H + ir % N;
        double dom = 1.0 + g->deltax[ip];
        double cf =
          g->stars[ip] / (M * p->Hubble_h * p->Hubble_h * dom) * (4.0 / 3.0) * M_PI * pow(R, 3.0) / pixel_volume;
        if (cf > 1.0)
          cf = 1.0;
        collapse_fraction += cf;
        x_e_ave += xe[ip];
        hmxb += (sm[((size_t)F + 0) * nreal + ir] / pixel_volume) * XRAY_LUMINOSITY_UNIT *
                pow(p->UnitLength_in_cm, -3.);
        agn_soft_ave += AGN_FIELD(ks, 0, ir);
        agn_hard_ave += AGN_FIELD(kh, 0, ir);
      }
      if (red0) {
        red0[0] = collapse_fraction / total_n_cells;
        red0[1] = x_e_ave / total_n_cells;
        red0[2] = hmxb / total_n_cells;
        red0[3] = agn_soft_ave / total_n_cells;
        red0[4] = agn_hard_ave / total_n_cells;
      }
      out->XrayEmissivity_HMXB = hmxb / total_n_cells;
      out->XrayEmissivity_soft = agn_soft_ave / total_n_cells;
      out->XrayEmissivity_hard = agn_hard_ave / total_n_cells;
    }
    if (smoothed)
      for (size_t i = 0; i < (size_t)G * F * nreal; i++)
        smoothed[i] = (float)sm[i];
    double J_alpha_ave = 0, xalpha_ave = 0, Xheat_ave = 0, Xion_ave = 0, Xheat_s = 0, Xheat_h = 0;
    double SFR[MXR_TS_MAX_STEPS], XR[MXR_TS_MAX_STEPS], fh[MXR_TS_MAX_STEPS], fi[MXR_TS_MAX_STEPS],
      fl[MXR_TS_MAX_STEPS];
    double XS[MXR_TS_MAX_STEPS], XH[MXR_TS_MAX_STEPS], fhs[MXR_TS_MAX_STEPS], fis[MXR_TS_MAX_STEPS],
      fls[MXR_TS_MAX_STEPS], fhh[MXR_TS_MAX_STEPS], fih[MXR_TS_MAX_STEPS], flh[MXR_TS_MAX_STEPS];
    const float* X = tb->x_int_XHII;
    for (long ir = 0; ir < nreal; ir++) {
      long ip = (ir / N) * 2 * H + ir % N;
      double ans[2] = {xe[ip], Tk[ir]}, d[7];
      for (int r = 0; r < F; r++) {
        SFR[r] = (sm[(size_t)r * nreal + ir] / pixel_volume) * (p->UnitMass_in_g / UT) *
                 pow(p->UnitLength_in_cm, -3.) / SOLAR_MASS;
        XR[r] = (sm[((size_t)F + r) * nreal + ir] / pixel_volume) * XRAY_LUMINOSITY_UNIT *
                pow(p->UnitLength_in_cm, -3.);
        double xc = xe[ip];
        if (xc > X[NXHII - 1] * 0.999)
          xc = X[NXHII - 1] * 0.999;
        else if (xc < X[0])
          xc = 1.001 * X[0];
        int lo = NXHII - 1;
        while ((float)xc < X[lo])
          lo--;
        int hi = lo + 1;
        fh[r] = (tb->freq_int_heat[hi][r] - tb->freq_int_heat[lo][r]) / (X[hi] - X[lo]);
        fh[r] *= (xc - X[lo]);
        fh[r] += tb->freq_int_heat[lo][r];
        fi[r] = (tb->freq_int_ion[hi][r] - tb->freq_int_ion[lo][r]) / (X[hi] - X[lo]);
        fi[r] *= (xc - X[lo]);
        fi[r] += tb->freq_int_ion[lo][r];
        fl[r] = (tb->freq_int_lya[hi][r] - tb->freq_int_lya[lo][r]) / (X[hi] - X[lo]);
        fl[r] *= (xc - X[lo]);
        fl[r] += tb->freq_int_lya[lo][r];
        /* the AGN bands' tables at x_e (ComputeTs.c:1260-1300), zero for a band left out */
        XS[r] = AGN_FIELD(ks, r, ir);
        XH[r] = AGN_FIELD(kh, r, ir);
#define INTERP(T) (((T)[hi][r] - (T)[lo][r]) / (X[hi] - X[lo]) * (xc - X[lo]) + (T)[lo][r])
        fhs[r] = agn_soft ? INTERP(tb->freq_int_heat_AGN_soft) : 0.0;
        fis[r] = agn_soft ? INTERP(tb->freq_int_ion_AGN_soft) : 0.0;
        fls[r] = agn_soft ? INTERP(tb->freq_int_lya_AGN_soft) : 0.0;
        fhh[r] = agn_hard ? INTERP(tb->freq_int_heat_AGN_hard) : 0.0;
        fih[r] = agn_hard ? INTERP(tb->freq_int_ion_AGN_hard) : 0.0;
        flh[r] = agn_hard ? INTERP(tb->freq_int_lya_AGN_hard) : 0.0;
#undef INTERP
      }
      evolve_int(e, tb, p, (float)zp, g->deltax[ip], SFR, XR, XS, XH, fh, fi, fl, fhs, fis, fls, fhh, fih, flh,
                 tb->NO_LIGHT, ans, d);
      Xheat_s += d[5];
      Xheat_h += d[6];
      xe[ip] += d[0] * tb->dzp;
      if (xe[ip] > 1)
        xe[ip] = (float)(1 - FRACT_FLOAT_ERR);
      else if (xe[ip] < 0)
        xe[ip] = 0;
      if (Tk[ir] < MAX_TK)
        Tk[ir] += d[1] * tb->dzp;
      if (Tk[ir] < MIN_TK)
        Tk[ir] = (float)(TCMB * (1 + zp));
      float xa;
      TS[ir] = get_Ts(e, (float)zp, g->deltax[ip], Tk[ir], xe[ip], (float)d[2], &xa);
      J_alpha_ave += d[2];
      xalpha_ave += xa;
      Xheat_ave += d[3];
      Xion_ave += d[4];
    }
    out->volume_ave_J_alpha = J_alpha_ave / total_n_cells;
    out->volume_ave_xalpha = xalpha_ave / total_n_cells;
    out->volume_ave_Xheat = Xheat_ave / total_n_cells;
    out->volume_ave_Xion = Xion_ave / total_n_cells;
    out->volume_ave_Xheat_AGN_soft = Xheat_s / total_n_cells;
    out->volume_ave_Xheat_AGN_hard = Xheat_h / total_n_cells;
#undef AGN_FIELD
    free(sm);
  }
  memcpy(g->x_e_box, xe, sizeof(float) * npad);
  for (long ir = 0; ir < nreal; ir++) {
    long ip = (ir / N) * 2 * H + ir % N;
    Ave_Ts += TS[ir];
    Ave_Tk += Tk[ir];
    Ave_xe += xe[ip];
  }
  out->volume_ave_TS = Ave_Ts / total_n_cells;
  out->volume_ave_TK = Ave_Tk / total_n_cells;
  out->volume_ave_xe = Ave_xe / total_n_cells;
  free(e);
  return 0;
}

int oracle_ts_step_f32(const mxr_ts_params* p, const mxr_ts_tables* tb, const mxr_ts_grids* g,
                       mxr_ts_globals* out, double* red0, float* smoothed)
{
  return ts_step(p, tb, g, out, red0, smoothed, 0);
}

int oracle_ts_step_f64(const mxr_ts_params* p, const mxr_ts_tables* tb, const mxr_ts_grids* g,
                       mxr_ts_globals* out, double* red0, float* smoothed)
{
  return ts_step(p, tb, g, out, red0, smoothed, 1);
}

/* The helpers on their own, for the tests that pin them (tables read from the params' directory):
 * what = 0 get_Ts(z, delta, TK, xe, Jalpha) -> out[0] T_S, out[1] x_alpha; 1 xcoll(z, TK,
 * delta, xe); 2 alpha_A(T); 3 dT_comp(z, TK, xe); 4 hubble(z); 5 dtdz(z). */
int oracle_ts_probe(const mxr_ts_params* p, int what, const double* a, double* out)
{
  ts_env* e = malloc(sizeof(ts_env));
  if (!e)
    return -3;
  if (env_init(e, p)) {
    free(e);
    return MXR_ERR_TABLES;
  }
  int rc = 0;
  switch (what) {
    case 0: {
      float xa = 0.f;
      out[0] = get_Ts(e, (float)a[0], (float)a[1], (float)a[2], (float)a[3], (float)a[4], &xa);
      out[1] = xa;
      break;
    }
    case 1: out[0] = xcoll(e, a[0], a[1], a[2], a[3]); break;
    case 2: out[0] = alpha_A(a[0]); break;
    case 3: out[0] = dT_comp(e, a[0], a[1], a[2]); break;
    case 4: out[0] = hubble(e, (float)a[0]); break;
    case 5: out[0] = dtdz(e, (float)a[0]); break;
    default: rc = MXR_ERR_INVALID_ARG;
  }
  free(e);
  return rc;
}
