// mxr_ts_host.hip -- host side of the spin-temperature step: shell geometry and the history
// plan (src/core/ComputeTs.c:32-83, 339-478), the per-shell frequency tables (:761-1011), the
// luminosity prefactors (:1013-1128) and the helper physics of
// src/core/XRayHeatingFunctions.c.  Everything here is grid-independent; the kernels that
// use it live in mxr_ts.hip.
//
// Arithmetic follows the reference's C types: functions the reference declares with a float
// argument (dtdz, hubble, drdz, zmax) take float here, `1 + z` with a float z is a float sum,
// the secondary-electron interpolation is float, and the adaptive integration and the root
// search restate the algorithms of GSL's gsl_integration_qag (Gauss-Kronrod 61) and
// gsl_root_fsolver_brent, which the reference calls.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "mxr_qag.hpp"
#include "mxr_ts.hpp"

namespace mxr {

constexpr double DBL_EPS = 2.2204460492503131e-16;
constexpr double DBL_MINV = 2.2250738585072014e-308;

// ---- cosmology ------------------------------------------------------------------------------
void TsCosmo::init(const mxr_ts_params* p)
{
  OmegaM = p->OmegaM;
  OmegaL = p->OmegaLambda;
  OmegaR = p->OmegaR;
  OmegaK = p->OmegaK;
  wL = p->wLambda;
  h = p->Hubble_h;
  // reionization.h:15-32 (RHOcrit_cgs, OMb, No, He_No, f_H, f_He, N_b0)
  const double rho_cgs = 3.0 * HUBBLE * HUBBLE * h * h / (8.0 * M_PI * GRAVITY);
  const double OMb = p->BaryonFrac * p->OmegaM;
  No = rho_cgs * OMb * (1 - p->Y_He) / PROTONMASS;
  He_No = rho_cgs * OMb * p->Y_He / (4.0 * PROTONMASS);
  f_H = No / (No + He_No);
  f_He = He_No / (No + He_No);
  N_b0 = No + He_No;
}

double TsCosmo::dtdz(float z) const
{
  const float opz = 1 + z;
  const double x = std::sqrt(OmegaL / OmegaM) * std::pow((double)opz, -3.0 / 2.0);
  const double dxdz = std::sqrt(OmegaL / OmegaM) * std::pow((double)opz, -5.0 / 2.0) * (-3.0 / 2.0);
  const double const1 = 2 * std::sqrt(1 + OmegaM / OmegaL) / (3.0 * HUBBLE * h);
  const double numer = dxdz * (1 + x * std::pow(std::pow(x, 2) + 1, -0.5));
  const double denom = x + std::sqrt(std::pow(x, 2) + 1);
  return (const1 * numer / denom);
}

double TsCosmo::hubble(float z) const
{
  const float opz = 1 + z;
  return HUBBLE * h * std::sqrt(OmegaM * std::pow((double)opz, 3) + OmegaR * std::pow((double)opz, 4) + OmegaL);
}

double TsCosmo::drdz(float z) const { return (1.0 + z) * SPEED_OF_LIGHT * dtdz(z); }

double TsCosmo::dicke(double z) const
{
  const double tiny = 1e-4;
  const double OMm = OmegaM, OMl = OmegaL, OMr = OmegaR, OMk = OmegaK;
  const double OMtot = OMm + OMl + OMr + OMk;
  if (std::fabs(OMm - 1.0) < tiny)
    return 1.0 / (1.0 + z);
  if ((OMl > (-tiny)) && (std::fabs(OMl + OMm + OMr - 1.0) < 0.01) && (std::fabs(wL + 1.0) < tiny)) {
    const double omegaM_z = OMm * std::pow(1 + z, 3) / (OMl + OMm * std::pow(1 + z, 3) + OMr * std::pow(1 + z, 4));
    const double dick_z =
      2.5 * omegaM_z / (1.0 / 70.0 + omegaM_z * (209 - omegaM_z) / 140.0 + std::pow(omegaM_z, 4.0 / 7.0));
    const double dick_0 = 2.5 * OMm / (1.0 / 70.0 + OMm * (209 - OMm) / 140.0 + std::pow(OMm, 4.0 / 7.0));
    return dick_z / (dick_0 * (1.0 + z));
  }
  if ((OMtot < (1 + tiny)) && (std::fabs(OMl) < tiny)) {
    const double x_0 = 1.0 / (OMm + 0.0) - 1.0;
    const double dick_0 = 1 + 3.0 / x_0 + 3 * std::log(std::sqrt(1 + x_0) - std::sqrt(x_0)) * std::sqrt(1 + x_0) /
                                             std::pow(x_0, 1.5);
    const double x = std::fabs(1.0 / (OMm + 0.0) - 1.0) / (1 + z);
    const double dick_z =
      1 + 3.0 / x + 3 * std::log(std::sqrt(1 + x) - std::sqrt(x)) * std::sqrt(1 + x) / std::pow(x, 1.5);
    return dick_z / dick_0;
  }
  return -1;
}

double TsCosmo::ddicke_dz(double z) const
{
  const float dz = 1e-10f;
  return (dicke(z + dz) - dicke(z)) / dz;
}

// ---- GSL-form natural cubic spline (gsl_interp_cspline) ---------------------------------------
// c = S''/2 at the knots from the symmetric tridiagonal system, solved by LDL^T; evaluation
// y_lo + t (b + t (c_i + t d)) on the bracketing interval.  Outside [x_0, x_n-1] it returns NaN.
struct CSpline {
  std::vector<double> x, y, c;
  void init(const double* xa, const double* ya, int n)
  {
    x.assign(xa, xa + n);
    y.assign(ya, ya + n);
    c.assign(n, 0.0);
    const int sys = n - 2;
    if (sys <= 0)
      return;
    std::vector<double> diag(sys), off(sys), g(sys);
    for (int i = 0; i < sys; i++) {
      const double h_i = xa[i + 1] - xa[i], h_ip1 = xa[i + 2] - xa[i + 1];
      const double g_i = h_i != 0.0 ? 1.0 / h_i : 0.0, g_ip1 = h_ip1 != 0.0 ? 1.0 / h_ip1 : 0.0;
      off[i] = h_ip1;
      diag[i] = 2.0 * (h_ip1 + h_i);
      g[i] = 3.0 * ((ya[i + 2] - ya[i + 1]) * g_ip1 - (ya[i + 1] - ya[i]) * g_i);
    }
    if (sys == 1) {
      c[1] = g[0] / diag[0];
      return;
    }
    std::vector<double> gam(sys), alp(sys), z(sys), cc(sys);
    alp[0] = diag[0];
    gam[0] = off[0] / alp[0];
    for (int i = 1; i < sys - 1; i++) {
      alp[i] = diag[i] - off[i - 1] * gam[i - 1];
      gam[i] = off[i] / alp[i];
    }
    alp[sys - 1] = diag[sys - 1] - off[sys - 2] * gam[sys - 2];
    z[0] = g[0];
    for (int i = 1; i < sys; i++)
      z[i] = g[i] - gam[i - 1] * z[i - 1];
    for (int i = 0; i < sys; i++)
      cc[i] = z[i] / alp[i];
    c[1 + sys - 1] = cc[sys - 1];
    for (int i = sys - 2; i >= 0; i--)
      c[1 + i] = cc[i] - gam[i] * c[1 + i + 1];
  }
  double eval(double xq) const
  {
    const int n = (int)x.size();
    if (xq < x[0] || xq > x[n - 1])
      return std::nan("");
    int ilo = 0, ihi = n - 1;
    while (ihi > ilo + 1) {
      const int i = (ihi + ilo) / 2;
      if (x[i] > xq)
        ihi = i;
      else
        ilo = i;
    }
    const double dx = x[ilo + 1] - x[ilo];
    const double dy = y[ilo + 1] - y[ilo];
    const double delx = xq - x[ilo];
    const double b = (dy / dx) - dx * (c[ilo + 1] + 2.0 * c[ilo]) / 3.0;
    const double d = (c[ilo + 1] - c[ilo]) / (3.0 * dx);
    return y[ilo] + delx * (b + delx * (c[ilo] + delx * d));
  }
};

// ---- data tables (read as the reference reads them) -------------------------------------------
struct TsData {
  std::string dir;
  bool ok = false;
  CSpline recfast_T, recfast_x;              // XRayHeatingFunctions.c:134-252
  CSpline kap10, kapeH, kappH;               // :1728-1966 (log-log)
  double kap10_x[KAP_N], kap10_y[KAP_N];
  double keH_x[20], keH_y[20], kpH_x[17], kpH_y[17];
  int spec_n[NSPEC_MAX];                     // :643-728
  float spec_N0[NSPEC_MAX], spec_alpha[NSPEC_MAX], spec_nu[NSPEC_MAX];
  double spec_pop2 = -1;                     // Pop2_ion the normalisation used
  float N0_raw[NSPEC_MAX];
  float x_int_XHII[XNH];                     // :861-968
  float E[XNE];
  float fheat[XNH][XNE], nLya[XNH][XNE], nHI[XNH][XNE], nHeI[XNH][XNE], nHeII[XNH][XNE];
};

static std::mutex g_data_mu;
static std::vector<TsData*> g_data;

static int load_data(const char* dir_c, TsData** out)
{
  if (!dir_c) {
    set_error("ComputeTs: TablesForXHeatingDir is NULL");
    return MXR_ERR_TABLES;
  }
  std::lock_guard<std::mutex> lk(g_data_mu);
  for (TsData* d : g_data)
    if (d->dir == dir_c) {
      *out = d;
      return MXR_OK;
    }
  TsData* d = new TsData();
  d->dir = dir_c;
  auto fail = [&](const std::string& what) {
    set_error("ComputeTs tables: %s", what.c_str());
    delete d;
    return MXR_ERR_TABLES;
  };
  // recfast_LCDM.dat: z, x_e, -, T_K, rows in decreasing z
  {
    const std::string fn = d->dir + "/recfast_LCDM.dat";
    FILE* F = std::fopen(fn.c_str(), "r");
    if (!F)
      return fail("unable to open " + fn);
    double zt[RECFAST_NPTS], xe[RECFAST_NPTS], TK[RECFAST_NPTS];
    for (int i = RECFAST_NPTS - 1; i >= 0; i--) {
      float z, x, trash, T;
      if (std::fscanf(F, "%f %E %E %E", &z, &x, &trash, &T) != 4) {
        std::fclose(F);
        return fail(fn + " is truncated");
      }
      zt[i] = z;
      xe[i] = x;
      TK[i] = T;
    }
    std::fclose(F);
    d->recfast_T.init(zt, TK, RECFAST_NPTS);
    d->recfast_x.init(zt, xe, RECFAST_NPTS);
  }
  // kappa_10 (Zygelman 2005, Table 2, column 4), hard-coded in the reference
  {
    static const double tk[KAP_N] = {1.0,  2.0,  4.0,  6.0,   8.0,   10.0,  15.0,  20.0,   25.0,
                                     30.0, 40.0, 50.0, 60.0,  70.0,  80.0,  90.0,  100.0,  200.0,
                                     300.0, 501.0, 701.0, 1000.0, 2000.0, 3000.0, 5000.0, 7000.0, 10000.0};
    static const double kp[KAP_N] = {1.38e-13, 1.43e-13, 2.71e-13, 6.60e-13, 1.47e-12, 2.88e-12, 9.10e-12,
                                     1.78e-11, 2.73e-11, 3.67e-11, 5.38e-11, 6.86e-11, 8.14e-11, 9.25e-11,
                                     1.02e-10, 1.11e-10, 1.19e-10, 1.75e-10, 2.09e-10, 2.565e-10, 2.91e-10,
                                     3.31e-10, 4.27e-10, 4.97e-10, 6.03e-10, 6.87e-10, 7.87e-10};
    for (int i = 0; i < KAP_N; i++) {
      d->kap10_x[i] = std::log(tk[i]);
      d->kap10_y[i] = std::log(kp[i]);
    }
    d->kap10.init(d->kap10_x, d->kap10_y, KAP_N);
  }
  auto read_kappa = [&](const char* name, int n, double* xs, double* ys, CSpline& sp) -> bool {
    const std::string fn = d->dir + "/" + name;
    FILE* F = std::fopen(fn.c_str(), "r");
    if (!F)
      return false;
    for (int i = 0; i < n; i++) {
      float t, k;
      if (std::fscanf(F, "%f %e", &t, &k) != 2) {
        std::fclose(F);
        return false;
      }
      xs[i] = t;
      ys[i] = k;
    }
    std::fclose(F);
    for (int i = 0; i < n; i++) {
      xs[i] = std::log(xs[i]);
      ys[i] = std::log(ys[i]);
    }
    sp.init(xs, ys, n);
    return true;
  };
  if (!read_kappa("kappa_eH_table.dat", 20, d->keH_x, d->keH_y, d->kapeH))
    return fail("unable to read kappa_eH_table.dat");
  if (!read_kappa("kappa_pH_table.dat", 17, d->kpH_x, d->kpH_y, d->kappH))
    return fail("unable to read kappa_pH_table.dat");
  // stellar_spectra.dat: n, N0_2, alpha_2, N0_3, alpha_3 for n = 2..23 (rows i = 1..22)
  {
    const std::string fn = d->dir + "/stellar_spectra.dat";
    FILE* F = std::fopen(fn.c_str(), "r");
    if (!F)
      return fail("unable to open " + fn);
    for (int i = 1; i < NSPEC_MAX; i++) {
      float n3, a3;
      if (std::fscanf(F, "%i %e %e %e %e", &d->spec_n[i], &d->N0_raw[i], &d->spec_alpha[i], &n3, &a3) != 5) {
        std::fclose(F);
        return fail(fn + " is truncated");
      }
    }
    std::fclose(F);
    for (int i = 1; i < NSPEC_MAX; i++)
      d->spec_nu[i] = (float)(4.0 / 3.0 * (1.0 - 1.0 / std::pow(d->spec_n[i], 2.0)));
  }
  // x_int_tables
  {
    const float xs[XNH] = {1.0e-4f, 2.318e-4f, 4.677e-4f, 1.0e-3f, 2.318e-3f, 4.677e-3f, 1.0e-2f,
                           2.318e-2f, 4.677e-2f, 1.0e-1f, 0.5f, 0.9f, 0.99f, 0.999f};
    std::memcpy(d->x_int_XHII, xs, sizeof(xs));
    for (int m = 0; m < XNH; m++) {
      char fn[1024];
      if (xs[m] < 0.3)
        std::snprintf(fn, sizeof(fn), "%s/x_int_tables/log_xi_%1.1f.dat", d->dir.c_str(), std::log10(xs[m]));
      else
        std::snprintf(fn, sizeof(fn), "%s/x_int_tables/xi_%1.3f.dat", d->dir.c_str(), (double)xs[m]);
      FILE* F = std::fopen(fn, "r");
      if (!F)
        return fail(std::string("unable to open ") + fn);
      char label[64];
      float a, b, c, zz, T, trash;
      bool okf = true;
      for (int i = 0; i < 5; i++)
        okf &= std::fscanf(F, "%63s", label) == 1;
      okf &= std::fscanf(F, "%g %g %g %g %g", &a, &b, &c, &zz, &T) == 5;
      for (int i = 0; i < 11; i++)
        okf &= std::fscanf(F, "%63s", label) == 1;
      for (int i = 0; i < XNE && okf; i++)
        okf &= std::fscanf(F, "%g %g %g %g %g %g %g %g %g", &d->E[i], &trash, &d->fheat[m][i], &trash,
                           &d->nLya[m][i], &d->nHI[m][i], &d->nHeI[m][i], &d->nHeII[m][i], &trash) == 9;
      std::fclose(F);
      if (!okf)
        return fail(std::string(fn) + " is malformed");
    }
  }
  d->ok = true;
  g_data.push_back(d);
  *out = d;
  return MXR_OK;
}

// spectral_emissivity(nu, 0, 2) after the flag-1 normalisation with Pop2_ion
// (XRayHeatingFunctions.c:643-728)
static void normalise_spectra(TsData* d, double pop2)
{
  if (d->spec_pop2 == pop2)
    return;
  for (int i = 1; i < NSPEC_MAX; i++)
    d->spec_N0[i] = d->N0_raw[i];
  for (int i = 1; i < (NSPEC_MAX - 1); i++) {
    const double n0_fac =
      (std::pow((double)d->spec_nu[i + 1], (double)(d->spec_alpha[i] + 1)) -
       std::pow((double)d->spec_nu[i], (double)(d->spec_alpha[i] + 1)));
    d->spec_N0[i] = (float)((double)d->spec_N0[i] * ((double)(d->spec_alpha[i] + 1) / n0_fac * pop2));
  }
  d->spec_pop2 = pop2;
}

static double spectral_emissivity(const TsData* d, double nu_norm)
{
  for (int i = 1; i < (NSPEC_MAX - 1); i++)
    if ((nu_norm >= d->spec_nu[i]) && (nu_norm < d->spec_nu[i + 1]))
      return d->spec_N0[i] * std::pow(nu_norm, (double)d->spec_alpha[i]) / NU_LA;
  const int i = NSPEC_MAX - 1;
  return d->spec_N0[i] * std::pow(nu_norm, (double)d->spec_alpha[i]) / NU_LA;
}

static double frecycle(int n)
{
  static const double f[31] = {1,      1,      1,      0,      0.2609, 0.3078, 0.3259, 0.3353, 0.3410, 0.3448, 0.3476,
                               0.3496, 0.3512, 0.3524, 0.3535, 0.3543, 0.3550, 0.3556, 0.3561, 0.3565, 0.3569, 0.3572,
                               0.3575, 0.3578, 0.3580, 0.3582, 0.3584, 0.3586, 0.3587, 0.3589, 0.3590};
  return (n >= 0 && n <= 30) ? f[n] : 0;
}

static float zmax(float z, int n)
{
  const double num = 1 - std::pow(n + 1, -2);
  const double denom = 1 - std::pow(n, -2);
  const float opz = 1 + z;
  return (float)(opz * num / denom - 1);
}

static double nu_n(int n)
{
  double ans = 1.0 - std::pow(n, -2.0);
  ans /= 0.75;
  return ans;
}

// ---- secondary-electron interpolation (float, XRayHeatingFunctions.c:970-1280) ---------------
static int locate_energy_index(float En)
{
  if (En < 1008.88)
    return (int)(std::log(En / 10.0) / 1.98026273e-2);
  return 232 + (int)(std::log(En / 1008.88) / 9.53101798e-2);
}

static int locate_xHII_index(const float* xint, float x)
{
  int m = XNH - 1;
  while (x < xint[m])
    m--;
  return m;
}

static float interp_tab(const TsData* d, const float (*tab)[XNE], float below, float En, float xHII)
{
#pragma clang fp contract(off)
  const float* E = d->E;
  const float* X = d->x_int_XHII;
  if (En > 0.999 * E[XNE - 1])
    En = (float)(E[XNE - 1] * 0.999);
  else if (En < E[0])
    return below;
  if (xHII > X[XNH - 1] * 0.999)
    xHII = (float)(X[XNH - 1] * 0.999);
  else if (xHII < X[0])
    xHII = (float)(1.001 * X[0]);
  const int nl = locate_energy_index(En), nh = nl + 1;
  const int ml = locate_xHII_index(X, xHII), mh = ml + 1;
  float lo = (tab[ml][nh] - tab[ml][nl]) / (E[nh] - E[nl]);
  lo *= (En - E[nl]);
  lo += tab[ml][nl];
  float hi = (tab[mh][nh] - tab[mh][nl]) / (E[nh] - E[nl]);
  hi *= (En - E[nl]);
  hi += tab[mh][nl];
  float fr = (hi - lo) / (X[mh] - X[ml]);
  fr *= (xHII - X[ml]);
  fr += lo;
  return fr;
}

// ---- cross sections (Verner 1996, Osterbrock; XRayHeatingFunctions.c:458-516) ---------------
static double HI_ion_crosssec(double nu)
{
  if (nu < NUIONIZATION)
    return 0;
  if (nu == NUIONIZATION)
    nu += TINY;
  const double Z = 1, eps = std::sqrt(nu / NUIONIZATION - 1);
  return SIGMA_HI / Z / Z * std::pow(NUIONIZATION / nu, 4) * std::exp(4 - (4 * std::atan(eps) / eps)) /
         (1 - std::exp(-2 * M_PI / eps));
}

static double HeII_ion_crosssec(double nu)
{
  if (nu < HeII_NUIONIZATION)
    return 0;
  if (nu == HeII_NUIONIZATION)
    nu += TINY;
  const double Z = 2, eps = std::sqrt(nu / HeII_NUIONIZATION - 1);
  return (6.3e-18) / Z / Z * std::pow(HeII_NUIONIZATION / nu, 4) * std::exp(4 - (4 * std::atan(eps) / eps)) /
         (1 - std::exp(-2 * M_PI / eps));
}

static double HeI_ion_crosssec(double nu)
{
  if (nu < HeI_NUIONIZATION)
    return 0;
  const double x = nu / NU_over_EV / 13.61 - 0.4434;
  const double y = std::sqrt(x * x + std::pow(2.136, 2));
  return 9.492e-16 * ((x - 1) * (x - 1) + 2.039 * 2.039) * std::pow(y, (0.5 * 3.188 - 5.5)) *
         std::pow(1.0 + std::sqrt(y / 1.469), -3.188);
}

// ---- adaptive Gauss-Kronrod (61) quadrature: GSL's QAG (mxr_qag.hpp) with the reference's
// workspace of 1000 intervals, one workspace per host thread
template <class F>
static int qag_gsl(const F& f, double a, double b, double epsabs, double epsrel, int limit, double* result,
                   double* abserr)
{
  static thread_local QagWork<1000> w;
  return qag<1000>(w, f, a, b, epsabs, epsrel, limit, result, abserr);
}

// ---- Brent's root bracketing (gsl_root_fsolver_brent + gsl_root_test_interval) --------------
struct Brent {
  double a, b, c, d, e, fa, fb, fc;
};

template <class F>
static bool brent_init(Brent& s, const F& f, double lo, double hi)
{
  const double flo = f(lo), fhi = f(hi);
  s.a = lo;
  s.fa = flo;
  s.b = hi;
  s.fb = fhi;
  s.c = hi;
  s.fc = fhi;
  s.d = hi - lo;
  s.e = hi - lo;
  return !((flo < 0.0 && fhi < 0.0) || (flo > 0.0 && fhi > 0.0));
}

template <class F>
static void brent_iterate(Brent& s, const F& f, double* root, double* x_lower, double* x_upper)
{
  double a = s.a, b = s.b, c = s.c, fa = s.fa, fb = s.fb, fc = s.fc, d = s.d, e = s.e;
  int ac_equal = 0;
  if ((fb < 0 && fc < 0) || (fb > 0 && fc > 0)) {
    ac_equal = 1;
    c = a;
    fc = fa;
    d = b - a;
    e = b - a;
  }
  if (std::fabs(fc) < std::fabs(fb)) {
    ac_equal = 1;
    a = b;
    b = c;
    c = a;
    fa = fb;
    fb = fc;
    fc = fa;
  }
  const double tol = 0.5 * DBL_EPS * std::fabs(b);
  const double m = 0.5 * (c - b);
  if (fb == 0) {
    *root = b;
    *x_lower = b;
    *x_upper = b;
    return;
  }
  if (std::fabs(m) <= tol) {
    *root = b;
    *x_lower = b < c ? b : c;
    *x_upper = b < c ? c : b;
    return;
  }
  if (std::fabs(e) < tol || std::fabs(fa) <= std::fabs(fb)) {
    d = m;
    e = m;
  } else {
    double p, q, r;
    const double sr = fb / fa;
    if (ac_equal) {
      p = 2 * m * sr;
      q = 1 - sr;
    } else {
      q = fa / fc;
      r = fb / fc;
      p = sr * (2 * m * q * (q - r) - (b - a) * (r - 1));
      q = (q - 1) * (r - 1) * (sr - 1);
    }
    if (p > 0)
      q = -q;
    else
      p = -p;
    if (2 * p < std::min(3 * m * q - std::fabs(tol * q), std::fabs(e * q))) {
      e = d;
      d = p / q;
    } else {
      d = m;
      e = m;
    }
  }
  a = b;
  fa = fb;
  if (std::fabs(d) > tol)
    b += d;
  else
    b += (m > 0 ? +tol : -tol);
  fb = f(b);
  s.a = a;
  s.b = b;
  s.c = c;
  s.d = d;
  s.e = e;
  s.fa = fa;
  s.fb = fb;
  s.fc = fc;
  *root = b;
  if ((fb < 0 && fc < 0) || (fb > 0 && fc > 0))
    c = a;
  *x_lower = b < c ? b : c;
  *x_upper = b < c ? c : b;
}

static bool root_test_interval(double lo, double hi, double epsabs, double epsrel)
{
  const double alo = std::fabs(lo), ahi = std::fabs(hi);
  double min_abs = ((lo > 0.0 && hi > 0.0) || (lo < 0.0 && hi < 0.0)) ? std::min(alo, ahi) : 0;
  return std::fabs(hi - lo) < epsabs + epsrel * min_abs;
}

// ---- shell geometry and the history plan (ComputeTs.c:32-83, 339-478) ---------------------
static double r_factor(double R0, int F) { return std::pow(R_XLy_MAX / R0, 1 / (float)F); }

int ts_history_depth(const mxr_ts_params* p, int last_snapshot)
{
  if (!p || !p->ZZ || p->TsNumFilterSteps <= 0 || p->ReionGridDim <= 0) {
    set_error("ts_history_depth: bad parameters");
    return MXR_ERR_INVALID_ARG;
  }
  TsCosmo cz;
  cz.init(p);
  const double box_size = p->BoxSize / p->Hubble_h;
  const int F = p->TsNumFilterSteps;
  int max_snaps = 0;
  for (int snapshot = 1; snapshot <= last_snapshot; snapshot++) {
    const double zp = p->ZZ[snapshot];
    double R = L_FACTOR * box_size / (float)p->ReionGridDim;
    const double Rf = r_factor(R, F);
    int scb = 1;
    double prev_R = 0, prev_zpp = zp, zpp_edge = 0;
    for (int r = 0; r < F; r++) {
      zpp_edge = prev_zpp - (R - prev_R) * MPC / (cz.drdz((float)prev_zpp));
      while (zpp_edge > p->ZZ[snapshot - scb]) {
        if (snapshot - scb == 0)
          break;
        scb += 1;
      }
      prev_R = R;
      prev_zpp = zpp_edge;
      R *= Rf;
    }
    max_snaps = std::max(max_snaps, scb);
  }
  return max_snaps;
}

static bool push_op(TsPlan* pl, int r, int kind, int hist, double w)
{
  if (pl->nops[r] >= TS_MAX_OPS) {
    set_error("ComputeTs: more than %d source operations before shell %d", TS_MAX_OPS, r);
    return false;
  }
  TsOp& o = pl->ops[r][pl->nops[r]++];
  o.kind = kind;
  o.hist = hist;
  o.wf = (float)w;
  o.wd = w;
  if (kind == TS_OP_LOAD || kind == TS_OP_ADD)
    pl->hist_max = std::max(pl->hist_max, hist);
  return true;
}

int ts_plan(const mxr_ts_params* p, TsPlan* pl)
{
  const int F = p->TsNumFilterSteps;
  const int snapshot = p->snapshot;
  const double* ZZ = p->ZZ;
  TsCosmo cz;
  cz.init(p);
  pl->F = F;
  pl->hist_max = -1;
  std::memset(pl->nops, 0, sizeof(pl->nops));
  if (p->Flag_IncludeAGNXray < MXR_AGN_XRAY_NONE || p->Flag_IncludeAGNXray > MXR_AGN_XRAY_SOFT) {
    set_error("ComputeTs: Flag_IncludeAGNXray %d is not one of 0..3", p->Flag_IncludeAGNXray);
    return MXR_ERR_INVALID_ARG;
  }
  pl->agn_soft = p->Flag_IncludeAGNXray == MXR_AGN_XRAY_BOTH || p->Flag_IncludeAGNXray == MXR_AGN_XRAY_SOFT;
  pl->agn_hard = p->Flag_IncludeAGNXray == MXR_AGN_XRAY_BOTH || p->Flag_IncludeAGNXray == MXR_AGN_XRAY_HARD;
  const double redshift = ZZ[snapshot];
  const double prev_redshift = snapshot == 0 ? ZZ[snapshot] : ZZ[snapshot - 1];
  const double zp = redshift;
  pl->zp = zp;
  pl->dzp = zp - prev_redshift;
  pl->dt_dzp = cz.dtdz((float)(float)zp);
  pl->homogeneous = (zp - p->ReionMaxHeatingRedshift) >= -0.0001;
  if (pl->homogeneous)
    return MXR_OK;

  const double box_size = p->BoxSize / p->Hubble_h;
  double R = L_FACTOR * box_size / (float)p->ReionGridDim;
  const double Rf = r_factor(R, F);
  int scb[TS_MAXF];
  scb[0] = 1;
  double weight = 1;
  for (int r = 0; r < F; r++) {
    pl->R[r] = R;
    double prev_zpp, prev_R;
    if (r == 0) {
      prev_zpp = zp;
      prev_R = 0;
    } else {
      prev_zpp = pl->zpp_edge[r - 1];
      prev_R = pl->R[r - 1];
      scb[r] = scb[r - 1];
    }
    pl->zpp_edge[r] = prev_zpp - (R - prev_R) * MPC / (cz.drdz((float)prev_zpp));
    const double zpp = (pl->zpp_edge[r] + prev_zpp) * 0.5;
    pl->zpp[r] = zpp;
    const double total_weight = (prev_zpp - pl->zpp_edge[r]) * cz.dtdz(zpp);
    if (snapshot - scb[r] >= 0) {
      if (snapshot - scb[r] == 0 && weight < 0) {
        if (!push_op(pl, r, TS_OP_CLEAR, 0, 0.0))
          return MXR_ERR_INVALID_ARG;
        scb[r] += 1;
      } else if (pl->zpp_edge[r] > ZZ[snapshot - scb[r]]) {
        if (r == 0) {
          weight = pl->dt_dzp * pl->dzp / total_weight;
          if (!push_op(pl, r, TS_OP_SCALE, 0, weight))
            return MXR_ERR_INVALID_ARG;
        } else if (ZZ[snapshot - scb[r]] > pl->zpp_edge[r - 1]) {
          weight = -cz.dtdz(0.5 * (ZZ[snapshot - scb[r]] + pl->zpp_edge[r - 1]));
          weight *= (ZZ[snapshot - scb[r]] - pl->zpp_edge[r - 1]);
          weight /= total_weight;
          if (!push_op(pl, r, TS_OP_LOAD, scb[r] - 1, weight))
            return MXR_ERR_INVALID_ARG;
        } else {
          if (!push_op(pl, r, TS_OP_CLEAR, 0, 0.0))
            return MXR_ERR_INVALID_ARG;
        }
        while (pl->zpp_edge[r] > ZZ[snapshot - scb[r]]) {
          if (snapshot - scb[r] == 0) {
            weight = -1;
            break;
          }
          if (pl->zpp_edge[r] < ZZ[snapshot - scb[r] - 1]) {
            const double zedge = prev_zpp > ZZ[snapshot - scb[r]] ? prev_zpp : ZZ[snapshot - scb[r]];
            weight = pl->zpp_edge[r] - zedge;
            weight *= -cz.dtdz(0.5 * (pl->zpp_edge[r] + zedge));
          } else {
            weight = ZZ[snapshot - scb[r] - 1] - ZZ[snapshot - scb[r]];
            weight *= -cz.dtdz(0.5 * (ZZ[snapshot - scb[r] - 1] + ZZ[snapshot - scb[r]]));
          }
          weight /= total_weight;
          if (!push_op(pl, r, TS_OP_ADD, scb[r], weight))
            return MXR_ERR_INVALID_ARG;
          scb[r] += 1;
        }
      } else if (r > 0 && weight != 1) {
        weight = 1;
        if (!push_op(pl, r, TS_OP_LOAD, scb[r] - 1, weight))
          return MXR_ERR_INVALID_ARG;
      }
    }
    R *= Rf;
  }
  pl->weight_end = weight;
  return MXR_OK;
}

// RtoM (find_HII_bubbles.c:27-46) with the internal units of set_units (init.c:203-225)
double ts_rtom(const mxr_ts_params* p, double R)
{
  const double UT = p->UnitLength_in_cm / p->UnitVelocity_in_cm_per_s;
  const double G = GRAVITY / std::pow(p->UnitLength_in_cm, 3) * p->UnitMass_in_g * std::pow(UT, 2);
  const double Hub = HUBBLE * UT;
  const double RhoCrit = 3 * Hub * Hub / (8 * M_PI * G);
  if (p->ReionRtoMFilterType == 0)
    return (4.0 / 3.0) * M_PI * std::pow(R, 3) * (p->OmegaM * RhoCrit);
  return std::pow(2 * M_PI, 1.5) * p->OmegaM * RhoCrit * std::pow(R, 3);
}

// ---- nu_tau_one and the frequency integrals (XRayHeatingFunctions.c:262-424, 742-856) -------
struct TauCtx {
  const TsCosmo* cz;
  const mxr_ts_params* p;
  const double* stored_fcoll;
  double x_e_ave;
};

static double interpolate_fcoll(const TauCtx& t, double redshift, int snap_i)
{
  const double* ZZ = t.p->ZZ;
  double f;
  if (snap_i == 0)
    f = t.stored_fcoll[snap_i];
  else
    f = t.stored_fcoll[snap_i - 1] + (redshift - ZZ[snap_i - 1]) * (t.stored_fcoll[snap_i] - t.stored_fcoll[snap_i - 1]) /
                                       (ZZ[snap_i] - ZZ[snap_i - 1]);
  return f < 0.0 ? 0.0 : f;
}

static double tauX(const TauCtx& t, double nu, double x_e, double zp, double zpp, int snap_i)
{
  const double nu_0 = nu / (1 + zp);
  const double ion_eff = t.p->ReionEfficiency;
  auto integrand = [&](double zhat) {
    const double drpropdz = SPEED_OF_LIGHT * t.cz->dtdz((float)zhat);
    const double n = t.cz->N_b0 * std::pow(1 + zhat, 3);
    const double nuhat = nu_0 * (1 + zhat);
    const double fcoll = interpolate_fcoll(t, zhat, snap_i);
    double ff = fcoll < 1e-20 ? 1 : 1 - (ion_eff * fcoll) / (1.0 - t.x_e_ave);
    if (ff < 1e-4)
      ff = 1e-4;
    const double sig = t.cz->f_H * (1 - x_e) * HI_ion_crosssec(nuhat) + t.cz->f_He * (1 - x_e) * HeI_ion_crosssec(nuhat) +
                       t.cz->f_He * x_e * HeII_ion_crosssec(nuhat);
    return drpropdz * n * ff * sig;
  };
  double result, err;
  if (qag_gsl(integrand, zpp, zp, 0, 0.005, 1000, &result, &err) != 0)
    return 1e10;
  return result;
}

static double nu_tau_one(const TauCtx& t, double zp, double zpp, double x_e, int snap_i)
{
  if (x_e > 0.9999)
    return -1;
  if (tauX(t, HeI_NUIONIZATION, x_e, zp, zpp, snap_i) < 1)
    return HeI_NUIONIZATION;
  auto f = [&](double nu) { return tauX(t, nu, x_e, zp, zpp, snap_i) - 1; };
  Brent s;
  double x_lo = HeI_NUIONIZATION, x_hi = 1e6 * NU_over_EV, r = 0;
  brent_init(s, f, x_lo, x_hi);  // the reference ignores the bracketing status and iterates
  int iter = 0;
  bool done;
  do {
    iter++;
    brent_iterate(s, f, &r, &x_lo, &x_hi);
    done = root_test_interval(x_lo, x_hi, 0, 0.02);
  } while (!done && iter < 100);
  return r;
}

static double integrate_over_nu(const TsData* d, const TsCosmo& cz, double zp, double local_x_e, double lo, double hi,
                                double thresh_energy, double spec_index, int flag)
{
  if (lo >= hi)
    return 0.0;
  const double xe = local_x_e;
  const float xf = (float)xe;
  const double f_H = cz.f_H, f_He = cz.f_He;
  auto powfac = [&](double nu) { return std::pow(nu / (thresh_energy * NU_over_EV), -spec_index - 1); };
  auto heat = [&](double nu) {
    double s = interp_tab(d, d->fheat, 1.0f, (float)((nu - NUIONIZATION) / NU_over_EV), xf) * PLANCK *
               (nu - NUIONIZATION) * f_H * (1 - xe) * HI_ion_crosssec(nu);
    s += interp_tab(d, d->fheat, 1.0f, (float)((nu - HeI_NUIONIZATION) / NU_over_EV), xf) * PLANCK *
         (nu - HeI_NUIONIZATION) * f_He * (1 - xe) * HeI_ion_crosssec(nu);
    s += interp_tab(d, d->fheat, 1.0f, (float)((nu - HeII_NUIONIZATION) / NU_over_EV), xf) * PLANCK *
         (nu - HeII_NUIONIZATION) * f_He * xe * HeII_ion_crosssec(nu);
    return s * powfac(nu);
  };
  auto Fi = [&](double nu, double thr) -> float {
    const float En = (float)((nu - thr) / NU_over_EV);
    float a = interp_tab(d, d->nHI, 0.0f, En, xf);
    a += interp_tab(d, d->nHeI, 0.0f, En, xf);
    a += interp_tab(d, d->nHeII, 0.0f, En, xf);
    a += 1;
    return a;
  };
  auto ion = [&](double nu) {
    double s = Fi(nu, NUIONIZATION) * f_H * (1 - xe) * HI_ion_crosssec(nu);
    s += Fi(nu, HeI_NUIONIZATION) * f_He * (1 - xe) * HeI_ion_crosssec(nu);
    s += Fi(nu, HeII_NUIONIZATION) * f_He * xe * HeII_ion_crosssec(nu);
    return s * powfac(nu);
  };
  auto lya = [&](double nu) {
    double s = interp_tab(d, d->nLya, 0.0f, (float)((nu - NUIONIZATION) / NU_over_EV), xf) * f_H * (1 - xe) *
               HI_ion_crosssec(nu);
    s += interp_tab(d, d->nLya, 0.0f, (float)((nu - HeI_NUIONIZATION) / NU_over_EV), xf) * f_He * (1 - xe) *
         HeI_ion_crosssec(nu);
    s += interp_tab(d, d->nLya, 0.0f, (float)((nu - HeII_NUIONIZATION) / NU_over_EV), xf) * f_He * xe *
         HeII_ion_crosssec(nu);
    return s * powfac(nu);
  };
  double result = 0, err = 0;
  int st;
  if (flag == 0)
    st = qag_gsl(heat, lo, hi, 0, 0.01, 1000, &result, &err);
  else if (flag == 1)
    st = qag_gsl(ion, lo, hi, 0, 0.01, 1000, &result, &err);
  else
    st = qag_gsl(lya, lo, hi, 0, 0.01, 1000, &result, &err);
  if (st != 0)
    return 0.0;
  if (flag == 2)
    return result * SPEED_OF_LIGHT / (4 * M_PI) / NU_LA / cz.hubble((float)zp);
  return result;
}

// ---- the device block -----------------------------------------------------------------------
static void fill_common(const mxr_ts_params* p, const TsCosmo& cz, const TsData* d, double zp, TsDev* dev)
{
  const int N = p->ReionGridDim;
  const double box_size = p->BoxSize / p->Hubble_h;
  dev->zp = zp;
  dev->zpf = (float)zp;
  const float zpf1 = 1 + dev->zpf;
  const double zd = (double)dev->zpf;
  dev->pv = std::pow(box_size / (double)N, 3);
  const double UT = p->UnitLength_in_cm / p->UnitVelocity_in_cm_per_s;
  dev->UMUT = p->UnitMass_in_g / UT;
  dev->ULm3 = std::pow(p->UnitLength_in_cm, -3.);
  dev->No = cz.No;
  dev->N_b0 = cz.N_b0;
  dev->f_H = cz.f_H;
  dev->f_He = cz.f_He;
  dev->h = p->Hubble_h;
  dev->HUBBLEh = HUBBLE * p->Hubble_h;
  dev->P2 = std::pow((double)zpf1, 2);
  dev->nb3 = std::pow((double)zpf1, 3);
  dev->Trad = TCMB * (1.0 + zd);
  dev->Trad4 = std::pow(dev->Trad, 4.0);
  dev->opz3 = std::pow(1.0 + zd, 3.0);
  dev->hub_z = cz.hubble(dev->zpf);
  dev->hub_ratio = cz.hubble(dev->zpf) / (HUBBLE * p->Hubble_h);
  dev->conv = (SPEED_OF_LIGHT / (4. * M_PI)) / (PROTONMASS / SOLAR_MASS);
  dev->iTrad = 1.0 / dev->Trad;
  dev->c_tgp = 1.342881e-7 / dev->hub_z * dev->No * dev->opz3;
  dev->c_xa = 1.66e11 / (1.0 + zd);
  dev->c_comp = (-1.51e-4) / dev->hub_ratio / p->Hubble_h * dev->Trad4 / (1.0 + zd);
  dev->c_adia = 3 / (1.0 + zd);
  const int kn[3] = {KAP_N, 20, 17};
  const CSpline* sp[3] = {&d->kap10, &d->kapeH, &d->kappH};
  for (int s = 0; s < 3; s++) {
    dev->kn[s] = kn[s];
    for (int i = 0; i < kn[s]; i++) {
      dev->kx[s][i] = sp[s]->x[i];
      dev->ky[s][i] = sp[s]->y[i];
      dev->kc[s][i] = sp[s]->c[i];
    }
  }
  dev->k10_E1 = std::exp(d->kap10_y[KAP_N - 1]);
  dev->k10_E2 = std::exp(d->kap10_x[KAP_N - 1]);
  std::memcpy(dev->xint, d->x_int_XHII, sizeof(dev->xint));
}

int ts_homogeneous(const mxr_ts_params* p, TsDev* dev)
{
  TsData* d = nullptr;
  if (int rc = load_data(p->TablesForXHeatingDir, &d))
    return rc;
  TsCosmo cz;
  cz.init(p);
  std::memset(dev, 0, sizeof(*dev));
  const double zp = p->ZZ[p->snapshot];
  fill_common(p, cz, d, zp, dev);
  const float zf = (float)zp;
  if (zf > d->recfast_T.x.back()) {
    set_error("T_RECFAST / xion_RECFAST called with z=%f beyond the table", (double)zf);
    return MXR_ERR_TABLES;
  }
  dev->xe_h = (float)d->recfast_x.eval(zf);
  dev->TK_h = (float)d->recfast_T.eval(zf);
  dev->cT_ad = cT_approx(zf);
  return MXR_OK;
}

int ts_tables(const mxr_ts_params* p, const TsPlan& pl, double collapse_fraction, double x_e_ave,
              const double* stored_fcoll, TsDev* dev, mxr_ts_tables* out)
{
  TsData* d = nullptr;
  if (int rc = load_data(p->TablesForXHeatingDir, &d))
    return rc;
  TsCosmo cz;
  cz.init(p);
  const int F = pl.F;
  const double zp = pl.zp;
  std::memset(dev, 0, sizeof(*dev));
  fill_common(p, cz, d, zp, dev);
  dev->F = F;
  dev->no_light = collapse_fraction > 0.0 ? 0 : 1;
  dev->dzp = pl.dzp;

  {
    std::lock_guard<std::mutex> lk(g_data_mu);
    normalise_spectra(d, p->ReionNionPhotPerBary);
  }

  // per-shell tables (ComputeTs.c:761-1011): nu_tau_one, then 14 x_e bins x {heat, ion, lya}
  double filling = 1. - p->ReionEfficiency * collapse_fraction / (1.0 - x_e_ave);
  const double filling_raw = filling;
  TauCtx tc{&cz, p, stored_fcoll, x_e_ave};
  std::vector<double> ntau(F), heat((size_t)XNH * F), ion((size_t)XNH * F), lya((size_t)XNH * F);
  const double lo_thr = p->NuXrayThreshold * NU_over_EV, hi_lim = p->NuXrayMax * NU_over_EV;
  // phase 1: nu_tau_one per shell; phase 2: the integrals per distinct lower limit and x_e bin
  // (shells whose tau = 1 frequency lies below the threshold share one set of integrals --
  // the same function of the same arguments, so computing it once changes nothing)
  auto parallel = [](int ntask, auto&& fn) {
    std::atomic<int> next{0};
    auto worker = [&]() {
      for (int k = next.fetch_add(1); k < ntask; k = next.fetch_add(1))
        fn(k);
    };
    unsigned nth = std::max(1u, std::min(std::thread::hardware_concurrency(), 16u));
    nth = std::min<unsigned>(nth, (unsigned)std::max(ntask, 1));
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < nth; t++)
      pool.emplace_back(worker);
    worker();
    for (auto& th : pool)
      th.join();
  };
  parallel(F, [&](int r) {
    const double prev_zpp = r == 0 ? zp : pl.zpp_edge[r - 1];
    const double zpp = (pl.zpp_edge[r] + prev_zpp) * 0.5;
    ntau[r] = nu_tau_one(tc, zp, zpp, x_e_ave, p->snapshot);
  });
  std::vector<double> lowers;
  std::vector<int> which(F);
  for (int r = 0; r < F; r++) {
    const double lower = std::fmax(ntau[r], lo_thr);
    int u = 0;
    while (u < (int)lowers.size() && lowers[u] != lower)
      u++;
    if (u == (int)lowers.size())
      lowers.push_back(lower);
    which[r] = u;
  }
  const int nu_ = (int)lowers.size();
  std::vector<double> uh((size_t)nu_ * XNH), ui((size_t)nu_ * XNH), ul((size_t)nu_ * XNH);
  parallel(nu_ * XNH, [&](int k) {
    const int u = k / XNH, m = k % XNH;
    const double xm = d->x_int_XHII[m];
    uh[k] = integrate_over_nu(d, cz, zp, xm, lowers[u], hi_lim, p->NuXrayThreshold, p->SpecIndexXrayGal, 0);
    ui[k] = integrate_over_nu(d, cz, zp, xm, lowers[u], hi_lim, p->NuXrayThreshold, p->SpecIndexXrayGal, 1);
    ul[k] = integrate_over_nu(d, cz, zp, xm, lowers[u], hi_lim, p->NuXrayThreshold, p->SpecIndexXrayGal, 2);
  });
  for (int r = 0; r < F; r++)
    for (int m = 0; m < XNH; m++) {
      const size_t k = (size_t)which[r] * XNH + m;
      heat[(size_t)m * F + r] = uh[k];
      ion[(size_t)m * F + r] = ui[k];
      lya[(size_t)m * F + r] = ul[k];
    }
  // AGN bands (ComputeTs.c:835-899): the limits scale with (1 + zp) / (1 + zpp), so every
  // shell has its own; soft = NuXrayThreshold..NuXraySoftCut, hard = NuXraySoftCut..NuXrayMax,
  // both floored at the shell's tau = 1 frequency
  const int agn_on[2] = {pl.agn_soft, pl.agn_hard};
  const double agn_alpha[2] = {p->SpecIndexXrayAGNSoft, p->SpecIndexXrayAGNHard};
  std::vector<double> agn_tab[2][3];
  for (int b = 0; b < 2; b++)
    for (int k = 0; k < 3; k++)
      agn_tab[b][k].assign((size_t)XNH * F, 0.0);
  {
    std::vector<int> tasks;  // (band, shell, bin) as band * F * XNH + r * XNH + m
    for (int b = 0; b < 2; b++)
      if (agn_on[b])
        for (int i = 0; i < F * XNH; i++)
          tasks.push_back(b * F * XNH + i);
    parallel((int)tasks.size(), [&](int t) {
      const int b = tasks[t] / (F * XNH), r = (tasks[t] / XNH) % F, m = tasks[t] % XNH;
      const double prev_zpp = r == 0 ? zp : pl.zpp_edge[r - 1];
      const double zpp = (pl.zpp_edge[r] + prev_zpp) * 0.5;
      double lo, hi;
      if (b == 0) {
        lo = std::fmax(ntau[r], p->NuXrayThreshold * NU_over_EV * (1. + zp) / (1. + zpp));
        hi = p->NuXraySoftCut * NU_over_EV * (1. + zp) / (1. + zpp);
      } else {
        lo = std::fmax(ntau[r], p->NuXraySoftCut * NU_over_EV * (1. + zp) / (1. + zpp));
        hi = p->NuXrayMax * NU_over_EV * (1. + zp) / (1. + zpp);
      }
      const double xm = d->x_int_XHII[m];
      for (int k = 0; k < 3; k++)
        agn_tab[b][k][(size_t)m * F + r] =
          integrate_over_nu(d, cz, zp, xm, lo, hi, p->NuXrayThreshold, agn_alpha[b], k);
    });
  }
  if (filling < 0)
    filling = 0;

  // sum over Lyman-n resonances (:918-954) and the kink correction (:966-1011)
  double sum_lyn[TS_MAXF];
  bool first_radii = true, first_zero = true;
  double weight = pl.weight_end;
  const int n_pts_radii = 1000;
  for (int r = 0; r < F; r++) {
    const double prev_zpp = r == 0 ? zp : pl.zpp_edge[r - 1];
    const double prev_R = r == 0 ? 0 : pl.R[r - 1];
    const double zpp = (pl.zpp_edge[r] + prev_zpp) * 0.5;
    sum_lyn[r] = 0;
    for (int n_ct = NSPEC_MAX; n_ct >= 2; n_ct--) {
      if (zpp > zmax((float)zp, n_ct))
        continue;
      const double nuprime = nu_n(n_ct) * (1 + zpp) / (1.0 + zp);
      sum_lyn[r] += frecycle(n_ct) * spectral_emissivity(d, nuprime);
    }
    if (r > 2 && sum_lyn[r] == 0.0 && sum_lyn[r - 1] > 0. && first_radii) {
      const double trial_zpp_max = (prev_zpp - (pl.R[r] - prev_R) * MPC / cz.drdz((float)prev_zpp) + prev_zpp) * 0.5;
      const double trial_zpp_min = (pl.zpp_edge[r - 2] -
                                    (pl.R[r - 1] - pl.R[r - 2]) * MPC / cz.drdz((float)pl.zpp_edge[r - 2]) +
                                    pl.zpp_edge[r - 2]) *
                                   0.5;
      for (int ii = 0; ii < n_pts_radii; ii++) {
        const double trial_zpp =
          trial_zpp_min + (trial_zpp_max - trial_zpp_min) * (float)ii / ((float)n_pts_radii - 1.);
        int counter = 0;
        for (int n_ct = NSPEC_MAX; n_ct >= 2; n_ct--) {
          if (trial_zpp > zmax((float)zp, n_ct))
            continue;
          counter += 1;
        }
        if (counter == 0 && first_zero) {
          first_zero = false;
          weight = (float)ii / (float)n_pts_radii;
        }
      }
      sum_lyn[r] = weight * sum_lyn[r - 1];
      first_radii = false;
    }
  }

  // prefactors (:1013-1128)
  const double growth = cz.dicke(zp);
  const double dgrowth = cz.ddicke_dz(zp);
  if (growth < 0) {
    set_error("ComputeTs: no growth function for these cosmological parameters");
    return MXR_ERR_UNSUPPORTED;
  }
  const double dt_dzp = cz.dtdz((float)zp);
  const double a = p->SpecIndexXrayGal;
  double L;
  if (std::fabs(a - 1.0) < REL_TOL) {
    L = (p->NuXrayThreshold * NU_over_EV) * std::log(p->NuXraySoftCut / p->NuXrayThreshold);
    L = 1. / L;
  } else {
    L = std::pow(p->NuXraySoftCut * NU_over_EV, 1. - a) - std::pow(p->NuXrayThreshold * NU_over_EV, 1. - a);
    L = 1. / L;
    L *= std::pow(p->NuXrayThreshold * NU_over_EV, -a) * (1 - a);
  }
  L *= 1.0 / PLANCK;
  const double prefac = L / (p->NuXrayThreshold * NU_over_EV) * SPEED_OF_LIGHT * std::pow(1 + zp, a + 3);
  // the AGN bands' conversions C = L0 / L_band, / PLANCK, and A_b = C / (h nu_th) c (1 + zp)^(alpha_b + 3),
  // computed (and checked) whatever the flag (ComputeTs.c:1045-1128)
  double prefac_agn[2];
  for (int b = 0; b < 2; b++) {
    const double ab = agn_alpha[b];
    const double nlo = b == 0 ? p->NuXrayThreshold : p->NuXraySoftCut;
    const double nhi = b == 0 ? p->NuXraySoftCut : p->NuXrayMax;
    double C;
    if (std::fabs(ab - 1.0) < REL_TOL) {
      C = (p->NuXrayThreshold * NU_over_EV) * std::log(nhi / nlo);
      C = 1.0 / C;
    } else {
      C = std::pow(nhi * NU_over_EV, 1.0 - ab) - std::pow(nlo * NU_over_EV, 1.0 - ab);
      C = 1.0 / C;
      C *= std::pow(p->NuXrayThreshold * NU_over_EV, -ab) * (1.0 - ab);
    }
    C /= (PLANCK);
    prefac_agn[b] = C / (p->NuXrayThreshold * NU_over_EV) * SPEED_OF_LIGHT * std::pow(1.0 + zp, ab + 3.0);
    if (!std::isfinite(prefac_agn[b])) {
      set_error("const_zp_prefactor_AGN_%s is not finite: check NuXrayThreshold / NuXraySoftCut / NuXrayMax / "
                "SpecIndexXrayAGN%s",
                b == 0 ? "soft" : "hard", b == 0 ? "Soft" : "Hard");
      return MXR_ERR_INVALID_ARG;
    }
  }
  dev->growth = growth;
  dev->dgrowth = dgrowth;
  dev->dt_dzp = dt_dzp;
  dev->c_heat = dt_dzp * 2.0 / 3.0 / BOLTZMANN;
  dev->prefac_gal = prefac;

  // evolveInt's per-shell constants with its float zp (XRayHeatingFunctions.c:1378-1433),
  // folded with the unit conversions of the smoothed fields (ComputeTs.c:596-602, 704-711)
  const double zpf = (double)dev->zpf;
  dev->cx = XRAY_LUMINOSITY_UNIT * dev->ULm3 / dev->pv;
  const double csfr = dev->UMUT * dev->ULm3 / SOLAR_MASS / dev->pv;
  dev->cagn = 1e10 * SOLAR_LUM / dev->pv * dev->ULm3;  // ComputeTs.c:617-618
  dev->nagn = 0;
  for (int b = 0; b < 2; b++)
    if (agn_on[b])
      dev->agn_kind[dev->nagn++] = b;
  for (int r = 0; r < F; r++) {
    double zpp, dzpp;
    if (r == 0) {
      zpp = (pl.zpp_edge[0] + zpf) * 0.5;
      dzpp = zpf - pl.zpp_edge[0];
    } else {
      zpp = (pl.zpp_edge[r] + pl.zpp_edge[r - 1]) * 0.5;
      dzpp = pl.zpp_edge[r - 1] - pl.zpp_edge[r];
    }
    const double dt_dzpp = cz.dtdz((float)zpp);
    const double wx = dt_dzpp * dzpp * std::pow(1 + zpp, -a) * dev->cx;
    dev->bs[r] = csfr * dev->P2 * (1 + zpp) * sum_lyn[r] * dt_dzpp * dzpp;
    for (int m = 0; m < XNH; m++) {
      const int mh = m + 1 < XNH ? m + 1 : m;
      const double dx = (double)(d->x_int_XHII[mh] - d->x_int_XHII[m]);  // float difference
      const double* tb[3] = {heat.data(), ion.data(), lya.data()};
      for (int k = 0; k < 3; k++) {
        const double lo = tb[k][(size_t)m * F + r], hi = tb[k][(size_t)mh * F + r];
        dev->fold[2 * k][m][r] = wx * (mh != m ? (hi - lo) / dx : 0.0);
        dev->fold[2 * k + 1][m][r] = wx * lo;
      }
      for (int j = 0; j < dev->nagn; j++) {
        const int b = dev->agn_kind[j];
        const double wb = dt_dzpp * dzpp * std::pow(1 + zpp, -agn_alpha[b]) * dev->cagn * prefac_agn[b];
        for (int k = 0; k < 3; k++) {
          const double lo = agn_tab[b][k][(size_t)m * F + r], hi = agn_tab[b][k][(size_t)mh * F + r];
          dev->fold_agn[j][2 * k][m][r] = wb * (mh != m ? (hi - lo) / dx : 0.0);
          dev->fold_agn[j][2 * k + 1][m][r] = wb * lo;
        }
      }
    }
  }

  // x_e_ave, collapse fraction and NO_LIGHT enter nothing else on the device
  if (out) {
    std::memset(out, 0, sizeof(*out));
    out->n = F;
    out->homogeneous = 0;
    out->NO_LIGHT = dev->no_light;
    out->zp = zp;
    out->dzp = pl.dzp;
    out->dt_dzp = dt_dzp;
    out->growth_factor_zp = growth;
    out->dgrowth_factor_dzp = dgrowth;
    out->const_zp_prefactor_GAL = prefac;
    out->const_zp_prefactor_AGN_soft = prefac_agn[0];
    out->const_zp_prefactor_AGN_hard = prefac_agn[1];
    out->collapse_fraction = collapse_fraction;
    out->x_e_ave = x_e_ave;
    out->filling_factor_of_HI_zp = filling_raw;
    for (int r = 0; r < F; r++) {
      out->R[r] = pl.R[r];
      out->zpp_edge[r] = pl.zpp_edge[r];
      out->zpp[r] = pl.zpp[r];
      out->nu_tau_one[r] = ntau[r];
      out->sum_lyn[r] = sum_lyn[r];
      for (int m = 0; m < XNH; m++) {
        out->freq_int_heat[m][r] = heat[(size_t)m * F + r];
        out->freq_int_ion[m][r] = ion[(size_t)m * F + r];
        out->freq_int_lya[m][r] = lya[(size_t)m * F + r];
        out->freq_int_heat_AGN_soft[m][r] = agn_tab[0][0][(size_t)m * F + r];
        out->freq_int_ion_AGN_soft[m][r] = agn_tab[0][1][(size_t)m * F + r];
        out->freq_int_lya_AGN_soft[m][r] = agn_tab[0][2][(size_t)m * F + r];
        out->freq_int_heat_AGN_hard[m][r] = agn_tab[1][0][(size_t)m * F + r];
        out->freq_int_ion_AGN_hard[m][r] = agn_tab[1][1][(size_t)m * F + r];
        out->freq_int_lya_AGN_hard[m][r] = agn_tab[1][2][(size_t)m * F + r];
      }
      out->n_ops[r] = pl.nops[r];
      for (int k = 0; k < pl.nops[r]; k++) {
        out->op_kind[r][k] = pl.ops[r][k].kind;
        out->op_hist[r][k] = pl.ops[r][k].hist;
        out->op_weight[r][k] = pl.ops[r][k].wd;
      }
    }
    std::memcpy(out->x_int_XHII, d->x_int_XHII, sizeof(out->x_int_XHII));
    out->hist_max = pl.hist_max;
  }
  return MXR_OK;
}

}  // namespace mxr
