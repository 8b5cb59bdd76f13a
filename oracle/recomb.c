/*
 * recomb.c -- CPU restatement of the inhomogeneous-recombination pieces of the reference's
 * find_HII_bubbles path.  TEST INFRASTRUCTURE ONLY.
 *
 *   src/core/recombinations.c:35-82   splined_recombination (per-cell table lookup)
 *   src/core/recombinations.c:397-403 Gamma_SS (Rahmati et al. 2013 self-shielding fit)
 *   src/core/recombinations.c:422-504 the four integrands over ln Delta (MHR_rr, MHR_cf_numerator,
 *                                     MHR_cf_denominator, MHR_rnh)
 *   src/core/recombinations.c:561-571 aux_function (the MHR00 PDF without its normalisation A)
 *   src/core/recombinations.c:619-700 C_MHR / beta_MHR (natural cubic splines over z = 2, 3, ...)
 *   src/core/recombinations.c:742-766 neutral_fraction, alpha_B
 *   src/core/reionization.h:15-31      No, the present-day hydrogen number density
 *   src/core/XRayHeatingFunctions.c:426-440 dtdz
 *
 * The tables themselves (init_MHR, recombinations.c:208-374: 200 x 100 x 150 GSL QAG integrals)
 * are not rebuilt here; the oracle's find_HII_bubbles takes the tables the library reports
 * (oracle_rec_set_tables), and the tests pin those tables separately against an independent
 * quadrature of the integrands below.  The splines over ln Gamma_12 are the oracle's own: a
 * natural cubic spline in its second-derivative form (the system solved by the Thomas
 * algorithm), which is the same interpolant as GSL's gsl_interp_cspline.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define HUBBLE 3.2407789e-18
#define GRAVITY 6.672e-8
#define PROTONMASS 1.6726e-24
#define TINY (double)(1e-30)
#define alphaB_10k (double)(2.59e-13)
#define FRACT_FLOAT_ERR (double)(1e-7)

/* recombinations.h:18-35 */
#define RR_Z_NPTS 200
#define RR_DEL_Z (float)(0.2)
#define RR_Z_END (float)(0.0)
#define RR_T_NPTS 100
#define RR_DEL_T (float)(0.03)
#define RR_T_STA (float)(2.0)
#define RR_lnGamma_NPTS 150
#define RR_lnGamma_min (float)(-10.0)
#define RR_DEL_lnGamma (float)(0.1)
#define RR_N ((long)RR_Z_NPTS * RR_T_NPTS * RR_lnGamma_NPTS)

/* ---------------------------------------------------------------- natural cubic spline */
/* Second derivatives M of the natural cubic spline through (x, y) (M_0 = M_{n-1} = 0). */
static void spline_M(const double* x, const double* y, int n, double* M, double* cp, double* dp)
{
  M[0] = M[n - 1] = 0.0;
  for (int i = 1; i < n - 1; i++) {
    const double h0 = x[i] - x[i - 1], h1 = x[i + 1] - x[i];
    const double a = h0, b = 2.0 * (h0 + h1), c = h1;
    const double r = 6.0 * ((y[i + 1] - y[i]) / h1 - (y[i] - y[i - 1]) / h0);
    const double den = b - (i > 1 ? a * cp[i - 1] : 0.0);
    cp[i] = c / den;
    dp[i] = (r - (i > 1 ? a * dp[i - 1] : 0.0)) / den;
  }
  for (int i = n - 2; i >= 1; i--)
    M[i] = dp[i] - ((i < n - 2) ? cp[i] * M[i + 1] : 0.0);
}

/* the interval x_lo <= xq < x_lo+1 by bisection (gsl_interp_bsearch), and the cubic on it */
static double spline_eval(const double* x, const double* y, const double* M, int n, double xq)
{
  int lo = 0, hi = n - 1;
  while (hi > lo + 1) {
    const int mid = (lo + hi) / 2;
    if (x[mid] > xq)
      hi = mid;
    else
      lo = mid;
  }
  const double h = x[lo + 1] - x[lo];
  const double t = xq - x[lo];
  const double b = (y[lo + 1] - y[lo]) / h - h * (2.0 * M[lo] + M[lo + 1]) / 6.0;
  const double c = M[lo] / 2.0;
  const double d = (M[lo + 1] - M[lo]) / (6.0 * h);
  return y[lo] + t * (b + t * (c + t * d));
}

/* ------------------------------------------------------------------------ tables */
static struct {
  int ok;
  double x[RR_lnGamma_NPTS];
  double* y[3];  /* ln rate, clumping, ln residual */
  double* M[3];
} rr;

int oracle_rec_set_tables(const double* lnGamma, const double* ln_rate, const double* clumping,
                          const double* ln_residual)
{
  const double* src[3] = {ln_rate, clumping, ln_residual};
  for (int q = 0; q < 3; q++) {
    if (!rr.y[q]) {
      rr.y[q] = (double*)malloc(sizeof(double) * RR_N);
      rr.M[q] = (double*)malloc(sizeof(double) * RR_N);
      if (!rr.y[q] || !rr.M[q])
        return MXR_ERR_INVALID_ARG;
    }
    memcpy(rr.y[q], src[q], sizeof(double) * RR_N);
  }
  memcpy(rr.x, lnGamma, sizeof rr.x);
#pragma omp parallel for schedule(static)
  for (long row = 0; row < 3L * RR_Z_NPTS * RR_T_NPTS; row++) {
    double cp[RR_lnGamma_NPTS], dp[RR_lnGamma_NPTS];
    const int q = (int)(row / (RR_Z_NPTS * RR_T_NPTS));
    const long off = (row % (RR_Z_NPTS * RR_T_NPTS)) * RR_lnGamma_NPTS;
    spline_M(rr.x, rr.y[q] + off, RR_lnGamma_NPTS, rr.M[q] + off, cp, dp);
  }
  rr.ok = 1;
  return 0;
}

/* recombinations.c:35-82 */
int oracle_splined_recombination(double z_eff, double gamma12_bg, double temp, double* recombination_rate,
                                 double* residual_xH, double* clumping_factor)
{
  if (!rr.ok)
    return MXR_ERR_TABLES;
  int z_ct = (int)((z_eff - RR_Z_END) / RR_DEL_Z + 0.5);
  int t_ct = (int)((log10(temp) - RR_T_STA) / RR_DEL_T + 0.5);
  double lnGamma = log(gamma12_bg);
  if (z_ct < 0)
    z_ct = 0;
  else if (z_ct >= RR_Z_NPTS)
    z_ct = RR_Z_NPTS - 1;
  if (t_ct < 0) {
    *recombination_rate = 0.;
    *residual_xH = 1e4;
    *clumping_factor = 1.0;
    return 0;
  } else if (t_ct >= RR_T_NPTS) {
    t_ct = RR_T_NPTS - 1;
  }
  if (lnGamma < RR_lnGamma_min) {
    *recombination_rate = 0.;
    *residual_xH = 1e4;
    *clumping_factor = 1.0;
    return 0;
  } else if (lnGamma >= (RR_lnGamma_min + RR_DEL_lnGamma * (RR_lnGamma_NPTS - 1))) {
    lnGamma = RR_lnGamma_min + RR_DEL_lnGamma * (RR_lnGamma_NPTS - 1) - FRACT_FLOAT_ERR;
  }
  const long off = ((long)z_ct * RR_T_NPTS + t_ct) * RR_lnGamma_NPTS;
  *recombination_rate = exp(spline_eval(rr.x, rr.y[0] + off, rr.M[0] + off, RR_lnGamma_NPTS, lnGamma));
  *residual_xH = exp(spline_eval(rr.x, rr.y[2] + off, rr.M[2] + off, RR_lnGamma_NPTS, lnGamma));
  *clumping_factor = spline_eval(rr.x, rr.y[1] + off, rr.M[1] + off, RR_lnGamma_NPTS, lnGamma);
  return 0;
}

/* ------------------------------------------------------------------ integrand pieces */
/* XRayHeatingFunctions.c:426-440 */
double oracle_dtdz(double OmegaM, double OmegaLambda, double Hubble_h, float z)
{
  double x = sqrt(OmegaLambda / OmegaM) * pow(1 + z, -3.0 / 2.0);
  double dxdz = sqrt(OmegaLambda / OmegaM) * pow(1 + z, -5.0 / 2.0) * (-3.0 / 2.0);
  double const1 = 2 * sqrt(1 + OmegaM / OmegaLambda) / (3.0 * HUBBLE * Hubble_h);
  double numer = dxdz * (1 + x * pow(pow(x, 2) + 1, -0.5));
  double denom = x + sqrt(pow(x, 2) + 1);
  return (const1 * numer / denom);
}

/* reionization.h:15-26 */
static double rec_No(const mxr_hii_params* p)
{
  const double rho = 3.0 * HUBBLE * HUBBLE * p->Hubble_h * p->Hubble_h / (8.0 * M_PI * GRAVITY);
  return rho * (p->BaryonFrac * p->OmegaM) * (1 - p->Y_He) / PROTONMASS;
}

static double alpha_B(double T) { return alphaB_10k * pow(T / 1.0e4, -0.75); }

/* recombinations.c:742-766 (case B) */
static double neutral_fraction(const mxr_hii_params* p, double density, double T4, double gamma12)
{
  double chi, b, alpha, corr_He = 1.0 / (4.0 / p->Y_He - 3);
  alpha = alpha_B(T4 * 1e4);
  gamma12 *= 1e-12;
  chi = (1 + corr_He) * density * alpha / gamma12;
  if (chi < TINY)
    return 0;
  if (chi < 1e-5)
    return chi;
  b = -2 - gamma12 / (density * (1 + corr_He) * alpha);
  chi = (-b - sqrt(b * b - 4)) / 2.0;
  return chi;
}

/* recombinations.c:397-403 */
static double Gamma_SS(double Gamma_bg, double Delta, double T_4, double z)
{
  double D_ss = 26.7 * pow(T_4, 0.17) * pow((1 + z) / 10.0, -3) * pow(Gamma_bg, 2.0 / 3.0);
  return Gamma_bg * (0.98 * pow((1.0 + pow(Delta / D_ss, 1.64)), -2.28) + 0.02 * pow(1.0 + Delta / D_ss, -0.84));
}

/* natural cubic spline over unit-spaced z nodes 2, 3, ... (recombinations.c:619-700) */
static double node_spline(const double* ys, int n, double z)
{
  double x[12], M[12], cp[12], dp[12];
  for (int i = 0; i < n; i++)
    x[i] = (float)i + 2.0;
  spline_M(x, ys, n, M, cp, dp);
  return spline_eval(x, ys, M, n, z);
}

static double C_MHR(double z)
{
  static const double C_table[12] = {0.558, 0.599, 0.611, 0.769, 0.868, 0.930,
                                     0.964, 0.983, 0.993, 0.998, 0.999, 1.00};
  if (z >= 13.0)
    return 1.0;
  if (z <= 2.0)
    return 0.558;
  return node_spline(C_table, 12, z);
}

static double beta_MHR(double z)
{
  static const double beta_table[5] = {-2.23, -2.35, -2.48, -2.49, -2.50};
  if (z >= 6.0)
    return -2.50;
  if (z <= 2.0)
    return -2.23;
  return node_spline(beta_table, 5, z);
}

/* Probes of the pieces, for the tests that pin them and for the tests' own quadrature of the
 * table integrals:
 *   0 neutral_fraction(density a0 [cm^-3], T4 a1, gamma12 a2)          -> out[0]
 *   1 Gamma_SS(Gamma_bg a0, Delta a1, T4 a2, z a3)                      -> out[0]
 *   2 C_MHR(z a0), beta_MHR(z a0)                                      -> out[0], out[1]
 *   3 aux_function(D a0; z a1): the PDF without A                       -> out[0]
 *   4 the four integrands at ln Delta a0 for z a1, gamma12 a2, T4 a3, A a4 (recombinations.c:
 *     422-504: MHR_rr, MHR_cf_numerator, MHR_cf_denominator, MHR_rnh)  -> out[0..3]
 *   5 No                                                                -> out[0]
 *   6 dtdz(float z a0)                                                  -> out[0] */
int oracle_rec_probe(const mxr_hii_params* p, int what, const double* a, double* out)
{
  switch (what) {
    case 0: out[0] = neutral_fraction(p, a[0], a[1], a[2]); return 0;
    case 1: out[0] = Gamma_SS(a[0], a[1], a[2], a[3]); return 0;
    case 2:
      out[0] = C_MHR(a[0]);
      out[1] = beta_MHR(a[0]);
      return 0;
    case 3: {
      const double D = a[0], z = a[1];
      out[0] = exp(-(pow(D, -2.0 / 3.0) - C_MHR(z)) * (pow(D, -2.0 / 3.0) - C_MHR(z)) /
                   (2.0 * (2.0 * 7.61 / (3.0 * (1.0 + z))) * (2.0 * 7.61 / (3.0 * (1.0 + z))))) *
               pow(D, beta_MHR(z));
      return 0;
    }
    case 4: {
      const double lnD = a[0], z = a[1], g12 = a[2], T4 = a[3], A = a[4];
      const double D = exp(lnD);
      const double gamma = Gamma_SS(g12, D, T4, z);
      const double n_H = rec_No(p) * pow(1 + z, 3) * D;
      const double x_HI = neutral_fraction(p, n_H, T4, gamma);
      const double x_e = 1.0 - x_HI;
      const double PDelta =
        A * exp(-0.5 * pow((pow(D, -2.0 / 3.0) - C_MHR(z)) / ((2.0 * 7.61 / (3.0 * (1.0 + z)))), 2)) *
        pow(D, beta_MHR(z));
      out[0] = n_H * PDelta * alpha_B(T4 * 1e4) * x_e * x_e * D * D;
      out[1] = PDelta * x_e * x_e * D * D * D;
      out[2] = PDelta * x_e * D * D;
      out[3] = 1e4 * PDelta * x_HI * D;
      return 0;
    }
    case 5: out[0] = rec_No(p); return 0;
    case 6: out[0] = oracle_dtdz(p->OmegaM, p->OmegaLambda, p->Hubble_h, (float)a[0]); return 0;
    default: return MXR_ERR_INVALID_ARG;
  }
}
