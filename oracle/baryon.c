/*
 * baryon.c -- CPU restatement of construct_baryon_grids for one rank (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ and bench.py's cpu_baseline leg as the checker.
 *
 * Follows src/core/reionization.c:2025-2591 without stochasticity or mini-haloes, with
 * pos_to_ngp (misc_tools.c:207-217), calc_metallicity (misc_tools.c:45-57),
 * galaxy_xray_band_fraction and calculate_galaxy_xray_luminosity (reionization.c:1967-2023).
 * One property at a time, as the reference loops: a float buffer per property accumulated in
 * galaxy order, then the per-property finalisation into the padded grids.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define O_SOLAR_MASS 1.989e33          /* src/meraxes.h:22 */
#define O_SEC_PER_YEAR 3.155e7         /* src/meraxes.h:34 */
#define O_XRAY_LUMINOSITY_UNIT 1.0e40  /* src/meraxes.h:36 */
#define O_REL_TOL 1e-5                 /* src/meraxes.h REL_TOL */

/* misc_tools.c:207-217 */
static int pos_to_ngp(double x, double side, int nx)
{
  int ind = (int)nearbyint(x / side * (double)nx);
  if (ind > nx - 1)
    ind = 0;
  return ind;
}

/* misc_tools.c:45-57 */
static double calc_metallicity(double total_gas, double metals)
{
  double Z;
  if ((total_gas > 0) && (metals > 0))
    Z = metals / total_gas;
  else
    Z = 0.0;
  if (Z < 0)
    Z = 0;
  if (Z > 1)
    Z = 1;
  return Z;
}

/* reionization.c:1967-1980 */
static double band_fraction(const mxr_baryon_params* p)
{
  double e_min = p->NuXrayThreshold, e_max = p->NuXraySoftCut, alpha = p->SpecIndexXrayGal;
  if (fabs(alpha - 1.0) < O_REL_TOL)
    return log(e_max / e_min) / log(2000.0 / 500.0);
  double exponent = 1.0 - alpha;
  return (pow(e_max, exponent) - pow(e_min, exponent)) / (pow(2000.0, exponent) - pow(500.0, exponent));
}

/* reionization.c:1982-2023, scatter_dex = 0 */
static double xray_luminosity(const mxr_baryon_params* p, const mxr_galaxies* g, long i)
{
  double source_sfr = p->Flag_InstantaneousSFR ? g->Sfr[i] : g->GrossStellarMass[i] / p->sfr_timescale;
  source_sfr *= p->UnitMass_in_g / O_SOLAR_MASS;
  source_sfr *= O_SEC_PER_YEAR / p->UnitTime_in_s;
  if (source_sfr <= 0.0)
    return 0.0;
  double log_lx_sfr, bf = 1.0;
  if (p->Flag_XrayMetallicity) {
    double metallicity = calc_metallicity(g->ColdGas[i], g->MetalsColdGas[i]);
    double z_solar = metallicity / p->XrayMetallicitySolar;
    double oxygen_abundance = p->XrayOxygenAbundanceFactor * z_solar;
    log_lx_sfr = p->XrayLogLNorm + p->XrayMetallicitySlope * oxygen_abundance;
    bf = band_fraction(p);
  } else {
    if (p->LXrayGal <= 0.0)
      return 0.0;
    log_lx_sfr = log10(p->LXrayGal);
  }
  return pow(10.0, log_lx_sfr) * bf * source_sfr / O_XRAY_LUMINOSITY_UNIT;
}

enum { P_STELLAR, P_BHM, P_BHAR, P_WSFR, P_XRAY, P_SFR, P_HARD, P_SOFT, P_N };

int oracle_construct_baryon_grids(const mxr_baryon_params* p, const mxr_galaxies* g, const mxr_baryon_grids* out,
                                  long* n_below_bh_limit)
{
  const int N = p->ReionGridDim;
  const long H = N / 2 + 1, npad = (long)N * N * 2 * H, nreal = (long)N * N * N;
  const int spin = p->Flag_IncludeSpinTemp, nstore = p->NstoreSnapshots_Heating;
  const int hard = p->Flag_IncludeAGNXray == 1 || p->Flag_IncludeAGNXray == 2;
  const int soft = p->Flag_IncludeAGNXray == 1 || p->Flag_IncludeAGNXray == 3;
  long below = 0;

  /* :2093-2141 init and history shift */
  for (long ii = 0; ii < npad; ii++) {
    out->stars[ii] = 0.0f;
    out->weighted_sfr[ii] = 0.0f;
  }
  if (spin) {
    for (long ii = 0; ii < npad; ii++) {
      out->sfr[ii] = 0.0f;
      for (int snap = nstore - 2; snap >= 0; snap--)
        out->sfr_histories[(snap + 1) * npad + ii] = out->sfr_histories[snap * npad + ii];
      out->xray_luminosity[ii] = 0.0f;
      for (int snap = nstore - 2; snap >= 0; snap--)
        out->xray_luminosity_histories[(snap + 1) * npad + ii] = out->xray_luminosity_histories[snap * npad + ii];
      if (hard) {
        out->BHXrayEmissivity_hard[ii] = 0.0f;
        for (int snap = nstore - 2; snap >= 0; snap--)
          out->bh_xray_histories_hard[(snap + 1) * npad + ii] = out->bh_xray_histories_hard[snap * npad + ii];
      }
      if (soft) {
        out->BHXrayEmissivity_soft[ii] = 0.0f;
        for (int snap = nstore - 2; snap >= 0; snap--)
          out->bh_xray_histories_soft[(snap + 1) * npad + ii] = out->bh_xray_histories_soft[snap * npad + ii];
      }
    }
  }

  float* buffer = (float*)malloc(sizeof(float) * nreal);
  if (!buffer)
    return MXR_ERR_OOM;
  for (int prop = 0; prop < P_N; prop++) {
    if (!spin && (prop == P_SFR || prop == P_XRAY))
      continue;
    if (prop == P_HARD && (!spin || !hard))
      continue;
    if (prop == P_SOFT && (!spin || !soft))
      continue;
    if (!p->Flag_BHFeedback && (prop == P_BHM || prop == P_BHAR))
      continue;
    memset(buffer, 0, sizeof(float) * nreal);
    for (long i = 0; i < g->n; i++) {
      if (g->Type[i] > 2)
        continue;
      const int ix = pos_to_ngp(g->Pos[3 * i + 0], p->BoxSize, N);
      const int iy = pos_to_ngp(g->Pos[3 * i + 1], p->BoxSize, N);
      const int iz = pos_to_ngp(g->Pos[3 * i + 2], p->BoxSize, N);
      if (ix < 0 || iy < 0 || iz < 0) {
        free(buffer);
        return MXR_ERR_INVALID_ARG;
      }
      const long ind = iz + (long)N * (iy + (long)N * ix);
      switch (prop) {
        case P_STELLAR: buffer[ind] += g->FescWeightedGSM[i]; break;
        case P_BHM:
          if (g->BlackHoleMass[i] >= p->BlackHoleMassLimitReion)
            buffer[ind] += g->EffectiveBHM[i];
          else
            below++;
          break;
        case P_WSFR: buffer[ind] += g->FescWeightedSfr[i]; break;
        case P_BHAR:
          if (g->BlackHoleMass[i] >= p->BlackHoleMassLimitReion)
            buffer[ind] += g->EffectiveBHAR[i];
          break;
        case P_XRAY: buffer[ind] += (float)xray_luminosity(p, g, i); break;
        case P_SFR: buffer[ind] += p->Flag_InstantaneousSFR ? g->Sfr[i] : g->GrossStellarMass[i]; break;
        case P_HARD:
          if (g->BlackHoleMass[i] >= p->BlackHoleMassLimitReion)
            buffer[ind] += g->BHXrayEmissivity_hard[i];
          break;
        case P_SOFT:
          if (g->BlackHoleMass[i] >= p->BlackHoleMassLimitReion)
            buffer[ind] += g->BHXrayEmissivity_soft[i];
          break;
      }
    }
    /* :2424-2569 finalisation into the padded grids */
    for (int ix = 0; ix < N; ix++)
      for (int iy = 0; iy < N; iy++)
        for (int iz = 0; iz < N; iz++) {
          const long ir = iz + (long)N * (iy + (long)N * ix), ip = iz + 2 * H * (iy + (long)N * ix);
          switch (prop) {
            case P_WSFR: {
              double val = (double)buffer[ir];
              if (val < 0)
                val = 0.0;
              out->weighted_sfr[ip] = (float)val;
              break;
            }
            case P_SFR: {
              double val = (double)buffer[ir];
              if (val < 0)
                val = 0.0;
              if (!p->Flag_InstantaneousSFR)
                val /= p->sfr_timescale;
              out->sfr[ip] = (float)val;
              out->sfr_histories[ip] = (float)val;
              break;
            }
            case P_XRAY: {
              const float val = buffer[ir];
              out->xray_luminosity[ip] = val;
              out->xray_luminosity_histories[ip] = val;
              break;
            }
            default: {
              float val = buffer[ir];
              if (val < 0)
                val = 0.0f;
              if (prop == P_STELLAR)
                out->stars[ip] = val;
              else if (prop == P_BHM)
                out->effective_bhm[ip] = val;
              else if (prop == P_BHAR)
                out->effective_bhar[ip] = val;
              else if (prop == P_HARD) {
                out->BHXrayEmissivity_hard[ip] = val;
                out->bh_xray_histories_hard[ip] = val;
              } else if (prop == P_SOFT) {
                out->BHXrayEmissivity_soft[ip] = val;
                out->bh_xray_histories_soft[ip] = val;
              }
            }
          }
        }
  }
  free(buffer);
  if (n_below_bh_limit)
    *n_below_bh_limit = below;
  return 0;
}
