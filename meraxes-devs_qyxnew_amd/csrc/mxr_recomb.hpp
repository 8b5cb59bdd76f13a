// mxr_recomb.hpp -- the recombination tables' grid and the per-cell lookup splined_recombination
// (src/core/recombinations.c:35-100; grid src/core/recombinations.h:18-35), shared by the table
// builder (mxr_recomb.hip) and the post-loop kernel (mxr_hii.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/mxr.h"

namespace mxr {

struct Ctx;

// node grids: z = z_ct 0.2 (0 .. 39.8), log10 T = 2 + t_ct 0.03 (2 .. 4.97), ln Gamma_12 = -10 +
// g 0.1 (-10 .. 4.9), each node evaluated in float as the reference's constants are floats
constexpr int RR_Z_NPTS = 200, RR_T_NPTS = 100, RR_G_NPTS = 150;
constexpr float RR_DEL_Z = 0.2f, RR_Z_END = 0.0f, RR_DEL_T = 0.03f, RR_T_STA = 2.0f;
constexpr float RR_LNG_MIN = -10.0f, RR_DEL_LNG = 0.1f;

// device tables: tab[q][z_ct][t_ct][g] for q = 0 ln(recombination rate per baryon, 1/s),
// 1 clumping factor, 2 ln(1e4 x residual neutral fraction); coef the natural-spline
// coefficients of each (q, z_ct, t_ct) row in ln Gamma_12; x the ln Gamma_12 nodes
struct RecTab {
  const double* tab;
  const double* coef;
  const double* x;
  long n;  // entries per table
};

int rec_tables(Ctx* c, const struct mxr_hii_params* p, RecTab* out);
// the tables copied to host arrays (any may be null): the ln Gamma_12 nodes, then each table
int rec_tables_copy(Ctx* c, const struct mxr_hii_params* p, double* lnGamma, double* ln_rate, double* clumping,
                    double* ln_residual);

// splined_recombination for one cell: z_eff, Gamma_12 of the background in 1e-12/s, T in K
__device__ inline void splined_recombination(const RecTab& t, double z_eff, double gamma12_bg, double temp,
                                             double* rate, double* residual, double* clumping)
{
#pragma clang fp contract(off)
  int z_ct = (int)((z_eff - (double)RR_Z_END) / (double)RR_DEL_Z + 0.5);
  int t_ct = (int)((log10(temp) - (double)RR_T_STA) / (double)RR_DEL_T + 0.5);
  double lnGamma = log(gamma12_bg);
  if (z_ct < 0)
    z_ct = 0;
  else if (z_ct >= RR_Z_NPTS)
    z_ct = RR_Z_NPTS - 1;
  if (t_ct < 0 || lnGamma < (double)RR_LNG_MIN) {
    *rate = 0.;
    *residual = 1e4;
    *clumping = 1.0;
    return;
  }
  if (t_ct >= RR_T_NPTS)
    t_ct = RR_T_NPTS - 1;
  // the last node, RR_lnGamma_min + RR_DEL_lnGamma (NPTS - 1) in float, less FRACT_FLOAT_ERR
  const double top = t.x[RR_G_NPTS - 1];
  if (lnGamma >= top)
    lnGamma = top - 1e-7;
  // the node interval x_i <= lnGamma < x_{i+1} (the nodes are float-rounded, so checked)
  int i = (int)floor((lnGamma - t.x[0]) / (double)RR_DEL_LNG);
  i = i < 0 ? 0 : (i > RR_G_NPTS - 2 ? RR_G_NPTS - 2 : i);
  while (i > 0 && t.x[i] > lnGamma)
    i--;
  while (i < RR_G_NPTS - 2 && t.x[i + 1] <= lnGamma)
    i++;
  const long row = ((long)z_ct * RR_T_NPTS + t_ct) * RR_G_NPTS;
  const double h = t.x[i + 1] - t.x[i], dx = lnGamma - t.x[i];
  double v[3];
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const double* y = t.tab + q * t.n + row;
    const double* c = t.coef + q * t.n + row;
    const double b = (y[i + 1] - y[i]) / h - h * (c[i + 1] + 2.0 * c[i]) / 3.0;
    const double d = (c[i + 1] - c[i]) / (3.0 * h);
    v[q] = y[i] + dx * (b + dx * (c[i] + dx * d));
  }
  *rate = exp(v[0]);
  *clumping = v[1];
  *residual = exp(v[2]);
}

}  // namespace mxr
