// mxr_recomb.hip -- inhomogeneous recombinations for find_HII_bubbles: the MHR00 density-PDF
// tables of recombination rate, clumping factor and residual neutral fraction over (z, T,
// ln Gamma_12) (src/core/recombinations.c:208-374, integrands :410-560, MHR parameters :561-740,
// neutral fraction :742-759), built on the GPU with one thread per table entry, and their
// natural cubic splines in ln Gamma_12, evaluated per cell by splined_recombination
// (recombinations.c:35-100).
//
// The reference reads the tables from an HDF5 cache or rebuilds them at start-up with GSL; here
// they are always rebuilt (once per context and cosmology), with the same grid of nodes, the same
// integrands in double and the same adaptive GK61 quadrature (mxr_qag.hpp).  Grid constants and
// the C / beta parameter nodes are the reference's (recombinations.h:18-35, recombinations.c).
#include <cmath>
#include <cstring>
#include <vector>

#include "mxr_internal.hpp"
#include "mxr_qag.hpp"
#include "mxr_recomb.hpp"

// every integrand and spline here rounds each operation as the reference's x86-64 build does
#pragma clang fp contract(off)

namespace mxr {

namespace {

constexpr int A_NPTS = 60, C_NPTS = 12, BETA_NPTS = 5;
constexpr double ALPHA_B_10K = 2.59e-13;  // case B at 1e4 K (Osterbrock)
constexpr double TINY_RR = 1e-30;
constexpr double HUBBLE_S = 3.2407789e-18;
constexpr double GRAVITY_CGS = 6.672e-8;
constexpr double PROTONMASS_G = 1.6726e-24;

// MHR00 parameters C(z) and beta(z) at z = 2, 3, ... (recombinations.c: init_C_MHR, init_beta_MHR)
constexpr double C_NODES[C_NPTS] = {0.558, 0.599, 0.611, 0.769, 0.868, 0.930, 0.964, 0.983, 0.993, 0.998, 0.999, 1.00};
constexpr double BETA_NODES[BETA_NPTS] = {-2.23, -2.35, -2.48, -2.49, -2.50};

// natural cubic spline through (x_i, y_i): the second-derivative halves c_i (c_0 = c_{n-1} = 0),
// as GSL 2.x's cspline_init (natural) computes them: right-hand sides 3 (dy_{i+1} / h_{i+1} -
// dy_i / h_i) with each 1 / h formed first, and the symmetric tridiagonal system (diagonal
// 2 (h_i + h_{i+1}), off-diagonal h_{i+1}) solved by its L D L^T factorisation
// (gsl_linalg_solve_symm_tridiag).  work holds the factor's lower diagonal (n - 3 entries).
template <class V>
__host__ __device__ inline void spline_c(const double* x, const double* y, int n, V& c, V& work)
{
  c[0] = 0.0;
  c[n - 1] = 0.0;
  const int m = n - 2;
  if (m <= 0)
    return;
  // forward: alpha_i = d_i - e_{i-1} gamma_{i-1}, gamma_i = e_i / alpha_i, z_i = g_i -
  // gamma_{i-1} z_{i-1}, and c_{i+1} = z_i / alpha_i for the back substitution
  double z_prev = 0.0;
  for (int i = 0; i < m; i++) {
    const double h_i = x[i + 1] - x[i], h_ip1 = x[i + 2] - x[i + 1];
    const double rh_i = h_i != 0.0 ? 1.0 / h_i : 0.0, rh_ip1 = h_ip1 != 0.0 ? 1.0 / h_ip1 : 0.0;
    const double g = 3.0 * ((y[i + 2] - y[i + 1]) * rh_ip1 - (y[i + 1] - y[i]) * rh_i);
    const double diag = 2.0 * (h_ip1 + h_i);
    const double alpha = i == 0 ? diag : diag - h_i * work[i - 1];  // e_{i-1} = h_i
    const double z = i == 0 ? g : g - work[i - 1] * z_prev;
    if (i < m - 1)
      work[i] = h_ip1 / alpha;
    c[i + 1] = z / alpha;
    z_prev = z;
  }
  for (int i = m - 2; i >= 0; i--)
    c[i + 1] = c[i + 1] - work[i] * c[i + 2];
}

// GSL's cspline evaluation on interval i (x_i <= xq < x_{i+1})
__host__ __device__ inline double spline_at(const double* x, const double* y, const double* c, int i, double xq)
{
  const double h = x[i + 1] - x[i], dx = xq - x[i];
  const double b = (y[i + 1] - y[i]) / h - h * (c[i + 1] + 2.0 * c[i]) / 3.0;
  const double d = (c[i + 1] - c[i]) / (3.0 * h);
  return y[i] + dx * (b + dx * (c[i] + dx * d));
}

// a spline over unit-spaced nodes x_i = 2 + i (the MHR parameter tables)
struct UnitSpline {
  int n;
  double x[A_NPTS], y[A_NPTS], c[A_NPTS];
  void build(const double* ys, int count)
  {
    n = count;
    double work[A_NPTS];
    for (int i = 0; i < n; i++) {
      x[i] = 2.0 + (double)(float)i;
      y[i] = ys[i];
    }
    spline_c(x, y, n, c, work);
  }
  double operator()(double z) const
  {
    int i = (int)std::floor(z - 2.0);
    if (i < 0)
      i = 0;
    if (i > n - 2)
      i = n - 2;
    return spline_at(x, y, c, i, z);
  }
};

struct MhrParams {
  UnitSpline A, C, beta;
  double C_z(double z) const { return z >= 13.0 ? 1.0 : (z <= 2.0 ? 0.558 : C(z)); }
  double beta_z(double z) const { return z >= 6.0 ? -2.50 : (z <= 2.0 ? -2.23 : beta(z)); }
  double A_z(double z) const
  {
    if (z >= 2.0 + (float)A_NPTS)
      return A(2.0 + (float)A_NPTS);
    if (z <= 2.0)
      return A(2.0);
    return A(z);
  }
};

// the PDF normalisation 1 / int P(Delta) dDelta at z = 2 .. 61 (A_aux_integral, QAG on
// [1e-25, 1e25] with relative tolerance 1e-3)
int build_mhr(MhrParams& m)
{
  m.C.build(C_NODES, C_NPTS);
  m.beta.build(BETA_NODES, BETA_NPTS);
  double A_table[A_NPTS];
  static thread_local QagWork<1000> w;
  for (int i = 0; i < A_NPTS; i++) {
    const double z = 2.0 + (float)i;
    const double C0 = m.C_z(z), be = m.beta_z(z);
    const double s = 2.0 * 7.61 / (3.0 * (1.0 + z));
    auto aux = [&](double D) {
      const double u = std::pow(D, -2.0 / 3.0) - C0;
      return std::exp(-u * u / (2.0 * s * s)) * std::pow(D, be);
    };
    // the reference turns GSL's error handler off (init.c:257) and uses the result whatever the
    // status; a workspace of `limit` intervals never reports status 4
    double result = 0, err = 0;
    (void)qag<1000>(w, aux, 1e-25, 1e25, 0, 0.001, 1000, &result, &err);
    if (!(result > 0)) {
      set_error("MHR normalisation integral at z = %g is not positive", z);
      return MXR_ERR_TABLES;
    }
    A_table[i] = 1.0 / result;
  }
  m.A.build(A_table, A_NPTS);
  return MXR_OK;
}

// per table row z: the MHR parameters and the mean hydrogen density at z
struct RowPar {
  double z, A, C0, beta, avenH;
};

struct RecArgs {
  const RowPar* rows;   // RR_Z_NPTS rows
  double corr_He;       // 1 / (4 / Y_He - 3)
  double* tab;          // [3][RR_Z_NPTS][RR_T_NPTS][RR_G_NPTS]: ln rate, clumping, ln residual
  int* failures;
};

// neutral_fraction (recombinations.c:742-759), case B
__device__ inline double neutral_fraction(double density, double T4, double gamma12, double corr_He)
{
  const double alpha = ALPHA_B_10K * pow((T4 * 1e4) / 1.0e4, -0.75);
  gamma12 *= 1e-12;
  double chi = (1 + corr_He) * density * alpha / gamma12;
  if (chi < TINY_RR)
    return 0;
  if (chi < 1e-5)
    return chi;
  const double b = -2 - gamma12 / (density * (1 + corr_He) * alpha);
  chi = (-b - sqrt(b * b - 4)) / 2.0;
  return chi;
}

// Gamma_SS (recombinations.c:397-403): self-shielded rate (Rahmati et al. 2013 fit)
__device__ inline double gamma_ss(double Gamma_bg, double Delta, double T_4, double z)
{
  const double D_ss = 26.7 * pow(T_4, 0.17) * pow((1 + z) / 10.0, -3) * pow(Gamma_bg, 2.0 / 3.0);
  return Gamma_bg * (0.98 * pow((1.0 + pow(Delta / D_ss, 1.64)), -2.28) + 0.02 * pow(1.0 + Delta / D_ss, -0.84));
}

// the MHR00 density PDF (recombinations.c:431, 447, 463, 479)
__device__ inline double pdelta(const RowPar& p, double D)
{
  const double u = (pow(D, -2.0 / 3.0) - p.C0) / ((2.0 * 7.61 / (3.0 * (1.0 + p.z))));
  return p.A * exp(-0.5 * pow(u, 2)) * pow(D, p.beta);
}

// one table entry: the four integrals over ln Delta in [ln 0.01, ln 200] (recombination_rate,
// clumping_factor's numerator and denominator, residual_neutral_hydrogen), each GSL QAG GK61
// with relative tolerance 0.01 and limit 1000.  Over this grid the adaptive scheme stops after one
// or two intervals; the per-thread workspace holds 16 and an entry that would need more is
// reported as an error rather than truncated.
constexpr int REC_QAG_CAP = 16;

__global__ void __launch_bounds__(64) k_rec_tables(RecArgs a)
{
  const long n = (long)RR_Z_NPTS * RR_T_NPTS * RR_G_NPTS;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n)
    return;
  const int g = (int)(e % RR_G_NPTS);
  const int t_ct = (int)((e / RR_G_NPTS) % RR_T_NPTS);
  const int z_ct = (int)(e / ((long)RR_G_NPTS * RR_T_NPTS));
  const RowPar p = a.rows[z_ct];
  const double T4 = pow(10.0, (double)((float)t_ct * RR_DEL_T + RR_T_STA) - 4.0);
  const double gamma = exp((double)(RR_LNG_MIN + (float)g * RR_DEL_LNG));
  const double alpha = ALPHA_B_10K * pow((T4 * 1e4) / 1.0e4, -0.75);
  const double corr_He = a.corr_He;
  // x_HI of one density, shared by the four integrands
  auto xHI = [&](double D) {
    return neutral_fraction(p.avenH * D, T4, gamma_ss(gamma, D, T4, p.z), corr_He);
  };
  auto rr = [&](double lnD) {
    const double D = exp(lnD);
    const double x_e = 1.0 - xHI(D);
    return p.avenH * D * pdelta(p, D) * alpha * x_e * x_e * D * D;
  };
  auto cf_num = [&](double lnD) {
    const double D = exp(lnD);
    const double x_e = 1.0 - xHI(D);
    return pdelta(p, D) * x_e * x_e * D * D * D;
  };
  auto cf_den = [&](double lnD) {
    const double D = exp(lnD);
    const double x_e = 1.0 - xHI(D);
    return pdelta(p, D) * x_e * D * D;
  };
  auto rnh = [&](double lnD) {
    const double D = exp(lnD);
    return 1e4 * pdelta(p, D) * xHI(D) * D;
  };
  QagWork<REC_QAG_CAP> w;
  const double lo = log(0.01), hi = log(200.0);
  double r_rate, r_num, r_den, r_res, err;
  int st = 0;
  st |= qag<REC_QAG_CAP>(w, rr, lo, hi, 0, 0.01, 1000, &r_rate, &err) == 4;
  st |= qag<REC_QAG_CAP>(w, cf_num, lo, hi, 0, 0.01, 1000, &r_num, &err) == 4;
  st |= qag<REC_QAG_CAP>(w, cf_den, lo, hi, 0, 0.01, 1000, &r_den, &err) == 4;
  st |= qag<REC_QAG_CAP>(w, rnh, lo, hi, 0, 0.01, 1000, &r_res, &err) == 4;
  if (st)
    atomicAdd(a.failures, 1);
  a.tab[e] = log(r_rate);
  a.tab[n + e] = r_num / r_den / r_den;
  a.tab[2 * n + e] = log(r_res);
}

// per (z, T) row: the natural-spline coefficients of the three tables in ln Gamma_12
__global__ void k_rec_splines(const double* __restrict__ x, const double* __restrict__ tab, double* __restrict__ coef,
                              double* __restrict__ work)
{
  const long nrow = (long)RR_Z_NPTS * RR_T_NPTS;
  const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= 3 * nrow)
    return;
  const double* y = tab + r * RR_G_NPTS;
  double* c = coef + r * RR_G_NPTS;
  double* wk = work + r * RR_G_NPTS;
  spline_c(x, y, RR_G_NPTS, c, wk);
}

}  // namespace

// The tables of this context for the params' cosmology, built on first use.
int rec_tables(Ctx* c, const mxr_hii_params* p, RecTab* out)
{
  const double key[4] = {p->Hubble_h, p->OmegaM, p->BaryonFrac, p->Y_He};
  if (!(p->Y_He > 0 && p->Y_He < 1)) {
    set_error("recombinations need 0 < Y_He < 1 (got %g)", p->Y_He);
    return MXR_ERR_INVALID_ARG;
  }
  const long n = (long)RR_Z_NPTS * RR_T_NPTS * RR_G_NPTS;
  if (!(c->rec_tab && std::memcmp(key, c->rec_key, sizeof key) == 0)) {
    MhrParams m;
    RC(build_mhr(m));
    // reionization.h:15-32: the present-day hydrogen number density
    const double rho_cgs = 3.0 * HUBBLE_S * HUBBLE_S * p->Hubble_h * p->Hubble_h / (8.0 * M_PI * GRAVITY_CGS);
    const double No = rho_cgs * (p->BaryonFrac * p->OmegaM) * (1 - p->Y_He) / PROTONMASS_G;
    std::vector<RowPar> rows(RR_Z_NPTS);
    for (int z_ct = 0; z_ct < RR_Z_NPTS; z_ct++) {
      const double z = (double)((float)z_ct * RR_DEL_Z + RR_Z_END);
      rows[z_ct] = RowPar{z, m.A_z(z), m.C_z(z), m.beta_z(z), No * std::pow(1 + z, 3)};
    }
    std::vector<double> x(RR_G_NPTS);
    for (int g = 0; g < RR_G_NPTS; g++)
      x[g] = (double)(RR_LNG_MIN + (float)g * RR_DEL_LNG);
    if (!c->rec_tab) {
      MXR_HIP(hipMalloc(&c->rec_tab, sizeof(double) * 3 * n));
      MXR_HIP(hipMalloc(&c->rec_coef, sizeof(double) * 3 * n));
      MXR_HIP(hipMalloc(&c->rec_x, sizeof(double) * RR_G_NPTS));
    }
    RowPar* drows = nullptr;
    int* dfail = nullptr;
    double* dwork = nullptr;
    MXR_HIP(hipMalloc(&drows, sizeof(RowPar) * RR_Z_NPTS));
    MXR_HIP(hipMalloc(&dfail, sizeof(int)));
    MXR_HIP(hipMalloc(&dwork, sizeof(double) * 3 * n));
    int rc = MXR_OK;
    auto run = [&]() -> int {
      MXR_HIP(hipMemcpyAsync(drows, rows.data(), sizeof(RowPar) * RR_Z_NPTS, hipMemcpyHostToDevice, c->stream));
      MXR_HIP(hipMemcpyAsync(c->rec_x, x.data(), sizeof(double) * RR_G_NPTS, hipMemcpyHostToDevice, c->stream));
      MXR_HIP(hipMemsetAsync(dfail, 0, sizeof(int), c->stream));
      RecArgs a{drows, 1.0 / (4.0 / p->Y_He - 3), c->rec_tab, dfail};
      hipLaunchKernelGGL(k_rec_tables, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, c->stream, a);
      MXR_HIP(hipGetLastError());
      const long nr = 3L * RR_Z_NPTS * RR_T_NPTS;
      hipLaunchKernelGGL(k_rec_splines, dim3((unsigned)((nr + 127) / 128)), dim3(128), 0, c->stream, c->rec_x,
                         c->rec_tab, c->rec_coef, dwork);
      MXR_HIP(hipGetLastError());
      int fails = 0;
      MXR_HIP(hipMemcpyAsync(&fails, dfail, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      MXR_HIP(hipStreamSynchronize(c->stream));
      if (fails) {
        set_error("%d recombination-table entries needed more than %d quadrature intervals", fails, REC_QAG_CAP);
        return MXR_ERR_TABLES;
      }
      return MXR_OK;
    };
    rc = run();
    (void)hipFree(drows);
    (void)hipFree(dfail);
    (void)hipFree(dwork);
    if (rc)
      return rc;
    std::memcpy(c->rec_key, key, sizeof key);
  }
  out->tab = c->rec_tab;
  out->coef = c->rec_coef;
  out->x = c->rec_x;
  out->n = n;
  return MXR_OK;
}

int rec_tables_copy(Ctx* c, const mxr_hii_params* p, double* lnGamma, double* ln_rate, double* clumping,
                    double* ln_residual)
{
  RecTab t;
  RC(rec_tables(c, p, &t));
  const long n = t.n;
  double* dst[4] = {lnGamma, ln_rate, clumping, ln_residual};
  const double* src[4] = {t.x, t.tab, t.tab + n, t.tab + 2 * n};
  const long cnt[4] = {RR_G_NPTS, n, n, n};
  for (int i = 0; i < 4; i++)
    if (dst[i])
      MXR_HIP(hipMemcpyAsync(dst[i], src[i], sizeof(double) * cnt[i], hipMemcpyDeviceToHost, c->stream));
  MXR_HIP(hipStreamSynchronize(c->stream));
  return MXR_OK;
}

}  // namespace mxr
