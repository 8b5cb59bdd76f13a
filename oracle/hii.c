/*
 * hii.c -- oracle restatement of find_HII_bubbles, filter, RtoM, T_RECFAST and the 21-cm
 * brightness box (TEST INFRASTRUCTURE ONLY; see oracle.h).
 *
 * Reference files followed:
 *   src/core/find_HII_bubbles.c:27-710      RtoM, ionised-gas temperatures, _find_HII_bubbles
 *   src/core/reionization.c:3373-3433       filter
 *   src/core/XRayHeatingFunctions.c:134-192 T_RECFAST (GSL natural cubic spline of
 *                                           recfast_LCDM.dat), :566-569 cT_approx
 *   src/core/init.c:203-225                 set_units (G, Hubble, RhoCrit)
 *   src/core/virial_properties.c:6-96       E_z, hubble_at_z, hubble_time
 *   src/core/BrightnessTemperature.c:24-549 ComputeBrightnessTemperatureBox
 *   src/meraxes.h:21-67, src/core/reionization.h:10-47, src/core/misc_tools.h:14-20 constants
 */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "oracle.h"

/* constants, src/meraxes.h:21-67 and src/core/reionization.h:10-47 */
#define GRAVITY 6.672e-8
#define SOLAR_MASS 1.989e33
#define PLANCK 6.6262e-27
#define PROTONMASS 1.6726e-24
#define HUBBLE 3.2407789e-18
#define SEC_PER_YEAR 3.155e7
#define SEC_PER_MEGAYEAR 3.155e13
#define SIGMA_HI (double)(6.3e-18)
#define TCMB 2.728
#define MPC 3.086e24
#define T_RE (double)(1e4)
#define REL_TOL (float)1e-5
#define ABS_TOL (float)1e-8
#define L_FACTOR 0.620350491
#define RECFAST_NPTS 501

/* misc_tools.h:14-20 */
#define CLAMP_0_1(x)                                                                                                   \
  do {                                                                                                                 \
    if ((x) < 0.0)                                                                                                     \
      (x) = 0.0;                                                                                                       \
    else if ((x) > 1.0)                                                                                                \
      (x) = 1.0;                                                                                                       \
  } while (0)

/* FFT kernels (shared with fft.c through the template) */
#define REAL float
#define SUFFIX f32
#include "fft_tmpl.h"
#undef REAL
#undef SUFFIX
#define REAL double
#define SUFFIX f64
#include "fft_tmpl.h"
#undef REAL
#undef SUFFIX

/* ---------------------------------------------------------------- units and cosmology */

static double unit_time_s(const mxr_hii_params* p) { return p->UnitLength_in_cm / p->UnitVelocity_in_cm_per_s; }

/* init.c:203-225 */
double oracle_rho_crit(const mxr_hii_params* p)
{
  const double UT = unit_time_s(p);
  const double G = GRAVITY / pow(p->UnitLength_in_cm, 3) * p->UnitMass_in_g * pow(UT, 2);
  const double Hubble = HUBBLE * UT;
  return 3 * Hubble * Hubble / (8 * M_PI * G);
}

/* virial_properties.c:6-20, 79-96 */
double oracle_hubble_time(const mxr_hii_params* p)
{
  const double z = p->redshift;
  const double opz = 1. + z;
  const double Ez = sqrt(p->OmegaM * opz * opz * opz + p->OmegaK * opz * opz + p->OmegaLambda);
  const double Hubble = HUBBLE * unit_time_s(p);
  return 1.0 / (Hubble * Ez);
}

/* find_HII_bubbles.c:27-46 */
double oracle_RtoM(const mxr_hii_params* p, double R)
{
  const double OmegaM = p->OmegaM;
  const double RhoCrit = oracle_rho_crit(p);
  switch (p->ReionRtoMFilterType) {
    case 0:
      return (4.0 / 3.0) * M_PI * pow(R, 3) * (OmegaM * RhoCrit);
    case 1:
      return pow(2 * M_PI, 1.5) * OmegaM * RhoCrit * pow(R, 3);
    default:
      return -1;
  }
}

/* find_HII_bubbles.c:145-148, 256-285, 546 */
int oracle_hii_radii(const mxr_hii_params* p, double* radii, int capacity)
{
  const int N = p->ReionGridDim;
  const double box_size = p->BoxSize;
  double cell_length_factor = L_FACTOR;
  if ((box_size / (double)N) < 1.0)
    cell_length_factor = 1.0;
  double Rmax;
  if (p->Flag_EvolvingReionRBubbleMax)
    Rmax = (p->redshift > 6.) ? 25.483241248322766 : 112. * pow((1. + p->redshift) / 5., -4.4);
  else
    Rmax = p->Flag_IncludeRecombinations ? p->ReionRBubbleMaxRecomb : p->ReionRBubbleMax;
  double R = fmin(Rmax, L_FACTOR * box_size);
  int n = 0, last = 0;
  while (!last) {
    if (((R / p->ReionDeltaRFactor) <= (cell_length_factor * box_size / (double)N)) ||
        ((R / p->ReionDeltaRFactor) <= p->ReionRBubbleMin)) {
      last = 1;
      R = cell_length_factor * box_size / (double)N;
    }
    if (radii && n < capacity)
      radii[n] = R;
    n++;
    R /= p->ReionDeltaRFactor;
  }
  return n;
}

/* XRayHeatingFunctions.c:566-569 */
static float cT_approx(float z) { return 0.58 - 0.006 * (z - 10.0); }

/* find_HII_bubbles.c:48-74 */
static float fully_ionized_T(float z_re, float z, float delta)
{
  float result, delta_re;
  if (fabs(z - z_re) < 1e-4)
    result = 1;
  else {
    delta_re = delta * (1. + z) / (1. + z_re);
    if (delta_re <= -1)
      delta_re = -1. + ABS_TOL;
    if (delta <= -1)
      delta = -1. + ABS_TOL;
    result = pow((1. + delta) / (1. + delta_re), 1.1333);
    result *= pow((1. + z) / (1. + z_re), 3.4);
    result *= expf(pow((1. + z) / 7.1, 2.5) - pow((1. + z_re) / 7.1, 2.5));
  }
  result *= pow(T_RE, 1.7);
  result += pow(1e4 * ((1. + z) / 4.), 1.7) * (1 + delta);
  result = pow(result, 0.5882);
  return result;
}

/* find_HII_bubbles.c:76-84 */
static float partially_ionized_T(float T_HI, float res_xH)
{
  if (res_xH <= 0.)
    return T_RE;
  if (res_xH >= 1)
    return T_HI;
  return T_HI * res_xH + T_RE * (1. - res_xH);
}

/* ------------------------------------------------------------ natural cubic spline */
/* The interpolant GSL's gsl_interp_cspline builds (natural end conditions, S'' = 0 at both
 * ends): S_i(x) = y_i + b_i t + c_i t^2 + d_i t^3, t = x - x_i, with c_i = M_i / 2 from the
 * tridiagonal system h_{i-1} M_{i-1} + 2 (h_{i-1} + h_i) M_i + h_i M_{i+1} = 6 (s_i - s_{i-1}).
 * Interval lookup: x_i <= x < x_{i+1}, the last interval for x = x_{n-1}. */
double oracle_spline_natural(const double* x, const double* y, int n, double xq)
{
  double* M = (double*)calloc((size_t)n, sizeof(double));
  double* cp = (double*)calloc((size_t)n, sizeof(double));
  double* dp = (double*)calloc((size_t)n, sizeof(double));
  /* Thomas algorithm on the n-2 interior unknowns */
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
  const double v = y[lo] + t * (b + t * (c + t * d));
  free(M);
  free(cp);
  free(dp);
  return v;
}

/* XRayHeatingFunctions.c:134-192: recfast_LCDM.dat rows "z x_e ? T_K", z descending from
 * 500; read as floats, stored reversed so z ascends. */
double oracle_T_RECFAST(const char* tables_dir, float z)
{
  if (!tables_dir)
    return -1;
  char fname[1024];
  snprintf(fname, sizeof(fname), "%s/recfast_LCDM.dat", tables_dir);
  FILE* F = fopen(fname, "r");
  if (!F)
    return -1;
  double zt[RECFAST_NPTS], TKt[RECFAST_NPTS];
  for (int i = RECFAST_NPTS - 1; i >= 0; i--) {
    float currz, trash, currTK;
    if (fscanf(F, "%f %E %E %E", &currz, &trash, &trash, &currTK) != 4) {
      fclose(F);
      return -1;
    }
    zt[i] = currz;
    TKt[i] = currTK;
  }
  fclose(F);
  if (z > zt[RECFAST_NPTS - 1])
    return -1;
  return oracle_spline_natural(zt, TKt, RECFAST_NPTS, z);
}

/* ---------------------------------------------------------------------------- filter */
/* reionization.c:3373-3433, float arithmetic as written (build with -ffp-contract=off). */
static void filter_f32(float* boxf, int local_ix_start, int slab_nx, int grid_dim, double BoxSize, float R,
                       int filter_type)
{
  float complex* box = (float complex*)boxf;
  int middle = grid_dim / 2;
  float box_size = (float)BoxSize;
  float delta_k = (float)(2.0 * M_PI / box_size);
#pragma omp parallel for schedule(static)
  for (int n_x = 0; n_x < slab_nx; n_x++) {
    float k_x;
    int n_x_global = n_x + local_ix_start;
    if (n_x_global > middle)
      k_x = (n_x_global - grid_dim) * delta_k;
    else
      k_x = n_x_global * delta_k;
    for (int n_y = 0; n_y < grid_dim; n_y++) {
      float k_y;
      if (n_y > middle)
        k_y = (n_y - grid_dim) * delta_k;
      else
        k_y = n_y * delta_k;
      for (int n_z = 0; n_z <= middle; n_z++) {
        float k_z = n_z * delta_k;
        float k_mag = sqrtf(k_x * k_x + k_y * k_y + k_z * k_z);
        float kR = k_mag * R;
        const long idx = n_z + (long)(middle + 1) * (n_y + (long)grid_dim * n_x);
        switch (filter_type) {
          case 0:
            if (kR > 1e-4)
              box[idx] *= (float complex)(3.0 * (sinf(kR) / powf(kR, 3) - cosf(kR) / powf(kR, 2)));
            break;
          case 1:
            kR *= 0.413566994;
            if (kR > 1)
              box[idx] = (float complex)0.0;
            break;
          case 2:
            kR *= 0.643;
            box[idx] *= (float complex)(powf((float)M_E, (float)(-kR * kR / 2.0)));
            break;
          default:
            break;
        }
      }
    }
  }
}

/* The same window evaluated entirely in double (FP64 oracle). */
static void filter_f64(double* boxd, int local_ix_start, int slab_nx, int grid_dim, double BoxSize, double R,
                       int filter_type)
{
  double complex* box = (double complex*)boxd;
  int middle = grid_dim / 2;
  double delta_k = 2.0 * M_PI / BoxSize;
#pragma omp parallel for schedule(static)
  for (int n_x = 0; n_x < slab_nx; n_x++) {
    int n_x_global = n_x + local_ix_start;
    double k_x = (n_x_global > middle) ? (n_x_global - grid_dim) * delta_k : n_x_global * delta_k;
    for (int n_y = 0; n_y < grid_dim; n_y++) {
      double k_y = (n_y > middle) ? (n_y - grid_dim) * delta_k : n_y * delta_k;
      for (int n_z = 0; n_z <= middle; n_z++) {
        double k_z = n_z * delta_k;
        double kR = sqrt(k_x * k_x + k_y * k_y + k_z * k_z) * R;
        const long idx = n_z + (long)(middle + 1) * (n_y + (long)grid_dim * n_x);
        switch (filter_type) {
          case 0:
            if (kR > 1e-4)
              box[idx] *= 3.0 * (sin(kR) / pow(kR, 3) - cos(kR) / pow(kR, 2));
            break;
          case 1:
            if (kR * 0.413566994 > 1)
              box[idx] = 0.0;
            break;
          case 2:
            box[idx] *= exp(-(0.643 * kR) * (0.643 * kR) / 2.0);
            break;
          default:
            break;
        }
      }
    }
  }
}

int oracle_filter_f32(float* box, int local_ix_start, int slab_nx, int grid_dim, double BoxSize, float R,
                      int filter_type)
{
  if (filter_type < 0 || filter_type > 2)
    return MXR_ERR_FILTER;
  filter_f32(box, local_ix_start, slab_nx, grid_dim, BoxSize, R, filter_type);
  return 0;
}

int oracle_filter_f64(double* box, int local_ix_start, int slab_nx, int grid_dim, double BoxSize, double R,
                      int filter_type)
{
  if (filter_type < 0 || filter_type > 2)
    return MXR_ERR_FILTER;
  filter_f64(box, local_ix_start, slab_nx, grid_dim, BoxSize, R, filter_type);
  return 0;
}

/* ------------------------------------------------------------------- the pipeline */
#define REAL float
#define SUFFIX f32
#define FILTER_R_T float
#define FMAXR fmaxf
#define FMINR fminf
#include "hii_tmpl.h"
#undef REAL
#undef SUFFIX
#undef FILTER_R_T
#undef FMAXR
#undef FMINR

#define REAL double
#define SUFFIX f64
#define FILTER_R_T double
#define FMAXR fmax
#define FMINR fmin
#include "hii_tmpl.h"
#undef REAL
#undef SUFFIX
#undef FILTER_R_T
#undef FMAXR
#undef FMINR

/* ------------------------------------------- the same pipeline on FFTW3's float API
 * The reference executes fftwf plans (src/core/reionization.c:907-997; find_HII_bubbles.c:
 * 166-229, 351-370).  No FFTW library is installed here, but MKL's libmkl_rt exports FFTW3's
 * float interface, so the CPU baseline runs the FP32 restatement above on that backend: the
 * forward transform out of place from the padded real grid (advanced interface, inembed
 * {N, N, 2(N/2+1)}), the inverse in place, as the reference plans them.  Symbols are resolved
 * at run time; without the library the entry point returns MXR_ERR_UNSUPPORTED. */
#include <dlfcn.h>
#include <omp.h>

#define FFTWF_ESTIMATE (1U << 6)

static struct {
  int tried, ok, N;
  void* (*plan_many_dft_r2c)(int, const int*, int, float*, const int*, int, int, void*, const int*, int, int,
                             unsigned);
  void* (*plan_dft_c2r_3d)(int, int, int, void*, float*, unsigned);
  void (*execute_dft_r2c)(void*, float*, void*);
  void (*execute_dft_c2r)(void*, void*, float*);
  void (*destroy_plan)(void*);
  int (*init_threads)(void);
  void (*plan_with_nthreads)(int);
  void* r2c;
  void* c2r;
} fw;

static const char* fftw_lib_path(void)
{
  const char* e = getenv("ORACLE_FFTW_LIB");
  return e ? e : "/opt/conda/lib/libmkl_rt.so.1";
}

static int fftw_load(void)
{
  if (fw.tried)
    return fw.ok;
  fw.tried = 1;
  /* MKL's GNU threading layer shares libgomp with the rest of the oracle */
  setenv("MKL_THREADING_LAYER", "GNU", 0);
  void* h = dlopen(fftw_lib_path(), RTLD_NOW | RTLD_LOCAL);
  if (!h)
    return 0;
  *(void**)&fw.plan_many_dft_r2c = dlsym(h, "fftwf_plan_many_dft_r2c");
  *(void**)&fw.plan_dft_c2r_3d = dlsym(h, "fftwf_plan_dft_c2r_3d");
  *(void**)&fw.execute_dft_r2c = dlsym(h, "fftwf_execute_dft_r2c");
  *(void**)&fw.execute_dft_c2r = dlsym(h, "fftwf_execute_dft_c2r");
  *(void**)&fw.destroy_plan = dlsym(h, "fftwf_destroy_plan");
  *(void**)&fw.init_threads = dlsym(h, "fftwf_init_threads");
  *(void**)&fw.plan_with_nthreads = dlsym(h, "fftwf_plan_with_nthreads");
  fw.ok = fw.plan_many_dft_r2c && fw.plan_dft_c2r_3d && fw.execute_dft_r2c && fw.execute_dft_c2r &&
          fw.destroy_plan && fw.init_threads && fw.plan_with_nthreads;
  if (fw.ok) {
    fw.init_threads();
    fw.plan_with_nthreads(omp_get_max_threads());
  }
  return fw.ok;
}

/* plans for dimension N, made on the first transform with these arrays' alignment (malloc'd
 * grids all share it) and executed on new arrays afterwards */
static int fftw_plans(int N, float* rin, void* cout, void* inplace)
{
  if (fw.N == N && fw.r2c && fw.c2r)
    return 0;
  if (fw.r2c)
    fw.destroy_plan(fw.r2c);
  if (fw.c2r)
    fw.destroy_plan(fw.c2r);
  fw.r2c = fw.c2r = NULL;
  const int H = N / 2 + 1;
  const int n[3] = {N, N, N}, inembed[3] = {N, N, 2 * H}, onembed[3] = {N, N, H};
  if (rin && cout)
    fw.r2c = fw.plan_many_dft_r2c(3, n, 1, rin, inembed, 1, 0, cout, onembed, 1, 0, FFTWF_ESTIMATE);
  if (inplace)
    fw.c2r = fw.plan_dft_c2r_3d(N, N, N, inplace, (float*)inplace, FFTWF_ESTIMATE);
  fw.N = N;
  return 0;
}

typedef cplx_f32 cplx_fftw;
#define filter_fftw filter_f32

static int r2c3d_fftw(int N, const float* in, cplx_fftw* out)
{
  if (fw.N != N || !fw.r2c) {
    fw.N = 0;
    fftw_plans(N, (float*)in, out, out);
  }
  fw.execute_dft_r2c(fw.r2c, (float*)in, out);
  return 0;
}

static int c2r3d_fftw(int N, cplx_fftw* data)
{
  if (fw.N != N || !fw.c2r) {
    fw.N = 0;
    fftw_plans(N, NULL, NULL, data);
  }
  fw.execute_dft_c2r(fw.c2r, data, (float*)data);
  return 0;
}

#define REAL float
#define SUFFIX fftw
#define FILTER_R_T float
#define FMAXR fmaxf
#define FMINR fminf
#include "hii_tmpl.h"
#undef REAL
#undef SUFFIX
#undef FILTER_R_T
#undef FMAXR
#undef FMINR

static int check_args(const mxr_hii_params* p, const mxr_hii_grids* g)
{
  if (!p || !g || !g->deltax || !g->stars || !g->weighted_sfr || !g->xH || !g->z_at_ionization ||
      !g->J_21_at_ionization || !g->r_bubble || !g->temp_kinetic_all_gas)
    return MXR_ERR_INVALID_ARG;
  if (p->ReionGridDim < 2 || (p->ReionGridDim % 2))
    return MXR_ERR_INVALID_ARG;
  if (p->Flag_IncludeSpinTemp && (!g->x_e_box || !g->Tk_box))
    return MXR_ERR_INVALID_ARG;
  if (p->Flag_BHFeedback && (!g->effective_bhm || !g->effective_bhar))
    return MXR_ERR_INVALID_ARG;
  return 0;
}

int oracle_find_hii_bubbles_f32(const mxr_hii_params* p, const mxr_hii_grids* g, mxr_hii_globals* out,
                                double* margin)
{
  int rc = check_args(p, g);
  if (rc)
    return rc;
  rc = hii_run_f32(p, g, out, margin, 0, NULL, NULL);
  return rc < 0 ? rc : 0;
}

int oracle_find_hii_bubbles_f64(const mxr_hii_params* p, const mxr_hii_grids* g, mxr_hii_globals* out,
                                double* margin)
{
  int rc = check_args(p, g);
  if (rc)
    return rc;
  rc = hii_run_f64(p, g, out, margin, 0, NULL, NULL);
  return rc < 0 ? rc : 0;
}

int oracle_fftw_available(void) { return fftw_load(); }
const char* oracle_fftw_library(void) { return fftw_lib_path(); }

int oracle_find_hii_bubbles_fftw(const mxr_hii_params* p, const mxr_hii_grids* g, mxr_hii_globals* out,
                                 double* seconds)
{
  int rc = check_args(p, g);
  if (rc)
    return rc;
  if (!fftw_load())
    return MXR_ERR_UNSUPPORTED;
  rc = hii_run_fftw(p, g, out, NULL, 0, seconds, NULL);
  return rc < 0 ? rc : 0;
}

int oracle_hii_sample_f32(const mxr_hii_params* p, const mxr_hii_grids* g, int n_radii, double* seconds,
                          double* seconds_fwd)
{
  int rc = check_args(p, g);
  if (rc)
    return rc;
  return hii_run_f32(p, g, NULL, NULL, n_radii > 0 ? n_radii : 0, seconds, seconds_fwd);
}

/* the temperature helpers on their own, for the tests that pin them against closed forms */
float oracle_fully_ionized_T(float z_re, float z, float delta) { return fully_ionized_T(z_re, z, delta); }
float oracle_partially_ionized_T(float T_HI, float res_xH) { return partially_ionized_T(T_HI, res_xH); }

/* ---------------------------------------------------------------- brightness box */
#define N_RSD_STEPS 50            /* reionization.h:34 */
#define MAX_DVDR (float)(0.2)     /* reionization.h:45 */
#define TB_FRACT_FLOAT_ERR (double)(1e-7)

/* velocity_gradient (reionization.c:3435-3450): each half-spectrum value times i k_z,
 * k_z = n_z (float)(2 pi / (float)BoxSize), as the complex product (a + ib)(0 + i k_z). */
int oracle_velocity_gradient(float* box, int slab_nx, int grid_dim, double BoxSize)
{
  if (!box || slab_nx < 0 || grid_dim < 2)
    return MXR_ERR_INVALID_ARG;
  const int middle = grid_dim / 2;
  const float box_size = (float)BoxSize;
  const float delta_k = (float)(2.0 * M_PI / box_size);
  for (int n_x = 0; n_x < slab_nx; n_x++)
    for (int n_y = 0; n_y < grid_dim; n_y++)
      for (int n_z = 0; n_z <= middle; n_z++) {
        const float k_z = n_z * delta_k;
        const long idx = n_z + (long)(middle + 1) * (n_y + (long)grid_dim * n_x);
        const float re = box[2 * idx], im = box[2 * idx + 1];
        box[2 * idx] = re * 0.0f - im * k_z;
        box[2 * idx + 1] = re * k_z + im * 0.0f;
      }
  return 0;
}

/* hubble(float z), XRayHeatingFunctions.c:443-453: 1 + z is a float sum */
static double tb_hubble(const mxr_tb_params* p, float z)
{
  return HUBBLE * p->Hubble_h * sqrt(p->OmegaM * pow(1 + z, 3) + p->OmegaR * pow(1 + z, 4) + p->OmegaLambda);
}

static long wrap_index(long k, int N) { return ((k % N) + N) % N; }

/* BrightnessTemperature.c:24-549 (no mini-haloes).  Destination cells of the line-of-sight
 * redistribution wrap modulo N, which equals the reference's periodic index wherever the
 * reference's own index stays inside the row. */
int oracle_brightness_temperature(const mxr_tb_params* p, const mxr_tb_grids* g, double* volume_ave_Tb)
{
  if (!p || !g || !g->xH || !g->deltax || !g->delta_T || (p->Flag_IncludeSpinTemp && !g->TS_box))
    return MXR_ERR_INVALID_ARG;
  if (p->Flag_IncludePecVelsFor21cm < 0 || p->Flag_IncludePecVelsFor21cm > 3)
    return MXR_ERR_INVALID_ARG;
  const int pecvel = p->Flag_IncludeSpinTemp ? p->Flag_IncludePecVelsFor21cm : 0;
  if (pecvel > 0 && !g->vel)
    return MXR_ERR_INVALID_ARG;
  const int N = p->ReionGridDim;
  const long H = N / 2 + 1;
  const double total_n_cells = pow((double)N, 3);
  const double BoxSize = p->BoxSize / p->Hubble_h;
  float redshift = (float)(float)p->redshift;
  float Hubble_h = (float)(float)p->Hubble_h;
  float OmegaM = (float)(float)p->OmegaM;
  float OmegaB = (float)(OmegaM * p->BaryonFrac);
  float const_factor = (float)(27.0 * (OmegaB * Hubble_h * Hubble_h / 0.023) *
                               sqrt((0.15 / OmegaM / Hubble_h / Hubble_h) * (1.0 + redshift) / 10.0));
  const float H_z = (float)(float)tb_hubble(p, redshift);
  double T_rad = TCMB * (1. + redshift);
  float* delta_T = g->delta_T;
  for (int ii = 0; ii < N; ii++)
    for (int jj = 0; jj < N; jj++)
      for (int kk = 0; kk < N; kk++) {
        const long i_real = kk + (long)N * (jj + (long)N * ii);
        const long i_padded = kk + 2 * H * (jj + (long)N * ii);
        double pixel_deltax = g->deltax[i_padded];
        double pixel_x_HI = g->xH[i_real];
        delta_T[i_real] = (float)(const_factor * pixel_x_HI * (1.0 + pixel_deltax));
        if (p->Flag_IncludeSpinTemp) {
          if (p->Flag_IncludePecVelsFor21cm > 1) {
            delta_T[i_real] *= (1. + redshift) / (1000. * g->TS_box[i_real]);
          } else {
            double pixel_Ts_factor = (1 - T_rad / g->TS_box[i_real]);
            delta_T[i_real] *= pixel_Ts_factor;
          }
        }
      }
  if (pecvel > 0) {
    const double subcell_width = (BoxSize / ((double)N)) / (double)N_RSD_STEPS;
    float* vel = g->vel;
    for (int ii = 0; ii < N; ii++)
      for (int jj = 0; jj < N; jj++)
        for (int kk = 0; kk < N; kk++) {
          const long i_padded = kk + 2 * H * (jj + (long)N * ii);
          switch (p->TreesID) {
            case MXR_TREES_GBPTREES:
              vel[i_padded] = (float)(sqrt(1. / (1. + redshift)) * vel[i_padded] / 1000.);
              break;
            case MXR_TREES_VELOCIRAPTOR:
            case MXR_TREES_VELOCIRAPTOR_AUG:
              vel[i_padded] = (float)((1. + redshift) * vel[i_padded]);
              break;
          }
          vel[i_padded] *= 1000. * 100. / MPC;
          vel[i_padded] = (float)(vel[i_padded] / (1. / (1. + redshift)));
        }
    float* grad = (float*)malloc(sizeof(float) * 2 * (size_t)N * N * H);
    if (!grad)
      return MXR_ERR_OOM;
    oracle_r2c3d_f32(N, vel, grad);
    for (long ii = 0; ii < (long)N * N * H; ii++) {
      grad[2 * ii] = (float)((double)grad[2 * ii] / total_n_cells);
      grad[2 * ii + 1] = (float)((double)grad[2 * ii + 1] / total_n_cells);
      grad[2 * ii] = (float)((double)grad[2 * ii] * p->Hubble_h);
      grad[2 * ii + 1] = (float)((double)grad[2 * ii + 1] * p->Hubble_h);
    }
    oracle_velocity_gradient(grad, N, N, p->BoxSize);
    oracle_c2r3d_f32(N, grad);
    const double max_v_deriv = fabs(MAX_DVDR * H_z);
    for (int ii = 0; ii < N; ii++)
      for (int jj = 0; jj < N; jj++)
        for (int kk = 0; kk < N; kk++) {
          const long i_real = kk + (long)N * (jj + (long)N * ii);
          const long i_padded = kk + 2 * H * (jj + (long)N * ii);
          if (pecvel == 1) {
            double dvdx = (double)grad[i_padded];
            if (fabs(dvdx) > max_v_deriv)
              dvdx = dvdx < 0 ? -max_v_deriv : max_v_deriv;
            delta_T[i_real] /= (dvdx / H_z + 1.0);
          } else {
            const double gradient_component = fabs((double)grad[i_padded] / H_z + 1.0);
            if (gradient_component < TB_FRACT_FLOAT_ERR)
              delta_T[i_real] = (float)(1000. * (g->TS_box[i_real] - T_rad) / (1. + redshift));
            else
              delta_T[i_real] = (float)((1. - exp(-delta_T[i_real] / gradient_component)) * 1000. *
                                        (g->TS_box[i_real] - T_rad) / (1. + redshift));
          }
        }
    free(grad);
    if (pecvel == 3) {
      float x_pos[N_RSD_STEPS], x_pos_offset[N_RSD_STEPS];
      const double x_val1 = 0., x_val2 = 1.;
      for (int iii = 0; iii < N_RSD_STEPS; iii++) {
        x_pos_offset[iii] = (float)(subcell_width * (float)iii + subcell_width / 2.);
        x_pos[iii] = (float)(x_pos_offset[iii] / (BoxSize / (float)N));
      }
      float* los = (float*)calloc((size_t)N, sizeof(float));
      if (!los)
        return MXR_ERR_OOM;
      const double cell = BoxSize / (float)N;
      const double frac_w = subcell_width / cell;
      for (int ii = 0; ii < N; ii++)
        for (int jj = 0; jj < N; jj++) {
          const long row = (long)N * (jj + (long)N * ii);
          const float* vrow = vel + 2 * H * (jj + (long)N * ii);
          for (int kk = 0; kk < N; kk++)
            los[kk] = 0.0f;
          for (int kk = 0; kk < N; kk++) {
            const float dT = delta_T[row + kk];
            if (!((fabs(dT) >= TB_FRACT_FLOAT_ERR) && (g->xH[row + kk] >= TB_FRACT_FLOAT_ERR)))
              continue;
            const double d1_low = vrow[kk == 0 ? N - 1 : kk - 1] / H_z;
            const double d2_low = vrow[kk] / H_z;
            const double d1_high = vrow[kk] / H_z;
            const double d2_high = vrow[kk == N - 1 ? 0 : kk + 1] / H_z;
            for (int iii = 0; iii < N_RSD_STEPS; iii++) {
              double disp;
              if (x_pos[iii] <= 0.5)
                disp = d1_low + ((x_pos[iii] + 0.5) - x_val1) * (d2_low - d1_low) / (x_val2 - x_val1);
              else
                disp = d1_high + ((x_pos[iii] - 0.5) - x_val1) * (d2_high - d1_high) / (x_val2 - x_val1);
              const double pos = (x_pos_offset[iii] + disp) / (BoxSize / (float)N);
              const double lo = pos - (subcell_width / 2.) / (BoxSize / (float)N);
              const double hi = pos + (subcell_width / 2.) / (BoxSize / (float)N);
              if (lo >= 0.0 && hi < 1.0) {
                los[kk] += dT / (float)N_RSD_STEPS;
              } else if (lo < 0.0 && hi < 0.0) {
                const double cd = ceil(fabs(lo)) - 1.;
                if (fabs(hi) > cd) {
                  los[wrap_index(kk - ((long)cd + 1), N)] += dT / (float)N_RSD_STEPS;
                } else {
                  const double fo = (fabs(lo) - cd) / frac_w;
                  const double fw = 1. - fo;
                  los[wrap_index(kk - (long)cd, N)] += fw * dT / (float)N_RSD_STEPS;
                  los[wrap_index(kk - ((long)cd + 1), N)] += fo * dT / (float)N_RSD_STEPS;
                }
              } else if (lo < 0.0 && (hi > 0.0 && hi < 1.0)) {
                const double fw = hi / frac_w;
                const double fo = 1. - fw;
                los[wrap_index(kk - 1, N)] += fo * dT / (float)N_RSD_STEPS;
                los[kk] += fw * dT / (float)N_RSD_STEPS;
              } else if ((lo >= 0.0 && lo < 1.0) && (hi >= 1.0)) {
                const double fo = (hi - 1.) / frac_w;
                const double fw = 1. - fo;
                los[kk] += fw * dT / (float)N_RSD_STEPS;
                los[wrap_index(kk + 1, N)] += fo * dT / (float)N_RSD_STEPS;
              } else {
                const double cd = floor(fabs(hi));
                if (lo >= cd) {
                  los[wrap_index(kk + (long)cd, N)] += dT / (float)N_RSD_STEPS;
                } else {
                  const double fo = (hi - cd) / frac_w;
                  const double fw = 1. - fo;
                  los[wrap_index(kk + (long)cd - 1, N)] += fw * dT / (float)N_RSD_STEPS;
                  los[wrap_index(kk + (long)cd, N)] += fo * dT / (float)N_RSD_STEPS;
                }
              }
            }
          }
          for (int kk = 0; kk < N; kk++)
            delta_T[row + kk] = los[kk];
        }
      free(los);
    }
  }
  double Ave_Tb = 0.0;
  for (long i = 0; i < (long)N * N * N; i++)
    Ave_Tb += (double)delta_T[i];
  Ave_Tb /= total_n_cells;
  if (volume_ave_Tb)
    *volume_ave_Tb = Ave_Tb;
  return 0;
}

/* ---------------------------------------------------------------- power spectrum */
/* Initialise_PowerSpectrum (ComputePowerSpectrum.c:18-42): bins of k_ceil = delta_k 1.35^i in
 * float while k_ceil < k_max = delta_k N, with delta_k = (float)(2 pi / (BoxSize / h)). */
int oracle_ps_length(int N, double BoxSize, double Hubble_h)
{
  const double box_size = BoxSize / Hubble_h;
  const float k_factor = 1.35;
  const float delta_k = (float)(2. * M_PI / box_size);
  const float k_max = delta_k * N;
  int num_bins = 0;
  float k_ceil = delta_k;
  while (k_ceil < k_max) {
    num_bins++;
    k_ceil *= k_factor;
  }
  return num_bins;
}

/* Compute_PS (ComputePowerSpectrum.c:44-255, no mini-haloes): delta_T (real layout) ->
 * PS_k, PS_data, PS_error (PS_Length each). */
int oracle_power_spectrum(int N, double BoxSize, double Hubble_h, const float* delta_T, float* PS_k, float* PS_data,
                          float* PS_error, int PS_Length)
{
  if (N < 2 || !delta_T || !PS_k || !PS_data || !PS_error || PS_Length != oracle_ps_length(N, BoxSize, Hubble_h))
    return MXR_ERR_INVALID_ARG;
  const long H = N / 2 + 1;
  const double box_size = BoxSize / Hubble_h;
  const float volume = powf((float)(float)box_size, 3);
  const double total_n_cells = pow((double)N, 3);
  double ave = 0.0;
  for (long i = 0; i < (long)N * N * N; i++)
    ave += delta_T[i];
  ave /= total_n_cells;
  float* dd = (float*)calloc((size_t)2 * N * N * H, sizeof(float));
  float* spec = (float*)malloc(sizeof(float) * 2 * (size_t)N * N * H);
  if (!dd || !spec) {
    free(dd);
    free(spec);
    return MXR_ERR_OOM;
  }
  for (int ii = 0; ii < N; ii++)
    for (int jj = 0; jj < N; jj++)
      for (int kk = 0; kk < N; kk++) {
        const long ir = kk + (long)N * (jj + (long)N * ii), ip = kk + 2 * H * (jj + (long)N * ii);
        dd[ip] = (float)((delta_T[ir] / ave - 1) * volume / (float)total_n_cells);
        dd[ip] *= ave;
      }
  oracle_r2c3d_f32(N, dd, spec);
  free(dd);
  const float k_factor = 1.35;
  const float delta_k = (float)(2. * M_PI / box_size);
  const float k_first_bin_ceil = delta_k;
  const float k_max = delta_k * N;
  double* p_box = (double*)calloc((size_t)PS_Length, sizeof(double));
  double* k_ave = (double*)calloc((size_t)PS_Length, sizeof(double));
  unsigned long long* in_bin_ct = (unsigned long long*)calloc((size_t)PS_Length, sizeof(unsigned long long));
  const int HII_middle = N / 2;
  for (int n_x = 0; n_x < N; n_x++) {
    float k_x = n_x * delta_k;
    if (n_x > HII_middle)
      k_x = (n_x - N) * delta_k;
    for (int n_y = 0; n_y < N; n_y++) {
      float k_y = n_y * delta_k;
      if (n_y > HII_middle)
        k_y = (n_y - N) * delta_k;
      for (int n_z = 0; n_z <= HII_middle; n_z++) {
        const float k_z = n_z * delta_k;
        const float k_mag = (float)sqrt(k_x * k_x + k_y * k_y + k_z * k_z);
        int ct = 0;
        float k_floor = 0;
        float k_ceil = k_first_bin_ceil;
        while (k_ceil < k_max) {
          if ((k_mag >= k_floor) && (k_mag < k_ceil)) {
            const long idx = n_z + H * (n_y + (long)N * n_x);
            const double re = spec[2 * idx], im = spec[2 * idx + 1];
            in_bin_ct[ct]++;
            p_box[ct] += pow(k_mag, 3) * pow(hypot(re, im), 2.) / (2.0 * M_PI * M_PI * volume);
            k_ave[ct] += k_mag;
            break;
          }
          ct++;
          k_floor = k_ceil;
          k_ceil *= k_factor;
        }
      }
    }
  }
  free(spec);
  for (int ii = 0; ii < PS_Length; ii++) {
    PS_k[ii] = (float)(k_ave[ii] / (double)in_bin_ct[ii]);
    PS_data[ii] = (float)(p_box[ii] / (double)in_bin_ct[ii]);
    PS_error[ii] = (float)((p_box[ii] / (double)in_bin_ct[ii]) / sqrt((double)in_bin_ct[ii]));
  }
  free(p_box);
  free(k_ave);
  free(in_bin_ct);
  return 0;
}
