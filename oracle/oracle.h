/*
 * oracle.h -- CPU restatement of the reference's reionization grid path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load liboracle.so, and only as the checker of (or the CPU timing beside)
 * the HIP path.  The product (libmxr.so) never links, loads or falls back to it.
 *
 * Two restatements of the same pipeline:
 *   *_f32 : faithful to the reference's types -- FP32 FFTs (FFTW's fftwf semantics), filter
 *           arithmetic in float exactly as written in src/core/reionization.c:3373-3433,
 *           clamps in float, the ionisation criterion in double.
 *   *_f64 : every field, filter and FFT in double (the FP64 oracle the errors are quoted
 *           against).
 */
#ifndef MXR_ORACLE_H
#define MXR_ORACLE_H

#include <stdio.h>
#include <stdlib.h>

#include "../include/mxr.h"

/* One line on stderr per radius or shell when ORACLE_PROGRESS is set, so that a long oracle
 * run at the benchmark sizes shows that it is moving. */
static inline void oracle_progress(const char* what, int i, int n)
{
  static int on = -1;
  if (on < 0)
    on = getenv("ORACLE_PROGRESS") != NULL;
  if (on) {
    fprintf(stderr, "oracle %s %d/%d\n", what, i, n);
    fflush(stderr);
  }
}

#ifdef __cplusplus
extern "C" {
#endif

/* FFT (FFTW semantics, unnormalised).  data is interleaved (re, im). */
int oracle_c2c1d_f32(int n, float* data, int sign);
int oracle_c2c1d_f64(int n, double* data, int sign);
int oracle_r2c3d_f32(int N, const float* in_padded, float* out_cplx);
int oracle_r2c3d_f64(int N, const double* in_padded, double* out_cplx);
int oracle_c2r3d_f32(int N, float* inout);
int oracle_c2r3d_f64(int N, double* inout);

/* filter() on a half spectrum (float complex / double complex). */
int oracle_filter_f32(float* box, int local_ix_start, int slab_nx, int grid_dim, double BoxSize, float R,
                      int filter_type);
int oracle_filter_f64(double* box, int local_ix_start, int slab_nx, int grid_dim, double BoxSize, double R,
                      int filter_type);

/* Physics helpers. */
double oracle_RtoM(const mxr_hii_params* p, double R);
int oracle_hii_radii(const mxr_hii_params* p, double* radii, int capacity);
double oracle_T_RECFAST(const char* tables_dir, float z);         /* < 0 on failure */
double oracle_spline_natural(const double* x, const double* y, int n, double xq);
double oracle_hubble_time(const mxr_hii_params* p);                /* internal units */
double oracle_rho_crit(const mxr_hii_params* p);                   /* internal units */

/* The pipeline.  Pointers are host arrays in the reference layouts (mxr.h).
 * margin (optional, N^3 doubles): max over radii of (LHS/RHS - 1) of the criterion, the
 * distance of each cell from the ionisation threshold. */
int oracle_find_hii_bubbles_f32(const mxr_hii_params* p, const mxr_hii_grids* g, mxr_hii_globals* out,
                                double* margin);
int oracle_find_hii_bubbles_f64(const mxr_hii_params* p, const mxr_hii_grids* g, mxr_hii_globals* out,
                                double* margin);

/* Bounded timing sample: forward transforms plus the first n_radii radii of the f32
 * restatement (n_radii <= 0: all).  Returns radii done; *seconds gets the wall time of the
 * sample and *seconds_fwd the part spent before the radius loop. */
int oracle_hii_sample_f32(const mxr_hii_params* p, const mxr_hii_grids* g, int n_radii, double* seconds,
                          double* seconds_fwd);

/* The f32 restatement with its transforms on FFTW3's float API (libmkl_rt's export, loaded at
 * run time; ORACLE_FFTW_LIB names another library): the CPU baseline of the benchmark.
 * *seconds gets the call's wall time (forward transforms to globals).  MXR_ERR_UNSUPPORTED when
 * the library is missing. */
int oracle_find_hii_bubbles_fftw(const mxr_hii_params* p, const mxr_hii_grids* g, mxr_hii_globals* out,
                                 double* seconds);
int oracle_fftw_available(void);
const char* oracle_fftw_library(void);

int oracle_brightness_temperature(const mxr_tb_params* p, const mxr_tb_grids* g, double* volume_ave_Tb);
int oracle_velocity_gradient(float* box, int slab_nx, int grid_dim, double BoxSize);
/* Input grids: smooth_grid + subsample_grid + the density step (one rank); out is padded N^3. */
int oracle_prepare_grid(int N, const float* hi, int n_hi, int property, double file_box, const mxr_grid_params* p,
                        float* out);
int oracle_read_grid_gbptrees(const char* fname, int property, int N, const mxr_grid_params* p, float* out);

/* construct_baryon_grids for one rank (reionization.c:2025-2591), host arrays. */
int oracle_construct_baryon_grids(const mxr_baryon_params* p, const mxr_galaxies* g, const mxr_baryon_grids* out,
                                  long* n_below_bh_limit);

int oracle_ps_length(int N, double BoxSize, double Hubble_h);
int oracle_power_spectrum(int N, double BoxSize, double Hubble_h, const float* delta_T, float* PS_k, float* PS_data,
                          float* PS_error, int PS_Length);

/* One spin-temperature step on host grids, with the grid-independent tables `tb` the library
 * reported for the same call (mxr_ts_tables), AGN X-ray bands as p->Flag_IncludeAGNXray asks.
 * red0 (optional, 5 doubles) receives the R_ct = 0 averages: collapse fraction, x_e, X-ray
 * emissivity, AGN soft and hard emissivity; smoothed (optional, G F N^3 floats, G = 2 + the AGN
 * bands) the smoothed SFR (clamped), X-ray, then AGN soft and hard (clamped) shells, real
 * layout. */
int oracle_ts_step_f32(const mxr_ts_params* p, const mxr_ts_tables* tb, const mxr_ts_grids* g,
                       mxr_ts_globals* out, double* red0, float* smoothed);
/* The same with the shells smoothed in FP64 (double source grids, FFTs, filter, clamp): the
 * FP64 anchor the FP32 paths' errors are quoted against. */
int oracle_ts_step_f64(const mxr_ts_params* p, const mxr_ts_tables* tb, const mxr_ts_grids* g,
                       mxr_ts_globals* out, double* red0, float* smoothed);

/* Helpers on their own, for the tests that pin them. */
float oracle_fully_ionized_T(float z_re, float z, float delta);
float oracle_partially_ionized_T(float T_HI, float res_xH);
int oracle_ts_probe(const mxr_ts_params* p, int what, const double* args, double* out);

int oracle_num_threads(void);
/* Recombinations (recomb.c): the tables find_HII_bubbles interpolates (the library's, as
 * mxr_recombination_tables reports them: 150 ln Gamma_12 nodes, then 200 x 100 x 150 doubles
 * each), the per-cell lookup, dt/dz, and probes of the table integrands' pieces. */
int oracle_rec_set_tables(const double* lnGamma, const double* ln_rate, const double* clumping,
                          const double* ln_residual);
int oracle_splined_recombination(double z_eff, double gamma12_bg, double temp, double* recombination_rate,
                                 double* residual_xH, double* clumping_factor);
double oracle_dtdz(double OmegaM, double OmegaLambda, double Hubble_h, float z);
int oracle_rec_probe(const mxr_hii_params* p, int what, const double* args, double* out);

#ifdef __cplusplus
}
#endif

#endif
