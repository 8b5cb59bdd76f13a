/*
 * mxr.h -- C ABI of the MI355X (gfx950) reionization grid path.
 *
 * Plain C types only: pointers, sizes, ints and doubles.  Every entry point returns
 * MXR_OK (0) on success and a negative MXR_ERR_* code otherwise; mxr_strerror() maps a
 * code to text, mxr_last_error() gives the detail of the last failure on this thread.
 *
 * Grid layouts are the reference's (src/core/misc_tools.c:244-264):
 *   padded real   I = k + 2*(N/2+1)*(j + N*i)     (deltax, stars, weighted_sfr, x_e_box)
 *   real          I = k + N*(j + N*i)             (xH, z_at_ionization, J_21_at_ionization,
 *                                                  r_bubble, temp_kinetic_all_gas, Tk_box, ...)
 *   half spectrum I = k + (N/2+1)*(j + N*i)       (complex64, FFTW r2c/c2r semantics)
 * with i the slowest (x) and k the fastest (z) index.  One GPU holds the whole grid:
 * slab_nix = N, slab_ix_start = 0.
 */
#ifndef MXR_H
#define MXR_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------------------- */
#define MXR_OK 0
#define MXR_ERR_INVALID_ARG (-1)  /* bad pointer, size or parameter value                 */
#define MXR_ERR_UNSUPPORTED (-2)  /* a flag combination this library does not implement    */
#define MXR_ERR_HIP (-3)          /* a HIP runtime call failed                              */
#define MXR_ERR_OOM (-4)          /* device allocation failed                               */
#define MXR_ERR_TABLES (-5)       /* a physics table could not be read or is out of range   */
#define MXR_ERR_NO_DEVICE (-6)    /* no usable gfx950 device                               */
#define MXR_ERR_FILTER (-7)       /* ReionFilterType not in {0,1,2} (reference aborts)      */

/* ---- memory location of the pointers handed to an entry point ------------------------- */
#define MXR_MEM_HOST 0
#define MXR_MEM_DEVICE 1

/* ---- execution path ------------------------------------------------------------------- */
/* FUSED: filter fused into the first inverse-FFT pass, clamp + criterion + state update
 *        fused into the last pass, FP64 reductions fused into the last radius, and for the
 *        sharp-k filter only the modes the filter keeps are read (the rest are exact zeros).
 * PLAIN: the reference's one-kernel-per-step shape (copy, filter, 3-pass c2r, clamp,
 *        criterion kernel per radius; src/core/find_HII_bubbles_gpu.cu:318-521).
 * Both paths produce bit-identical filtered fields; they exist so the speed-up of the fused
 * design can be measured on the same GPU.                                                    */
#define MXR_PATH_FUSED 0
#define MXR_PATH_PLAIN 1
#define MXR_PATH_FUSED_DENSE 2 /* FUSED without skipping the sharp-k zeroed modes */

/* ---- parameters (names follow run_params_t / physics_params_t / run_units_t,
 *      src/meraxes.h:142-201, 262-398) ------------------------------------------------- */
typedef struct mxr_hii_params {
  int ReionGridDim;             /* N; N even, N = 2^a 3^b 5^c, 16 <= N <= 1024           */
  double BoxSize;               /* Mpc/h                                                  */
  double Hubble_h;
  double OmegaM;
  double OmegaLambda;
  double OmegaK;
  double BaryonFrac;
  double ReionDeltaRFactor;     /* radius step factor (1.1)                              */
  int ReionFilterType;          /* 0 real-space top-hat, 1 sharp-k, 2 Gaussian            */
  int ReionRtoMFilterType;      /* 0 top-hat, 1 Gaussian                                  */
  int ReionUVBFlag;             /* 0, 1 (J21 at ionisation) or 2 (recombinations only)   */
  int Flag_IncludeSpinTemp;     /* x_e_box becomes a 4th filtered grid, Tk_box is used    */
  int Flag_IncludeRecombinations; /* must be 0 (MXR_ERR_UNSUPPORTED otherwise)            */
  int Flag_EvolvingReionRBubbleMax;
  int Flag_BHFeedback;          /* must be 0 (MXR_ERR_UNSUPPORTED otherwise)              */
  double ReionSfrTimescale;     /* in units of the Hubble time                            */
  double ReionEfficiency;
  double ReionNionPhotPerBary;
  double ReionGammaHaloBias;
  double ReionAlphaUV;
  double ReionRBubbleMax;       /* Mpc/h                                                  */
  double ReionRBubbleMaxRecomb; /* Mpc/h                                                  */
  double ReionRBubbleMin;       /* Mpc/h                                                  */
  double UnitLength_in_cm;
  double UnitMass_in_g;
  double UnitVelocity_in_cm_per_s;
  double redshift;              /* ZZ[snapshot]                                           */
  double prev_redshift;         /* ZZ[snapshot-1] (ZZ[0] for snapshot 0)                  */
  double TK_recfast;            /* T_RECFAST(redshift) in K; if < 0 the library computes
                                   it from TablesForXHeatingDir/recfast_LCDM.dat          */
  const char* TablesForXHeatingDir; /* directory holding recfast_LCDM.dat (may be NULL
                                       when TK_recfast >= 0)                              */
} mxr_hii_params;

/* In/out grids of one find_HII_bubbles call (reion_grids_t fields, src/meraxes.h:448-681). */
typedef struct mxr_hii_grids {
  int mem;                      /* MXR_MEM_HOST or MXR_MEM_DEVICE for every pointer below  */
  const float* deltax;          /* in  padded real                                        */
  const float* stars;           /* in  padded real                                        */
  const float* weighted_sfr;    /* in  padded real                                        */
  const float* x_e_box;         /* in  padded real (Flag_IncludeSpinTemp only, else NULL) */
  const float* Tk_box;          /* in  real        (Flag_IncludeSpinTemp only, else NULL) */
  float* xH;                    /* out real                                               */
  float* z_at_ionization;       /* in/out real (persists across calls, starts at -1)      */
  float* J_21_at_ionization;    /* in/out real                                            */
  float* r_bubble;              /* out real                                               */
  float* temp_kinetic_all_gas;  /* in/out real                                            */
} mxr_hii_grids;

/* Globals written back by the reference (src/core/find_HII_bubbles.c:686-702). */
typedef struct mxr_hii_globals {
  double volume_weighted_global_xH;
  double volume_weighted_global_Gamma12;
  double volume_weighted_global_r_bubble;
  double volume_weighted_global_weighted_sfr;
  double volume_weighted_global_effective_bhar;
  double volume_weighted_global_temp_kinetic_all_gas;
  double volume_weighted_global_N_rec;
  double volume_weighted_global_residual_xH;
  double volume_weighted_global_clumping_factor;
  double mass_weighted_global_xH;
  double mass_weighted_global_Gamma12;
  double mass_weighted_global_r_bubble;
  double mass_weighted_global_temp_kinetic_all_gas;
  double mass_weighted_global_N_rec;
  double mass_weighted_global_residual_xH;
  double mass_weighted_global_clumping_factor;
  int n_radii;                  /* filter radii processed (the last one unfiltered)       */
  double device_ms;             /* device time of the call, inputs resident in HBM        */
} mxr_hii_globals;

/* ---- context ---------------------------------------------------------------------------- */
typedef struct mxr_ctx mxr_ctx;

/* Create a context on HIP device `device` for grids of dimension `grid_dim`.  The context
 * owns the stream, twiddle tables and every device buffer the path needs, allocated once
 * (the reference allocates per call: src/core/find_HII_bubbles_gpu.cu:150-189).
 * Replaces init_CUDA (src/meraxes_gpu.h:55, src/core/meraxes_gpu.cu:127-194) plus the
 * grid/plan set-up of malloc_reionization_grids (src/core/reionization.c:718-1000).      */
int mxr_create(mxr_ctx** ctx, int device, int grid_dim);
void mxr_destroy(mxr_ctx* ctx);
const char* mxr_strerror(int code);
const char* mxr_last_error(void);

/* Select MXR_PATH_FUSED (default) or MXR_PATH_PLAIN. */
int mxr_set_path(mxr_ctx* ctx, int path);

/* ---- the hot path ------------------------------------------------------------------------ */

/* One excursion-set ionisation step: forward r2c of deltax/stars/weighted_sfr (+x_e_box),
 * then for each radius from R_max down to the cell scale the k-space filter, c2r, aliasing
 * clamp and per-cell criterion with xH / z_at_ionization / J_21_at_ionization updates, then
 * the FP64 volume- and mass-weighted globals.
 * Replaces _find_HII_bubbles (src/core/find_HII_bubbles.c:86-710) and
 * _find_HII_bubbles_gpu (src/core/find_HII_bubbles_gpu.cu:68-633); called where
 * find_HII_bubbles (src/core/find_HII_bubbles.c:713-747) dispatches.                     */
int mxr_find_hii_bubbles(mxr_ctx* ctx, const mxr_hii_params* params, const mxr_hii_grids* grids,
                         mxr_hii_globals* globals);

/* In-place k-space window on a half spectrum of dimension grid_dim (Nc complex64 values,
 * `mem` says where `box` lives).  Replaces filter() (src/core/reionization.c:3373-3433,
 * declared src/core/reionization.h:87); BoxSize is run_globals.params.BoxSize.          */
int mxr_filter(mxr_ctx* ctx, void* box, int mem, int local_ix_start, int slab_nx, int grid_dim,
               double BoxSize, float R, int filter_type);

/* Mass enclosed by a filter of radius R (internal units).  Replaces RtoM
 * (src/core/find_HII_bubbles.c:27-46, declared src/core/find_HII_bubbles.h:11).          */
double mxr_RtoM(const mxr_hii_params* params, double R);

/* Number of filter radii find_HII_bubbles visits for these parameters, and (if radii is not
 * NULL and capacity allows) the radii themselves in Mpc/h, last (unfiltered) one included
 * (src/core/find_HII_bubbles.c:256-285, 546).                                             */
int mxr_hii_radii(const mxr_hii_params* params, double* radii, int capacity);

/* ---- 21-cm brightness temperature ---------------------------------------------------------- */
typedef struct mxr_tb_params {
  int ReionGridDim;
  double Hubble_h;
  double OmegaM;
  double BaryonFrac;
  double redshift;
  int Flag_IncludeSpinTemp;
  int Flag_IncludePecVelsFor21cm; /* must be 0 (MXR_ERR_UNSUPPORTED otherwise)            */
} mxr_tb_params;

typedef struct mxr_tb_grids {
  int mem;
  const float* xH;      /* real                                  */
  const float* deltax;  /* padded real                           */
  const float* TS_box;  /* real (Flag_IncludeSpinTemp only)      */
  float* delta_T;       /* out real                              */
} mxr_tb_grids;

/* delta_T = T0 xH (1+delta) [(1 - T_CMB/T_S)] per cell and its FP64 volume mean.
 * Replaces ComputeBrightnessTemperatureBox (src/core/BrightnessTemperature.c:24-549,
 * declared src/core/BrightnessTemperature.h:9) for Flag_IncludePecVelsFor21cm = 0.       */
int mxr_brightness_temperature(mxr_ctx* ctx, const mxr_tb_params* params, const mxr_tb_grids* grids,
                               double* volume_ave_Tb);

/* ---- per-kernel timing (hipEvents on the context's stream) ------------------------------ */
/* One row per kernel class of the path: launches, summed device time and the algorithmic
 * HBM bytes those launches must move (inputs read + outputs written, modes the sharp-k window
 * zeroes excluded).  Replaces the reference's mlog/timer_info timing of the solver call
 * (src/core/find_HII_bubbles.c:713-747, src/core/utils.c:41-56) at kernel granularity.     */
typedef struct mxr_kernel_stat {
  char name[32];
  int launches;
  double total_ms;
  double alg_bytes;
} mxr_kernel_stat;

/* on != 0 enables event timing of every launch (and clears the counters); 0 disables it. */
int mxr_profile_enable(mxr_ctx* ctx, int on);
/* Copies up to `capacity` rows into out; returns the number of rows (kernel classes). */
int mxr_profile_read(mxr_ctx* ctx, mxr_kernel_stat* out, int capacity);

#ifdef __cplusplus
}
#endif

#endif /* MXR_H */
