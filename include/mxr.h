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
#define MXR_ERR_NO_WISDOM (-8)    /* the wisdom file holds no entry for this device and shape */

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
  int ReionGridDim;             /* N: one of the dimensions mxr_supported_dims() lists   */
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
  int Flag_IncludeRecombinations; /* N_rec becomes a filtered grid; Gamma12, r_bubble,
                                   UVBFlag 2 and the post-loop recombination update apply */
  int Flag_EvolvingReionRBubbleMax;
  int Flag_BHFeedback;          /* effective_bhm / effective_bhar become filtered grids    */
  double ReionSfrTimescale;     /* in units of the Hubble time                            */
  double ReionEfficiency;
  double ReionNionPhotPerBary;
  double ReionGammaHaloBias;
  double ReionAlphaUV;
  double ReionAlphaUVBH;        /* J21 spectral index of the BH sources (Flag_BHFeedback) */
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
  int Flag_TemperatureDependentRec; /* recombinations: the post-loop rate at the cell's gas
                                       temperature (1) or at 1e4 K (0)                     */
  double Y_He;                  /* helium mass fraction (physics.Y_He; recombinations only) */
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
  int storage;                  /* MXR_STORAGE_F32 (0), or MXR_STORAGE_BF16: stars and
                                   weighted_sfr hold bfloat16 bit patterns (uint16, same padded
                                   layout), widened to FP32 on load; fused paths only        */
  const float* effective_bhm;   /* in  padded real (Flag_BHFeedback only, else NULL)      */
  const float* effective_bhar;  /* in  padded real (Flag_BHFeedback only, else NULL)      */
  /* Flag_IncludeRecombinations only (else NULL): */
  float* N_rec;                 /* in/out padded real: recombinations per baryon (starts 0) */
  float* Gamma12;               /* in/out real: ionising background at the first crossing
                                   (1e-12/s h^-2; persists across calls, starts 0)          */
  float* residual_xH;           /* out real: 1e4 x residual neutral fraction               */
  float* clumping_factor;       /* out real                                              */
  float* t_resp;                /* out real: 1 / (Gamma + recombination rate), Myr        */
} mxr_hii_grids;
#define MXR_STORAGE_F32 0
#define MXR_STORAGE_BF16 1

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
/* The grid dimensions this build compiles kernels for, in increasing order:
 *   16, 32, 64, 96, 128, 160, 256, 512, 1024
 * (the reference plans FFTW for any ReionGridDim, src/core/reionization.c:907-997; here each
 * transform length is a compile-time radix plan).  Writes up to `capacity` of them into dims
 * (may be NULL) and returns how many there are.  mxr_create rejects any other dimension with
 * MXR_ERR_INVALID_ARG.  Needs no GPU.                                                      */
int mxr_supported_dims(int* dims, int capacity);
void mxr_destroy(mxr_ctx* ctx);
const char* mxr_strerror(int code);
const char* mxr_last_error(void);

/* Select MXR_PATH_FUSED (default) or MXR_PATH_PLAIN. */
int mxr_set_path(mxr_ctx* ctx, int path);

/* Launch shapes of the fused path, per context.  No counterpart in the reference, where the
 * FFTW plans and their wisdom file fill this role (src/core/reionization.c:726-758, 907-997).
 * Every field has a gfx950 default.                                                        */
typedef struct mxr_tuning {
  int zdirect_kmax; /* z pass by direct summation on sharp-k radii keeping <= this many k_z
                       (0: never; at most 64; applies when the grid dimension is a multiple
                       of 256)                                                               */
  int zdirect_rpb;  /* rows per wavefront of the direct z pass (1 .. 4096)                  */
  int live_tiles;   /* 1: x and y pass grids cover only the tiles that can hold kept sharp-k
                       columns; 0: every tile is launched and empty ones return at once    */
  int col_tc_x;     /* columns per workgroup of the per-radius x pass (4, 8, 16 or 32)      */
  int col_tc_y;     /* columns per workgroup of the per-radius y pass (4, 8, 16 or 32)      */
} mxr_tuning;
int mxr_get_tuning(const mxr_ctx* ctx, mxr_tuning* t);
int mxr_set_tuning(mxr_ctx* ctx, const mxr_tuning* t);

/* Autotuner: time candidate launch shapes with find_HII_bubbles on these inputs (device or
 * host grids; the calls run on scratch state, the caller's state grids are left untouched),
 * keep the fastest in the context, return it in `chosen` (may be NULL) and, when
 * `wisdom_path` is not NULL, merge it into that JSON wisdom file under the key (device
 * architecture, compute units, grid dimension, ReionFilterType, Flag_IncludeSpinTemp).
 * Takes the place of the reference's FFTW planning with its wisdom file
 * (src/core/reionization.c:726-758, 907-997).                                              */
int mxr_autotune(mxr_ctx* ctx, const mxr_hii_params* params, const mxr_hii_grids* grids, const char* wisdom_path,
                 mxr_tuning* chosen);
/* Apply the wisdom file's entry for this context's device and grid dimension and the
 * params' filter type and spin flag; MXR_ERR_NO_WISDOM (tuning unchanged) when it has none. */
int mxr_load_wisdom(mxr_ctx* ctx, const mxr_hii_params* params, const char* wisdom_path);
/* The wisdom file itself, without a device: look an entry up, or add / replace one.      */
int mxr_wisdom_lookup(const char* wisdom_path, const char* arch, int cus, int grid_dim, int filter_type, int spin,
                      mxr_tuning* t);
int mxr_wisdom_store(const char* wisdom_path, const char* arch, int cus, int grid_dim, int filter_type, int spin,
                     const mxr_tuning* t, double ms);

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

/* velocity_gradient (src/core/reionization.c:3435-3450, declared reionization.h:88): each
 * half-spectrum value of slab_nx x-planes times i k_z, k_z = n_z (float)(2 pi/(float)BoxSize).
 * box: complex64 half spectrum, host or device (mem).                                      */
int mxr_velocity_gradient(mxr_ctx* ctx, void* box, int mem, int slab_nx, int grid_dim, double BoxSize);

/* Mass enclosed by a filter of radius R (internal units).  Replaces RtoM
 * (src/core/find_HII_bubbles.c:27-46, declared src/core/find_HII_bubbles.h:11).          */
double mxr_RtoM(const mxr_hii_params* params, double R);

/* The recombination tables find_HII_bubbles uses with Flag_IncludeRecombinations, for the
 * params' cosmology (Hubble_h, OmegaM, BaryonFrac, Y_He): per (z_ct, t_ct, g) with z = 0.2 z_ct
 * (z_ct < 200), log10 T = 2 + 0.03 t_ct (t_ct < 100), ln Gamma_12 = -10 + 0.1 g (g < 150), the
 * integrals over the MHR00 density PDF with self-shielding of the recombination rate per baryon
 * (ln, 1/s), the clumping factor, and 1e4 times the residual neutral fraction (ln), case B.
 * Built on the GPU on first use and kept with the context.  Each output (may be NULL) receives
 * 200 x 100 x 150 doubles, z_ct slowest; lnGamma the 150 nodes.  No counterpart entry point in
 * the reference, which keeps them in recombinations.c (init_MHR, :208-374) or an HDF5 cache. */
int mxr_recombination_tables(mxr_ctx* ctx, const mxr_hii_params* params, double* lnGamma, double* ln_rate,
                             double* clumping, double* ln_residual);

/* Number of filter radii find_HII_bubbles visits for these parameters, and (if radii is not
 * NULL and capacity allows) the radii themselves in Mpc/h, last (unfiltered) one included
 * (src/core/find_HII_bubbles.c:256-285, 546).                                             */
int mxr_hii_radii(const mxr_hii_params* params, double* radii, int capacity);

/* ---- batches of independent realisations ----------------------------------------------------- */
/* A batch owns one context per realisation (grid dimensions may differ).  Modes:
 *   MXR_BATCH_GRAPH   every realisation's launch sequence captured as one branch of a single
 *                     hipGraph (fork/join), instantiated once and launched; a call with the
 *                     same parameters and grid pointers as the previous one relaunches the
 *                     instantiated graph without capturing again;
 *   MXR_BATCH_STREAMS the same sequences launched directly, one stream per realisation;
 *   MXR_BATCH_SERIAL  one realisation at a time, each call synchronised (the comparator).
 * Grids must be device memory.  globals[i].device_ms holds the whole batch's device time.
 * No counterpart in the reference (it runs one grid per call); the realisations' results are
 * identical to n separate mxr_find_hii_bubbles calls.                                        */
#define MXR_BATCH_GRAPH 0
#define MXR_BATCH_STREAMS 1
#define MXR_BATCH_SERIAL 2
typedef struct mxr_batch mxr_batch;
int mxr_batch_create(mxr_batch** batch, int device, int n, const int* grid_dims);
void mxr_batch_destroy(mxr_batch* batch);
int mxr_find_hii_bubbles_batched(mxr_batch* batch, const mxr_hii_params* params, const mxr_hii_grids* grids,
                                 mxr_hii_globals* globals, int mode);
/* mxr_autotune once per distinct (grid dimension, filter type, spin flag) of the batch, on the
 * first realisation with that shape, applied to every realisation of that shape; and
 * mxr_load_wisdom for each realisation (MXR_ERR_NO_WISDOM if any has no entry; the others
 * are applied).                                                                              */
int mxr_batch_autotune(mxr_batch* batch, const mxr_hii_params* params, const mxr_hii_grids* grids,
                       const char* wisdom_path);
int mxr_batch_load_wisdom(mxr_batch* batch, const mxr_hii_params* params, const char* wisdom_path);

/* ---- 21-cm brightness temperature ---------------------------------------------------------- */
/* run_params_t.TreesID (src/meraxes.h:254-258): selects the velocity grid's units */
#define MXR_TREES_VELOCIRAPTOR 0
#define MXR_TREES_GBPTREES 1
#define MXR_TREES_VELOCIRAPTOR_AUG 2

typedef struct mxr_tb_params {
  int ReionGridDim;
  double BoxSize;                 /* Mpc/h                                                   */
  double Hubble_h;
  double OmegaM;
  double OmegaLambda;
  double OmegaR;
  double BaryonFrac;
  double redshift;
  int Flag_IncludeSpinTemp;
  int Flag_IncludePecVelsFor21cm; /* 0..3; velocity terms apply with Flag_IncludeSpinTemp    */
  int TreesID;                    /* MXR_TREES_*                                             */
} mxr_tb_params;

typedef struct mxr_tb_grids {
  int mem;
  const float* xH;      /* real                                                          */
  const float* deltax;  /* padded real                                                   */
  const float* TS_box;  /* real (Flag_IncludeSpinTemp only)                              */
  float* vel;           /* in/out padded real, line-of-sight (z) velocity (PecVels > 0 with
                           Flag_IncludeSpinTemp only): converted in place to comoving
                           velocity as the reference does                                */
  float* delta_T;       /* out real                                                      */
} mxr_tb_grids;

/* delta_T = T0 xH (1+delta) with the spin-temperature factor (1 - T_CMB/T_S), or, for
 * Flag_IncludePecVelsFor21cm > 1, the optical-depth form; then the peculiar-velocity terms:
 * the velocity gradient dv/dz from the velocity grid (r2c, /N^3, x Hubble_h,
 * velocity_gradient, c2r), capped at 0.2 H(z) for PecVels = 1, uncapped in the optical
 * depth for 2 and 3, and for 3 the line-of-sight redshift-space redistribution in 50
 * sub-cells per cell.  Returns the FP64 volume mean.
 * Replaces ComputeBrightnessTemperatureBox (src/core/BrightnessTemperature.c:24-549,
 * declared src/core/BrightnessTemperature.h:9), without mini-haloes.                    */
int mxr_brightness_temperature(mxr_ctx* ctx, const mxr_tb_params* params, const mxr_tb_grids* grids,
                               double* volume_ave_Tb);

/* The same brightness box for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  No context: grids->mem must be MXR_MEM_DEVICE, volume_ave_Tb is a device pointer
 * that receives the FP64 mean, and the work is enqueued on `stream` (a hipStream_t; NULL for the
 * null stream) without synchronising.  Flag_IncludePecVelsFor21cm must be 0 when
 * Flag_IncludeSpinTemp is set (the velocity terms need a context's FFT workspace;
 * MXR_ERR_UNSUPPORTED); ReionGridDim even and at most 2048.  It calls no allocating routine: its
 * scratch (8192 FP64 partial sums and the total, 64 KiB) is a fixed array in device memory, one
 * per GPU, held until the process ends and shared by every call on that GPU, so calls that
 * overlap on one GPU race.  Results equal mxr_brightness_temperature's bit for bit.          */
int mxr_brightness_temperature_noalloc(const mxr_tb_params* params, const mxr_tb_grids* grids,
                                       double* volume_ave_Tb, void* stream);

/* ---- 21-cm power spectrum -------------------------------------------------------------------- */
/* Number of k bins: ceilings delta_k 1.35^i below delta_k ReionGridDim, delta_k = (float)(2 pi
 * / (BoxSize / Hubble_h)).  Replaces Initialise_PowerSpectrum's PS_Length
 * (src/core/ComputePowerSpectrum.c:18-41).                                              */
int mxr_ps_length(int ReionGridDim, double BoxSize, double Hubble_h);

/* Dimensional power spectrum (mK^2) of a real grid delta_T (host or device, `mem`): the FP64
 * mean <dT>, dd = (dT/<dT> - 1) V / N^3 <dT> on the padded grid, an unnormalised r2c, then
 * per bin the averages of k and of k^3 |dd(k)|^2 / (2 pi^2 V) over the half spectrum, and
 * PS_error = PS_data / sqrt(count).  PS_Length must equal mxr_ps_length(); the three outputs
 * are host arrays of PS_Length floats.  Works for any field (delta_TII too).
 * Replaces Compute_PS (src/core/ComputePowerSpectrum.c:44-255, declared
 * src/core/ComputePowerSpectrum.h:10).                                                 */
int mxr_power_spectrum(mxr_ctx* ctx, const float* delta_T, int mem, double BoxSize, double Hubble_h,
                       int PS_Length, float* PS_k, float* PS_data, float* PS_error);

/* The per-bin mode counts behind PS_error (in_bin_ct, src/core/ComputePowerSpectrum.c:140,
 * 176), PS_Length doubles into `counts`.  The reference keeps them internal; they are exposed
 * for callers that combine spectra of several boxes and to verify the binning.              */
int mxr_ps_counts(mxr_ctx* ctx, double BoxSize, double Hubble_h, int PS_Length, double* counts);

/* ---- input density and velocity grids (read_grids*.c) --------------------------------------- */
/* enum grid_prop (src/core/read_grids.h) */
#define MXR_GRID_DENSITY 0
#define MXR_GRID_X_VELOCITY 1
#define MXR_GRID_Y_VELOCITY 2
#define MXR_GRID_Z_VELOCITY 3

/* Header of a gbpTrees grid file (read_header, src/core/read_grids-gbptrees.c:38-86): three
 * ints n_cell, three doubles box_size (cMpc/h), ints n_grids (4) and ma_scheme, then per grid a
 * 32-byte identifier and n_cell^3 floats (x slowest).                                      */
typedef struct mxr_grid_file_info {
  int n_cell[3];
  double box_size[3];
  int n_grids;
  int ma_scheme;
} mxr_grid_file_info;

int mxr_grid_file_read_info(const char* fname, mxr_grid_file_info* info);

typedef struct mxr_grid_params {
  double BoxSize;                  /* run_globals.params.BoxSize (Mpc/h): the smoothing scale  */
  double Hubble_h;
  long long NPart;                 /* density only                                            */
  double PartMass;                 /* density only                                            */
  int TreesID;                     /* MXR_TREES_*: selects the density normalisation and floor */
} mxr_grid_params;

/* One rank's read_grid post-processing on a grid already in memory: hi is the file's grid at
 * n_hi^3 in the real layout (host or device, mem_in); n_hi a multiple of the context's
 * dimension N and, when larger, itself a supported dimension.  Larger grids are smoothed
 * (r2c, /n_hi^3, real-space top-hat of radius BoxSize/N/2, c2r) and every (n_hi/N)-th point
 * kept; then, for MXR_GRID_DENSITY, the overdensity max(rho mean_inv - 1, floor) with the
 * file's box size file_box.  slab: the padded real N^3 grid (mem_out); its padding entries
 * are zero.  For callers with their own reader (the VELOCIraptor HDF5 grids).  Smoothing keeps
 * a context of dimension n_hi and its window table ((n_hi/2+1)^3 floats, rebuilt when the box
 * size changes) with ctx until mxr_destroy.
 * Replaces smooth_grid and subsample_grid (src/core/read_grids.c:128-287) and the density
 * step of read_grid__gbptrees (read_grids-gbptrees.c:201-218) and read_grid__velociraptor
 * (read_grids-velociraptor.c:213-230).                                                    */
int mxr_prepare_grid(mxr_ctx* ctx, const float* hi, int n_hi, int mem_in, int property, double file_box,
                     const mxr_grid_params* params, float* slab, int mem_out);

/* read_grid__gbptrees (src/core/read_grids-gbptrees.c:88-231) for one rank: reads grid
 * `property` of the file and prepares it as mxr_prepare_grid does (TreesID is taken as
 * MXR_TREES_GBPTREES).  MXR_ERR_INVALID_ARG for an unreadable file, n_grids != 4, a
 * non-cubic grid or one not a multiple of N (the reference aborts).                       */
int mxr_read_grid_gbptrees(mxr_ctx* ctx, const char* fname, int property, const mxr_grid_params* params, float* slab,
                           int mem_out);

/* ---- galaxies to grids (construct_baryon_grids) -------------------------------------------- */
/* Flag_IncludeAGNXray (run_params_t.physics): which AGN X-ray emissivity grids are built     */
#define MXR_AGN_XRAY_NONE 0
#define MXR_AGN_XRAY_BOTH 1
#define MXR_AGN_XRAY_HARD 2
#define MXR_AGN_XRAY_SOFT 3

typedef struct mxr_baryon_params {
  int ReionGridDim;
  double BoxSize;                  /* Mpc/h, the side pos_to_ngp divides positions by       */
  int Flag_IncludeSpinTemp;        /* sfr, X-ray and AGN grids and their histories            */
  int Flag_BHFeedback;             /* effective_bhm and effective_bhar grids                  */
  int Flag_InstantaneousSFR;       /* sfr grids from Sfr (1), or GrossStellarMass / t_sfr (0)  */
  int Flag_IncludeAGNXray;         /* MXR_AGN_XRAY_*                                          */
  double BlackHoleMassLimitReion;  /* BH grids count galaxies with BlackHoleMass >= this      */
  double sfr_timescale;            /* ReionSfrTimescale * hubble_time(snapshot), internal     */
  int NstoreSnapshots_Heating;     /* padded grids in each *_histories array, slot 0 newest   */
  /* galaxy X-ray luminosity (calculate_galaxy_xray_luminosity, reionization.c:1982-2023)    */
  int Flag_XrayMetallicity;
  double LXrayGal;
  double XrayLogLNorm;
  double XrayMetallicitySlope;
  double XrayMetallicitySolar;
  double XrayOxygenAbundanceFactor;
  double NuXrayThreshold;
  double NuXraySoftCut;
  double SpecIndexXrayGal;
  double UnitMass_in_g;            /* run_globals.units                                        */
  double UnitTime_in_s;
} mxr_baryon_params;

/* Galaxy properties as structure of arrays (galaxy_t fields, src/meraxes.h), n entries each,
 * in the order the reference accumulates them (its galaxy_to_slab_map order).  Arrays a
 * call's flags do not read may be NULL.                                                     */
typedef struct mxr_galaxies {
  int mem;                         /* MXR_MEM_HOST or MXR_MEM_DEVICE                          */
  long n;
  const float* Pos;                /* [n][3], Mpc/h                                            */
  const int* Type;                 /* > 2: dead, skipped                                       */
  const double* FescWeightedGSM;
  const double* FescWeightedSfr;
  const double* Sfr;
  const double* GrossStellarMass;
  const double* ColdGas;           /* Flag_XrayMetallicity                                     */
  const double* MetalsColdGas;
  const double* BlackHoleMass;     /* Flag_BHFeedback or AGN X-ray grids                      */
  const double* EffectiveBHM;
  const double* EffectiveBHAR;
  const double* BHXrayEmissivity_hard;
  const double* BHXrayEmissivity_soft;
} mxr_galaxies;

/* Output grids in the padded real layout (reion_grids_t names); a *_histories array holds
 * NstoreSnapshots_Heating padded grids.  Grids a call's flags do not build may be NULL.    */
typedef struct mxr_baryon_grids {
  int mem;
  float* stars;
  float* weighted_sfr;
  float* effective_bhm;
  float* effective_bhar;
  float* sfr;
  float* sfr_histories;
  float* xray_luminosity;
  float* xray_luminosity_histories;
  float* BHXrayEmissivity_hard;
  float* bh_xray_histories_hard;
  float* BHXrayEmissivity_soft;
  float* bh_xray_histories_soft;
  int* history_head;    /* NULL: the reference's layout, each history shifted by one slot (a
                           move of NstoreSnapshots_Heating - 1 grids per history).  Otherwise the
                           histories are rings (history s in slot (*history_head + s) % n, as
                           mxr_ts_grids.history_head reads them): the new grid goes into slot
                           (*history_head + n - 1) % n, which becomes *history_head, and nothing
                           moves.  The ring's padding entries are unspecified.               */
} mxr_baryon_grids;

/* Nearest-grid-point deposit of the galaxies into the grids, for one rank holding the whole
 * grid, without stochasticity or mini-haloes: every cell sums its galaxies in the given order
 * with the reference's float/double arithmetic, negative sums clamp to zero, sfr becomes a
 * rate, and each history array shifts by one slot before slot 0 takes the new grid.
 * *n_below_bh_limit (may be NULL) gets the live galaxies under BlackHoleMassLimitReion
 * (Flag_BHFeedback).  A position outside [0, BoxSize) that pos_to_ngp would map below 0 is
 * MXR_ERR_INVALID_ARG (the reference asserts).
 * Replaces construct_baryon_grids (src/core/reionization.c:2025-2591, declared
 * reionization.h:73), with pos_to_ngp (misc_tools.c:207-217).                            */
int mxr_construct_baryon_grids(mxr_ctx* ctx, const mxr_baryon_params* params, const mxr_galaxies* galaxies,
                               const mxr_baryon_grids* grids, long* n_below_bh_limit);

/* ---- spin temperature ----------------------------------------------------------------------- */
#define MXR_TS_MAX_STEPS 64  /* TsNumFilterSteps upper bound                                  */
#define MXR_TS_MAX_OPS 16    /* source-grid operations before one shell's forward FFT          */
#define MXR_X_INT_NXHII 14   /* x_e bins of the secondary-electron tables                     */
/* source-grid operations (ComputeTs.c:366-478, load_reion_sfr_grids reionization.c:2750-2797) */
#define MXR_TS_OP_CLEAR 0    /* grid = 0                                                       */
#define MXR_TS_OP_SCALE 1    /* grid = (float)((double)grid * weight)                          */
#define MXR_TS_OP_LOAD 2     /* grid = history[hist] * (float)weight                           */
#define MXR_TS_OP_ADD 3      /* grid += history[hist] * (float)weight                          */

typedef struct mxr_ts_params {
  int ReionGridDim;
  double BoxSize;                 /* Mpc/h                                                 */
  double Hubble_h;
  double OmegaM;
  double OmegaLambda;
  double OmegaK;
  double OmegaR;
  double wLambda;
  double BaryonFrac;
  double Y_He;
  int ReionRtoMFilterType;        /* RtoM of the R_ct = 0 collapse fraction                */
  int TsNumFilterSteps;           /* shells (<= MXR_TS_MAX_STEPS)                          */
  int TsHeatingFilterType;        /* window of the shells (reionization.c:3373-3433)       */
  double ReionMaxHeatingRedshift; /* homogeneous branch at z >= this                       */
  double ReionEfficiency;
  double ReionNionPhotPerBary;    /* Pop II Lyman-band photons per baryon                  */
  double NuXrayThreshold;         /* eV                                                    */
  double NuXraySoftCut;           /* eV                                                    */
  double NuXrayMax;               /* eV                                                    */
  double SpecIndexXrayGal;
  int Flag_IncludeAGNXray;        /* MXR_AGN_XRAY_*: AGN X-ray bands heating and ionising  */
  double SpecIndexXrayAGNSoft;    /* soft band NuXrayThreshold..NuXraySoftCut               */
  double SpecIndexXrayAGNHard;    /* hard band NuXraySoftCut..NuXrayMax                     */
  double UnitLength_in_cm;
  double UnitMass_in_g;
  double UnitVelocity_in_cm_per_s;
  int snapshot;                   /* index of the current snapshot in ZZ                   */
  const double* ZZ;               /* host array: redshift of snapshots 0..snapshot         */
  const char* TablesForXHeatingDir; /* recfast_LCDM.dat, kappa_*_table.dat,
                                       stellar_spectra.dat, x_int_tables/                  */
} mxr_ts_params;

typedef struct mxr_ts_grids {
  int mem;                        /* MXR_MEM_HOST or MXR_MEM_DEVICE (grids below)          */
  const float* deltax;            /* in  padded real                                        */
  const float* stars;             /* in  padded real (collapse fraction of the shell R_0)   */
  const float* sfr_histories;     /* in  n_histories padded grids, [s] = s snapshots back;
                                     [0] is the current sfr grid (reionization.c:2106-2141) */
  const float* xray_luminosity_histories; /* in, same layout (1e40 erg/s units)             */
  const float* bh_xray_histories_hard; /* in, same layout (1e10 L_sun per cell): read when
                                     Flag_IncludeAGNXray is BOTH or HARD, else may be NULL  */
  const float* bh_xray_histories_soft; /* in, same layout: read when BOTH or SOFT            */
  int n_histories;
  float* x_e_box_prev;            /* in/out padded real                                     */
  float* x_e_box;                 /* out padded real (x_e_box_prev after the step)          */
  float* Tk_box;                  /* in/out real                                            */
  float* TS_box;                  /* out real                                               */
  double* stored_fcoll;           /* host, in/out: collapse fraction per snapshot (the entry
                                     [snapshot] is written, [snapshot - 1] read)            */
  int history_head;               /* slot of the newest history grid: history s lives in slot
                                     (history_head + s) % n_histories.  0 is the reference's
                                     layout; a ring kept by mxr_construct_baryon_grids moves it */
} mxr_ts_grids;

/* Grid-independent per-shell tables and the history plan ComputeTs builds on the host
 * (ComputeTs.c:339-478, 749-1132).  Filled by mxr_compute_ts when asked.                  */
typedef struct mxr_ts_tables {
  int n;                          /* shells                                                */
  int homogeneous;                /* 1: the ReionMaxHeatingRedshift branch ran             */
  int NO_LIGHT;
  double zp, dzp, dt_dzp, growth_factor_zp, dgrowth_factor_dzp, const_zp_prefactor_GAL;
  double collapse_fraction, x_e_ave, filling_factor_of_HI_zp;
  double R[MXR_TS_MAX_STEPS];         /* Mpc                                               */
  double zpp_edge[MXR_TS_MAX_STEPS];
  double zpp[MXR_TS_MAX_STEPS];
  double nu_tau_one[MXR_TS_MAX_STEPS];
  double sum_lyn[MXR_TS_MAX_STEPS];
  double freq_int_heat[MXR_X_INT_NXHII][MXR_TS_MAX_STEPS];
  double freq_int_ion[MXR_X_INT_NXHII][MXR_TS_MAX_STEPS];
  double freq_int_lya[MXR_X_INT_NXHII][MXR_TS_MAX_STEPS];
  /* AGN bands (ComputeTs.c:835-899, 1045-1107): the same integrals over each band's limits
   * with its spectral index (zero for a band the flag leaves out), and the band prefactors
   * (computed whatever the flag, as the reference does) */
  double const_zp_prefactor_AGN_soft, const_zp_prefactor_AGN_hard;
  double freq_int_heat_AGN_soft[MXR_X_INT_NXHII][MXR_TS_MAX_STEPS];
  double freq_int_ion_AGN_soft[MXR_X_INT_NXHII][MXR_TS_MAX_STEPS];
  double freq_int_lya_AGN_soft[MXR_X_INT_NXHII][MXR_TS_MAX_STEPS];
  double freq_int_heat_AGN_hard[MXR_X_INT_NXHII][MXR_TS_MAX_STEPS];
  double freq_int_ion_AGN_hard[MXR_X_INT_NXHII][MXR_TS_MAX_STEPS];
  double freq_int_lya_AGN_hard[MXR_X_INT_NXHII][MXR_TS_MAX_STEPS];
  float x_int_XHII[MXR_X_INT_NXHII];
  int n_ops[MXR_TS_MAX_STEPS];        /* source operations before each shell's FFT         */
  int op_kind[MXR_TS_MAX_STEPS][MXR_TS_MAX_OPS];
  int op_hist[MXR_TS_MAX_STEPS][MXR_TS_MAX_OPS];
  double op_weight[MXR_TS_MAX_STEPS][MXR_TS_MAX_OPS];
  int hist_max;                   /* largest history index read (-1: none)                 */
} mxr_ts_tables;

typedef struct mxr_ts_globals {
  double volume_ave_J_alpha;      /* the four below are not written in the homogeneous     */
  double volume_ave_xalpha;       /* branch (the reference leaves them unchanged there)    */
  double volume_ave_Xheat;
  double volume_ave_Xion;
  double volume_ave_TS;
  double volume_ave_TK;
  double volume_ave_xe;
  double XrayEmissivity_HMXB;     /* volume-averaged galaxy X-ray emissivity density of the
                                     unfiltered shell (stored_XrayEmissivity_HMXB[snapshot]);
                                     not written in the homogeneous branch               */
  double XrayEmissivity_hard;     /* the same for the AGN bands (stored_XrayEmissivity_hard /
                                     _soft[snapshot], erg/s/cm^3): 0 for a band the flag   */
  double XrayEmissivity_soft;     /* leaves out; not written in the homogeneous branch     */
  double volume_ave_Xheat_AGN_soft; /* AGN heating terms of dT_K/dz (reion_grids_t); 0 for */
  double volume_ave_Xheat_AGN_hard; /* a band left out; not written when homogeneous      */
  double device_ms;
} mxr_ts_globals;

/* One spin-temperature step: the homogeneous RECFAST branch above ReionMaxHeatingRedshift,
 * otherwise TsNumFilterSteps shells of history-weighted SFR, X-ray luminosity and (with
 * Flag_IncludeAGNXray) AGN hard / soft X-ray emissivity grids, each r2c -> window -> c2r on
 * the GPU and kept in HBM, the host's per-shell frequency tables, and the per-cell
 * forward-Euler update of x_e and T_K with the spin temperature.  The AGN grids follow the
 * history plan's loads and accumulations only: clearing and rescaling the SFR grids leaves
 * them as they are (load_reion_bh_grids, reionization.c:2799-2870).  `tables` may be NULL;
 * otherwise it receives the tables used.  Replaces _ComputeTs (src/core/ComputeTs.c:85-1554;
 * entry ComputeTs, src/core/ComputeTs.h:13) without mini-haloes or Lyman-Werner feedback.
 * A non-finite AGN prefactor is MXR_ERR_INVALID_ARG whatever the flag (the reference aborts). */
int mxr_compute_ts(mxr_ctx* ctx, const mxr_ts_params* params, const mxr_ts_grids* grids,
                   mxr_ts_globals* globals, mxr_ts_tables* tables);

/* The host-side, grid-independent part of one spin-temperature step on its own: the shell
 * geometry and history plan (ComputeTs.c:339-478) and, given the R_ct = 0 averages the grid
 * pass produces (collapse_fraction, x_e_ave) and stored_fcoll (entry [snapshot] must already
 * hold collapse_fraction), the frequency tables, sum_lyn and prefactors (:749-1132).  Needs
 * no GPU; mxr_compute_ts runs the same code.                                              */
int mxr_ts_build_tables(const mxr_ts_params* params, double collapse_fraction, double x_e_ave,
                        const double* stored_fcoll, mxr_ts_tables* tables);

/* Number of history snapshots the shells of snapshots 1..last_snapshot reach back to
 * (params->ZZ must hold last_snapshot + 1 redshifts).  Replaces set_sfr_history
 * (src/core/ComputeTs.c:32-83), which sizes NstoreSnapshots_Heating.                       */
int mxr_ts_history_depth(const mxr_ts_params* params, int last_snapshot);

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
