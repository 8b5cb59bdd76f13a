/*
 * mxr_meraxes.c -- the reference-side binding of libmxr.so: the file a Meraxes maintainer adds
 * to src/core/ to run the reionization grid path on one MI355X per rank.
 *
 * Each function below has the signature of the reference function it replaces:
 *   find_HII_bubbles_mxr              find_HII_bubbles   src/core/find_HII_bubbles.h:12
 *   RtoM_mxr                          RtoM               src/core/find_HII_bubbles.h:11
 *   ComputeTs_mxr                     ComputeTs          src/core/ComputeTs.h:13
 *   ComputeBrightnessTemperatureBox_mxr
 *                                     ComputeBrightnessTemperatureBox
 *                                                        src/core/BrightnessTemperature.h:9
 *   filter_mxr                        filter             src/core/reionization.h:87
 *   velocity_gradient_mxr             velocity_gradient  src/core/reionization.h:88
 * Built with -DMXR_REFERENCE_NAMES the file also defines the reference's own names as calls of
 * these, for a build that leaves the reference's definitions out.
 *
 * Parameters come from run_globals field by field (the library's structs use the same names),
 * the grids are handed over in their FFTW layouts (host memory), and a failure becomes the
 * reference's ABORT.  C99.
 */
#include <stdlib.h>

#include "XRayHeatingFunctions.h"
#include "meraxes.h"
#include "mxr.h"
#include "reionization.h"
#include "utils.h"

/* one context per grid dimension in use (the reionization grid, and the dimensions filter()
 * is called with, e.g. by the input-grid smoothing), created on first use, freed at exit */
#define MXR_SHIM_MAX_CTX 4
static mxr_ctx* mxr_ctx_of[MXR_SHIM_MAX_CTX];
static int mxr_dim_of[MXR_SHIM_MAX_CTX];

static void mxr_check(int rc, const char* where)
{
  if (rc != MXR_OK) {
    mlog_error("%s: %s (%s)", where, mxr_strerror(rc), mxr_last_error());
    ABORT(EXIT_FAILURE);
  }
}

static void mxr_release_all(void)
{
  for (int i = 0; i < MXR_SHIM_MAX_CTX; i++)
    if (mxr_ctx_of[i]) {
      mxr_destroy(mxr_ctx_of[i]);
      mxr_ctx_of[i] = NULL;
    }
}

static mxr_ctx* mxr_context(int grid_dim)
{
  static int registered = 0;
  for (int i = 0; i < MXR_SHIM_MAX_CTX; i++)
    if (mxr_ctx_of[i] && mxr_dim_of[i] == grid_dim)
      return mxr_ctx_of[i];
  if (run_globals.mpi_size != 1) {
    mlog_error("the MI355X path runs the whole grid on one rank");
    ABORT(EXIT_FAILURE);
  }
  for (int i = 0; i < MXR_SHIM_MAX_CTX; i++)
    if (!mxr_ctx_of[i]) {
      mxr_check(mxr_create(&mxr_ctx_of[i], 0, grid_dim), "mxr_create");
      mxr_dim_of[i] = grid_dim;
      if (!registered) {
        atexit(mxr_release_all);
        registered = 1;
      }
      return mxr_ctx_of[i];
    }
  mlog_error("more than %d grid dimensions in use", MXR_SHIM_MAX_CTX);
  ABORT(EXIT_FAILURE);
  return NULL;
}

static void mxr_fill_hii_params(mxr_hii_params* p, int snapshot)
{
  const run_params_t* rp = &run_globals.params;
  const physics_params_t* ph = &run_globals.params.physics;
  p->ReionGridDim = rp->ReionGridDim;
  p->BoxSize = rp->BoxSize;
  p->Hubble_h = rp->Hubble_h;
  p->OmegaM = rp->OmegaM;
  p->OmegaLambda = rp->OmegaLambda;
  p->OmegaK = rp->OmegaK;
  p->BaryonFrac = rp->BaryonFrac;
  p->ReionDeltaRFactor = rp->ReionDeltaRFactor;
  p->ReionFilterType = rp->ReionFilterType;
  p->ReionRtoMFilterType = rp->ReionRtoMFilterType;
  p->ReionUVBFlag = rp->ReionUVBFlag;
  p->Flag_IncludeSpinTemp = rp->Flag_IncludeSpinTemp;
  p->Flag_IncludeRecombinations = rp->Flag_IncludeRecombinations;
  p->Flag_EvolvingReionRBubbleMax = rp->Flag_EvolvingReionRBubbleMax;
  p->Flag_BHFeedback = ph->Flag_BHFeedback;
  p->ReionSfrTimescale = rp->ReionSfrTimescale;
  p->ReionEfficiency = ph->ReionEfficiency;
  p->ReionNionPhotPerBary = ph->ReionNionPhotPerBary;
  p->ReionGammaHaloBias = ph->ReionGammaHaloBias;
  p->ReionAlphaUV = ph->ReionAlphaUV;
  p->ReionAlphaUVBH = ph->ReionAlphaUVBH;
  p->ReionRBubbleMax = ph->ReionRBubbleMax;
  p->ReionRBubbleMaxRecomb = ph->ReionRBubbleMaxRecomb;
  p->ReionRBubbleMin = ph->ReionRBubbleMin;
  p->UnitLength_in_cm = run_globals.units.UnitLength_in_cm;
  p->UnitMass_in_g = run_globals.units.UnitMass_in_g;
  p->UnitVelocity_in_cm_per_s = run_globals.units.UnitVelocity_in_cm_per_s;
  p->redshift = run_globals.ZZ[snapshot];
  p->prev_redshift = run_globals.ZZ[snapshot > 0 ? snapshot - 1 : 0];
  p->TK_recfast = -1.0; /* the library evaluates T_RECFAST from the tables directory */
  p->TablesForXHeatingDir = rp->TablesForXHeatingDir;
  p->Flag_TemperatureDependentRec = rp->Flag_TemperatureDependentRec;
  p->Y_He = ph->Y_He;
}

/* _find_HII_bubbles (src/core/find_HII_bubbles.c:86-710) on the GPU, timed as the reference's
 * dispatcher find_HII_bubbles (:713-747) times it */
void find_HII_bubbles_mxr(int snapshot, timer_info* timer_total)
{
  reion_grids_t* g = &run_globals.reion_grids;
  const int spin = run_globals.params.Flag_IncludeSpinTemp;
  const int bh = run_globals.params.physics.Flag_BHFeedback;
  const int rec = run_globals.params.Flag_IncludeRecombinations;
  timer_info timer;
  timer_start(&timer);
  mxr_hii_params p;
  mxr_fill_hii_params(&p, snapshot);
  mxr_hii_grids io;
  io.mem = MXR_MEM_HOST;
  io.deltax = g->deltax;
  io.stars = g->stars;
  io.weighted_sfr = g->weighted_sfr;
  io.x_e_box = spin ? g->x_e_box : NULL;
  io.Tk_box = spin ? g->Tk_box : NULL;
  io.xH = g->xH;
  io.z_at_ionization = g->z_at_ionization;
  io.J_21_at_ionization = g->J_21_at_ionization;
  io.r_bubble = g->r_bubble;
  io.temp_kinetic_all_gas = g->temp_kinetic_all_gas;
  io.storage = MXR_STORAGE_F32;
  io.effective_bhm = bh ? g->effective_bhm : NULL;
  io.effective_bhar = bh ? g->effective_bhar : NULL;
  io.N_rec = rec ? g->N_rec : NULL;
  io.Gamma12 = rec ? g->Gamma12 : NULL;
  io.residual_xH = rec ? g->residual_xH : NULL;
  io.clumping_factor = rec ? g->clumping_factor : NULL;
  io.t_resp = rec ? g->t_resp : NULL;
  mxr_hii_globals out;
  mxr_check(mxr_find_hii_bubbles(mxr_context(p.ReionGridDim), &p, &io, &out), "mxr_find_hii_bubbles");
  g->volume_weighted_global_xH = out.volume_weighted_global_xH;
  g->mass_weighted_global_xH = out.mass_weighted_global_xH;
  g->volume_weighted_global_r_bubble = out.volume_weighted_global_r_bubble;
  g->mass_weighted_global_r_bubble = out.mass_weighted_global_r_bubble;
  g->volume_weighted_global_weighted_sfr = out.volume_weighted_global_weighted_sfr;
  g->volume_weighted_global_effective_bhar = out.volume_weighted_global_effective_bhar;
  g->volume_weighted_global_temp_kinetic_all_gas = out.volume_weighted_global_temp_kinetic_all_gas;
  g->mass_weighted_global_temp_kinetic_all_gas = out.mass_weighted_global_temp_kinetic_all_gas;
  g->volume_weighted_global_Gamma12 = out.volume_weighted_global_Gamma12;
  g->mass_weighted_global_Gamma12 = out.mass_weighted_global_Gamma12;
  g->volume_weighted_global_N_rec = out.volume_weighted_global_N_rec;
  g->volume_weighted_global_residual_xH = out.volume_weighted_global_residual_xH;
  g->volume_weighted_global_clumping_factor = out.volume_weighted_global_clumping_factor;
  g->mass_weighted_global_N_rec = out.mass_weighted_global_N_rec;
  g->mass_weighted_global_residual_xH = out.mass_weighted_global_residual_xH;
  g->mass_weighted_global_clumping_factor = out.mass_weighted_global_clumping_factor;
  timer_stop(&timer);
  timer_stop(timer_total);
  timer_gpu += timer_delta(timer);
}

/* RtoM (src/core/find_HII_bubbles.c:27-46) from run_globals' cosmology and units */
double RtoM_mxr(double R)
{
  mxr_hii_params p;
  mxr_fill_hii_params(&p, 0);
  return mxr_RtoM(&p, R);
}

/* ComputeBrightnessTemperatureBox (src/core/BrightnessTemperature.c:24-549), no mini-haloes */
void ComputeBrightnessTemperatureBox_mxr(int snapshot)
{
  reion_grids_t* g = &run_globals.reion_grids;
  const run_params_t* rp = &run_globals.params;
  mxr_tb_params p;
  p.ReionGridDim = rp->ReionGridDim;
  p.BoxSize = rp->BoxSize;
  p.Hubble_h = rp->Hubble_h;
  p.OmegaM = rp->OmegaM;
  p.OmegaLambda = rp->OmegaLambda;
  p.OmegaR = rp->OmegaR;
  p.BaryonFrac = rp->BaryonFrac;
  p.redshift = run_globals.ZZ[snapshot];
  p.Flag_IncludeSpinTemp = rp->Flag_IncludeSpinTemp;
  p.Flag_IncludePecVelsFor21cm = rp->Flag_IncludePecVelsFor21cm;
  p.TreesID = (int)rp->TreesID; /* enum tree_ids 0, 1, 2 = MXR_TREES_* */
  mxr_tb_grids io;
  io.mem = MXR_MEM_HOST;
  io.xH = g->xH;
  io.deltax = g->deltax;
  io.TS_box = g->TS_box;
  io.vel = g->vel;
  io.delta_T = g->delta_T;
  mxr_check(mxr_brightness_temperature(mxr_context(p.ReionGridDim), &p, &io, &g->volume_ave_Tb),
            "mxr_brightness_temperature");
}

/* _ComputeTs (src/core/ComputeTs.c:85-1554) without mini-haloes, AGN X-ray bands as
 * Flag_IncludeAGNXray asks, timed as the reference's dispatcher ComputeTs (:1558-1579) times it */
void ComputeTs_mxr(int snapshot, timer_info* timer_total)
{
  reion_grids_t* g = &run_globals.reion_grids;
  const run_params_t* rp = &run_globals.params;
  const physics_params_t* ph = &run_globals.params.physics;
  timer_info timer;
  timer_start(&timer);
  mxr_ts_params p;
  p.ReionGridDim = rp->ReionGridDim;
  p.BoxSize = rp->BoxSize;
  p.Hubble_h = rp->Hubble_h;
  p.OmegaM = rp->OmegaM;
  p.OmegaLambda = rp->OmegaLambda;
  p.OmegaK = rp->OmegaK;
  p.OmegaR = rp->OmegaR;
  p.wLambda = rp->wLambda;
  p.BaryonFrac = rp->BaryonFrac;
  p.Y_He = ph->Y_He;
  p.ReionRtoMFilterType = rp->ReionRtoMFilterType;
  p.TsNumFilterSteps = rp->TsNumFilterSteps;
  p.TsHeatingFilterType = rp->TsHeatingFilterType;
  p.ReionMaxHeatingRedshift = ph->ReionMaxHeatingRedshift;
  p.ReionEfficiency = ph->ReionEfficiency;
  p.ReionNionPhotPerBary = ph->ReionNionPhotPerBary;
  p.NuXrayThreshold = ph->NuXrayThreshold;
  p.NuXraySoftCut = ph->NuXraySoftCut;
  p.NuXrayMax = ph->NuXrayMax;
  p.SpecIndexXrayGal = ph->SpecIndexXrayGal;
  p.Flag_IncludeAGNXray = ph->Flag_IncludeAGNXray;
  p.SpecIndexXrayAGNSoft = ph->SpecIndexXrayAGNSoft;
  p.SpecIndexXrayAGNHard = ph->SpecIndexXrayAGNHard;
  p.UnitLength_in_cm = run_globals.units.UnitLength_in_cm;
  p.UnitMass_in_g = run_globals.units.UnitMass_in_g;
  p.UnitVelocity_in_cm_per_s = run_globals.units.UnitVelocity_in_cm_per_s;
  p.snapshot = snapshot;
  p.ZZ = run_globals.ZZ;
  p.TablesForXHeatingDir = rp->TablesForXHeatingDir;
  mxr_ts_grids io;
  io.mem = MXR_MEM_HOST;
  io.deltax = g->deltax;
  io.stars = g->stars;
  io.sfr_histories = g->sfr_histories;
  io.xray_luminosity_histories = g->xray_luminosity_histories;
  io.bh_xray_histories_hard = g->bh_xray_histories_hard; /* NULL unless the flag allocated them */
  io.bh_xray_histories_soft = g->bh_xray_histories_soft;
  io.n_histories = run_globals.NstoreSnapshots_Heating;
  io.x_e_box_prev = g->x_e_box_prev;
  io.x_e_box = g->x_e_box;
  io.Tk_box = g->Tk_box;
  io.TS_box = g->TS_box;
  io.stored_fcoll = stored_fcoll; /* one entry per snapshot */
  io.history_head = 0;            /* the reference's shifted history layout */
  /* the homogeneous branch leaves these four and the emissivity as they are */
  mxr_ts_globals out;
  out.volume_ave_J_alpha = g->volume_ave_J_alpha;
  out.volume_ave_xalpha = g->volume_ave_xalpha;
  out.volume_ave_Xheat = g->volume_ave_Xheat;
  out.volume_ave_Xion = g->volume_ave_Xion;
  out.volume_ave_TS = g->volume_ave_TS;
  out.volume_ave_TK = g->volume_ave_TK;
  out.volume_ave_xe = g->volume_ave_xe;
  out.XrayEmissivity_HMXB = stored_XrayEmissivity_HMXB[snapshot];
  out.XrayEmissivity_hard = stored_XrayEmissivity_hard[snapshot];
  out.XrayEmissivity_soft = stored_XrayEmissivity_soft[snapshot];
  out.volume_ave_Xheat_AGN_soft = g->volume_ave_Xheat_AGN_soft;
  out.volume_ave_Xheat_AGN_hard = g->volume_ave_Xheat_AGN_hard;
  out.device_ms = 0.0;
  mxr_check(mxr_compute_ts(mxr_context(p.ReionGridDim), &p, &io, &out, NULL), "mxr_compute_ts");
  g->volume_ave_J_alpha = out.volume_ave_J_alpha;
  g->volume_ave_xalpha = out.volume_ave_xalpha;
  g->volume_ave_Xheat = out.volume_ave_Xheat;
  g->volume_ave_Xion = out.volume_ave_Xion;
  g->volume_ave_TS = out.volume_ave_TS;
  g->volume_ave_TK = out.volume_ave_TK;
  g->volume_ave_xe = out.volume_ave_xe;
  stored_XrayEmissivity_HMXB[snapshot] = out.XrayEmissivity_HMXB;
  stored_XrayEmissivity_hard[snapshot] = out.XrayEmissivity_hard;
  stored_XrayEmissivity_soft[snapshot] = out.XrayEmissivity_soft;
  g->volume_ave_Xheat_AGN_soft = out.volume_ave_Xheat_AGN_soft;
  g->volume_ave_Xheat_AGN_hard = out.volume_ave_Xheat_AGN_hard;
  timer_stop(&timer);
  timer_stop(timer_total);
  timer_gpu += timer_delta(timer);
}

/* filter (src/core/reionization.c:3373-3433): the window on a host half spectrum of grid_dim
 * (slab_nx x-planes from local_ix_start), BoxSize from run_globals as the reference takes it */
void filter_mxr(fftwf_complex* box, int local_ix_start, int slab_nx, int grid_dim, float R, int filter_type)
{
  mxr_check(mxr_filter(mxr_context(grid_dim), box, MXR_MEM_HOST, local_ix_start, slab_nx, grid_dim,
                       run_globals.params.BoxSize, R, filter_type),
            "mxr_filter");
}

/* velocity_gradient (src/core/reionization.c:3435-3450) */
void velocity_gradient_mxr(fftwf_complex* box, int slab_nx, int grid_dim)
{
  mxr_check(mxr_velocity_gradient(mxr_context(grid_dim), box, MXR_MEM_HOST, slab_nx, grid_dim,
                                  run_globals.params.BoxSize),
            "mxr_velocity_gradient");
}

#ifdef MXR_REFERENCE_NAMES
void find_HII_bubbles(int snapshot, timer_info* timer_total) { find_HII_bubbles_mxr(snapshot, timer_total); }
double RtoM(double R) { return RtoM_mxr(R); }
void ComputeTs(int snapshot, timer_info* timer_total) { ComputeTs_mxr(snapshot, timer_total); }
void ComputeBrightnessTemperatureBox(int snapshot) { ComputeBrightnessTemperatureBox_mxr(snapshot); }
void filter(fftwf_complex* box, int local_ix_start, int slab_nx, int grid_dim, float R, int filter_type)
{
  filter_mxr(box, local_ix_start, slab_nx, grid_dim, R, filter_type);
}
void velocity_gradient(fftwf_complex* box, int slab_nx, int grid_dim) { velocity_gradient_mxr(box, slab_nx, grid_dim); }
#endif
