/* Minimal stand-in for the declarations of src/meraxes.h that mxr_meraxes.c reads, with the
 * reference's field names and types, so the binding's C can be type-checked and linked against
 * libmxr.so without the reference's build.  Test infrastructure only. */
#ifndef MXR_SHIM_ENV_MERAXES_H
#define MXR_SHIM_ENV_MERAXES_H

#include <stdlib.h>

#define STRLEN 256

typedef float fftwf_complex[2]; /* FFTW3's single-precision complex */

enum tree_ids { VELOCIRAPTOR_TREES, GBPTREES_TREES, VELOCIRAPTOR_TREES_AUG };

typedef struct physics_params_t {
  double ReionEfficiency;
  double ReionNionPhotPerBary;
  double Y_He;
  double NuXrayThreshold;
  double SpecIndexXrayGal;
  double NuXraySoftCut;
  double NuXrayMax;
  int Flag_IncludeAGNXray;
  double SpecIndexXrayAGNSoft;
  double SpecIndexXrayAGNHard;
  double ReionMaxHeatingRedshift;
  double ReionGammaHaloBias;
  double ReionAlphaUV;
  double ReionAlphaUVBH;
  double ReionRBubbleMin;
  double ReionRBubbleMax;
  double ReionRBubbleMaxRecomb;
  int Flag_BHFeedback;
} physics_params_t;

typedef struct run_params_t {
  char TablesForXHeatingDir[STRLEN];
  physics_params_t physics;
  double BoxSize;
  double Hubble_h;
  double BaryonFrac;
  double OmegaM;
  double OmegaK;
  double OmegaR;
  double OmegaLambda;
  double wLambda;
  double ReionDeltaRFactor;
  int ReionGridDim;
  int ReionFilterType;
  int TsHeatingFilterType;
  int ReionRtoMFilterType;
  int ReionUVBFlag;
  enum tree_ids TreesID;
  int Flag_IncludeSpinTemp;
  int Flag_IncludeRecombinations;
  int Flag_EvolvingReionRBubbleMax;
  int Flag_TemperatureDependentRec;
  int Flag_IncludePecVelsFor21cm;
  int TsNumFilterSteps;
  double ReionSfrTimescale;
} run_params_t;

typedef struct run_units_t {
  double UnitLength_in_cm;
  double UnitVelocity_in_cm_per_s;
  double UnitMass_in_g;
} run_units_t;

typedef struct reion_grids_t {
  float* stars;
  float* effective_bhm;
  float* effective_bhar;
  float* deltax;
  float* sfr_histories;
  float* weighted_sfr;
  float* xray_luminosity_histories;
  float* bh_xray_histories_hard;
  float* bh_xray_histories_soft;
  float* xH;
  float* z_at_ionization;
  float* J_21_at_ionization;
  float* r_bubble;
  float* x_e_box;
  float* x_e_box_prev;
  float* Tk_box;
  float* TS_box;
  float* temp_kinetic_all_gas;
  float* delta_T;
  float* vel;
  float* N_rec;
  float* Gamma12;
  float* residual_xH;
  float* clumping_factor;
  float* t_resp;
  double volume_weighted_global_xH;
  double volume_weighted_global_Gamma12;
  double volume_weighted_global_r_bubble;
  double volume_weighted_global_temp_kinetic_all_gas;
  double volume_weighted_global_weighted_sfr;
  double volume_weighted_global_effective_bhar;
  double mass_weighted_global_xH;
  double mass_weighted_global_Gamma12;
  double mass_weighted_global_r_bubble;
  double mass_weighted_global_temp_kinetic_all_gas;
  double volume_weighted_global_N_rec;
  double volume_weighted_global_residual_xH;
  double volume_weighted_global_clumping_factor;
  double mass_weighted_global_N_rec;
  double mass_weighted_global_residual_xH;
  double mass_weighted_global_clumping_factor;
  double volume_ave_J_alpha;
  double volume_ave_xalpha;
  double volume_ave_Xheat;
  double volume_ave_Xheat_AGN_soft;
  double volume_ave_Xheat_AGN_hard;
  double volume_ave_Xion;
  double volume_ave_TS;
  double volume_ave_TK;
  double volume_ave_xe;
  double volume_ave_Tb;
} reion_grids_t;

typedef struct run_globals_t {
  struct run_params_t params;
  reion_grids_t reion_grids;
  struct run_units_t units;
  int mpi_size;
  double* ZZ;
  int NstoreSnapshots_Heating;
} run_globals_t;

extern run_globals_t run_globals;

void myexit(int signum);
#define ABORT(sigterm) myexit(sigterm)
void mlog_error(const char* fmt, ...);

#endif
