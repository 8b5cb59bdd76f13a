// mxr_internal.hpp -- context, error plumbing and host-side physics shared by the library's
// translation units.  Not part of the public ABI (include/mxr.h is).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/mxr.h"

namespace mxr {

// Constants restated from src/meraxes.h:21-67 and src/core/reionization.h:10-47.
constexpr double GRAVITY = 6.672e-8;
constexpr double SOLAR_MASS = 1.989e33;
constexpr double PLANCK = 6.6262e-27;
constexpr double PROTONMASS = 1.6726e-24;
constexpr double HUBBLE = 3.2407789e-18;
constexpr double SEC_PER_YEAR = 3.155e7;
constexpr double SEC_PER_MEGAYEAR = 3.155e13;
constexpr double SIGMA_HI = 6.3e-18;  // HI ionisation cross section at 13.6 eV, cm^2
constexpr double TCMB = 2.728;
constexpr double MPC = 3.086e24;          // cm
constexpr double T_RE = 1e4;
constexpr float REL_TOL = 1e-5f;
constexpr float ABS_TOL = 1e-8f;
constexpr double L_FACTOR = 0.620350491;
constexpr int RECFAST_NPTS = 501;

void set_error(const char* fmt, ...);
int hip_fail(hipError_t e, const char* what, const char* file, int line);

#define MXR_HIP(call)                                                                                                \
  do {                                                                                                               \
    hipError_t e_ = (call);                                                                                          \
    if (e_ != hipSuccess)                                                                                            \
      return ::mxr::hip_fail(e_, #call, __FILE__, __LINE__);                                                         \
  } while (0)

// propagate a nonzero status
#define RC(x)                                                                                                        \
  do {                                                                                                               \
    int rc_ = (x);                                                                                                   \
    if (rc_)                                                                                                         \
      return rc_;                                                                                                    \
  } while (0)

// Kernel classes timed by the profiler.
enum KClass {
  K_FWD_R2C = 0,  // z-pass r2c of the inputs
  K_FWD_Y,        // forward y-pass
  K_FWD_X,        // forward x-pass (+ /N^3 when fused)
  K_INV_X,        // inverse x-pass (+ window when fused)
  K_INV_Y,        // inverse y-pass
  K_INV_Z,        // inverse z-pass c2r (+ clamp/criterion/state when fused)
  K_COPY,         // plain: unfiltered -> filtered copy
  K_FILTER,       // plain: filter()
  K_CLAMP,        // plain: sanity_check_aliasing
  K_CRIT,         // plain: criterion / state kernel
  K_SCALE,        // plain: /N^3, xH and r_bubble fills
  K_POST,         // post-loop temperatures and FP64 sums
  K_TS_SRC,       // spin temperature: history-weighted source grids
  K_TS_FWD,       // spin temperature: forward r2c + y + x (/N^3) of a changed source
  K_TS_INV_X,     // spin temperature: inverse x-pass (+ window)
  K_TS_INV_Y,     // spin temperature: inverse y-pass
  K_TS_INV_Z,     // spin temperature: inverse z-pass c2r storing the smoothed shell
  K_TS_RED,       // spin temperature: R_ct = 0 collapse fraction, <x_e>, <L_X>
  K_TS_EVOLVE,    // spin temperature: per-cell x_e / T_K / T_S update (or homogeneous branch)
  K_NCLASS
};

struct KStat {
  int launches = 0;
  double total_ms = 0.0;
  double alg_bytes = 0.0;
};

// One grid dimension's worth of persistent device state.
constexpr int MXR_PS_MAXB = 40;  // power-spectrum bins: 1 + log(N)/log(1.35), 21 at N = 512

constexpr int BG_STREAMS = 4;  // history arrays of galaxies to grids, one copy stream each

struct Ctx {
  // profiling: an event pair per timed launch, resolved after the call's final sync
  bool prof = false;
  std::vector<hipEvent_t> evpool;
  struct Rec {
    int cls;
    int ev;
    double bytes;
  };
  std::vector<Rec> recs;
  int evnext = 0;
  KStat kstat[K_NCLASS];

  int device = 0;
  int N = 0;
  int path = MXR_PATH_FUSED;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float2* tw = nullptr;            // tw[m] = exp(-2 pi i m / N), m < N (FP64-evaluated, FP32-stored)
  float2* spec[7] = {};            // unfiltered half spectra (deltax, stars, weighted_sfr,
                                   // [x_e], [effective_bhm, effective_bhar], [N_rec])
  float2* work[7] = {};            // filtered spectra / FFT work (same size)
  float* in_stage[7] = {};         // device copies of host inputs: 6 padded grids + Tk_box
  float* out_stage[10] = {};       // device copies of host state grids: xH, z_in, J21, r_bubble,
                                   // temp, and with recombinations N_rec (padded), Gamma12,
                                   // residual_xH, clumping_factor, t_resp
  // recombination tables (mxr_recomb.hip), built for the cosmology in rec_key
  double* rec_tab = nullptr;
  double* rec_coef = nullptr;
  double* rec_x = nullptr;
  double rec_key[4] = {};
  uint8_t* ion = nullptr;          // per-cell "ionised during this call" flags
  double* partials = nullptr;      // per-block FP64 partial sums
  size_t partials_cap = 0;
  double* sums = nullptr;          // final FP64 sums (device)
  double* sums_host = nullptr;     // pinned copy
  // spin temperature
  float* ts_src[4] = {};           // current sfr, xray_luminosity and AGN band source grids (padded real)
  float* ts_store = nullptr;       // smoothed shells: [grid][shell][cell] FP32, real layout
  size_t ts_store_cap = 0;         // floats
  float* ts_hist = nullptr;        // staged host histories (sfr then xray), padded grids
  size_t ts_hist_cap = 0;          // floats
  void* ts_dev = nullptr;          // device copy of the per-call table block
  void* ts_dev_host = nullptr;     // pinned host copy
  hipEvent_t ev_red = nullptr;     // R_ct = 0 averages ready
  // batched find_HII_bubbles: enqueue without timing events or a final synchronisation (so the
  // sequence can be captured into a graph); the globals follow from sums_host afterwards
  bool enqueue_only = false;
  // launch shapes (defaults, or the autotuner's choice)
  int zdirect_kmax = 16;           // z pass by direct summation when a radius keeps <= this many k_z
  int zdirect_rpb = 4;             // rows per wavefront of the direct z pass
  int live_tiles = 1;              // x/y pass grids over the tiles that can hold kept columns
  int col_tc_x = 16, col_tc_y = 16; // columns per workgroup of the per-radius x and y passes
  // power spectrum: per-block bin partials, and the data-independent per-bin mode counts and
  // |k| sums of the last (grid, delta_k) seen
  double* ps_part = nullptr;
  size_t ps_part_cap = 0;          // doubles
  float ps_geom_dk = 0;
  int ps_geom_nb = 0;
  double ps_geom[2 * MXR_PS_MAXB] = {};
  // a context of an input grid's dimension, for smoothing it (created on first use)
  mxr_ctx* hires = nullptr;
  // the filter window of types 0 and 2 by (|n_x|, |n_y|, n_z), for the last (dk, R, type) seen
  float* wtab = nullptr;
  float wtab_dk = 0, wtab_R = 0;
  int wtab_ftype = -1;
  // galaxies to grids: sort keys and values, the sort's storage, staged galaxy arrays; streams
  // (default priority) for the history shift's copies, one per history array, and the events
  // ordering them
  void* bg_ws = nullptr;
  size_t bg_ws_cap = 0;            // bytes
  hipStream_t bg_stream[BG_STREAMS] = {};
  hipEvent_t bg_ev_in = nullptr, bg_ev_shift[BG_STREAMS] = {};
  double fin_total_n_cells = 0, fin_sfr_conv = 0;
  int fin_n_radii = 0;
  int fin_bh = 0;
  int fin_rec = 0;
};

inline long n_complex(int N) { return (long)N * N * (N / 2 + 1); }
// Row pitch, in complex values, of the internal half spectra (spec, work): the N/2 + 1 modes
// of a row rounded up to a multiple of 16, i.e. 128 bytes, so that every row starts on a cache
// line and a column tile's 16 adjacent modes are one aligned line (with 257 modes at N = 512 the
// 2056-byte rows put most 128-byte segments across two lines).  The caller's grids keep the
// reference's layouts; only these buffers are padded.  Padding entries are never read.
constexpr int spec_pitch(int N) { return ((N / 2 + 1) + 15) / 16 * 16; }
inline long n_spec(int N) { return (long)N * N * spec_pitch(N); }
inline long n_real(int N) { return (long)N * N * N; }

// host physics (mxr_api.hip)
double rho_crit(const mxr_hii_params* p);
double hubble_time(const mxr_hii_params* p);
double T_RECFAST(const char* tables_dir, float z);
float cT_approx(float z);
int check_hii_params(const mxr_hii_params* p, int N);

// grid-dimension dispatch (mxr_hii.hip)
// grid dimensions the library instantiates (include/mxr.h lists them at mxr_supported_dims)
#define MXR_DIMS(X) X(16) X(32) X(64) X(96) X(128) X(160) X(256) X(512) X(1024)
bool supported_dim(int N);
int run_brightness_temperature_noalloc(const mxr_tb_params* p, const mxr_tb_grids* g, double* ave_dev,
                                       hipStream_t stream);
int run_find_hii_bubbles(Ctx* c, const mxr_hii_params* p, const mxr_hii_grids* g, mxr_hii_globals* out);
int run_filter(Ctx* c, float2* box, int local_ix_start, int slab_nx, int grid_dim, float dk, float R, int ftype);
// device table of the type 0 / 2 window at (dk, R) over (|n_x|, |n_y|, n_z) in 0..N/2, kept in
// the context and rebuilt (on the context's stream) when (dk, R, type) changes
int window_table(Ctx* c, float dk, float R, int ftype, const float** tab);
int run_velocity_gradient(Ctx* c, float2* box, int slab_nx, int grid_dim, float dk);
int run_brightness_temperature(Ctx* c, const mxr_tb_params* p, const mxr_tb_grids* g, double* ave);
int launch_final_sum(Ctx* c, const double* partials, int nblk, int nv, double* out);  // fixed-order FP64 sum
int launch_sum_f(Ctx* c, const float* x, long n, int nblk);  // FP64 block sums of x into c->partials
void hii_globals(const Ctx* c, mxr_hii_globals* out);  // after a synchronised enqueue_only call
int run_compute_ts(Ctx* c, const mxr_ts_params* p, const mxr_ts_grids* g, mxr_ts_globals* out,
                   mxr_ts_tables* tables);
int ensure_buffers(Ctx* c, int ngrid);

// power spectrum (mxr_ps.hip)
int ps_length(int N, double BoxSize, double Hubble_h);
int run_power_spectrum(Ctx* c, const float* dT, int mem, double BoxSize, double Hubble_h, float* PS_k,
                       float* PS_data, float* PS_error, int nb);
int run_ps_counts(Ctx* c, double BoxSize, double Hubble_h, int nb, double* counts);

// input grids (mxr_grid.hip)
int run_c2r_store(Ctx* c, float2* spec);  // in-place c2r of one half spectrum (mxr_hii.hip)
int read_grid_file_info(const char* fname, mxr_grid_file_info* info, long* data_offset_of_grid0);
int run_prepare_grid(Ctx* c, const float* hi, int n_hi, int mem_in, int property, double file_box,
                     const mxr_grid_params* p, float* out);
int run_read_grid_gbptrees(Ctx* c, const char* fname, int property, const mxr_grid_params* p, float* out);

// galaxies to grids (mxr_baryon.hip)
int run_construct_baryon_grids(Ctx* c, const mxr_baryon_params* p, const mxr_galaxies* gal,
                               const mxr_baryon_grids* grids, long* n_below);
int prof_begin(Ctx* c);                              // returns the event index (or -1)
int prof_end(Ctx* c, int ev, int cls, double bytes); // records the launch
int prof_resolve(Ctx* c);                            // after a stream sync
const char* kclass_name(int cls);

// launch-shape tuning and its wisdom file (mxr_tune.hip)
int tuning_check(const mxr_tuning* t);
void tuning_get(const Ctx& c, mxr_tuning* t);
void tuning_set(Ctx& c, const mxr_tuning& t);
int device_key(const Ctx& c, std::string& arch, int& cus);
int wisdom_lookup(const char* path, const char* arch, int cus, int N, int filter, int spin, mxr_tuning* t);
int wisdom_store(const char* path, const char* arch, int cus, int N, int filter, int spin, const mxr_tuning& t,
                 double ms);
int autotune(Ctx* c, const mxr_hii_params* p, const mxr_hii_grids* g, mxr_tuning* best, double* best_ms);

}  // namespace mxr

struct mxr_ctx {
  mxr::Ctx c;
};
