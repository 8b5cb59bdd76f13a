// mxr_api.hip -- C ABI entry points, context lifetime, error reporting and the host-side
// physics (units, RtoM, radius schedule, T_RECFAST) of the reionization grid path.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "mxr_internal.hpp"
#include "mxr_recomb.hpp"
#include "mxr_ts.hpp"

namespace mxr {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int hip_fail(hipError_t e, const char* what, const char* file, int line)
{
  set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
  return (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? MXR_ERR_OOM : MXR_ERR_HIP;
}

// ---- units and cosmology -----------------------------------------------------------------
// set_units (src/core/init.c:203-225): UnitTime = L/V, G in internal units, Hubble = HUBBLE UT,
// RhoCrit = 3 Hubble^2 / (8 pi G).
static double unit_time_s(const mxr_hii_params* p) { return p->UnitLength_in_cm / p->UnitVelocity_in_cm_per_s; }

double rho_crit(const mxr_hii_params* p)
{
  const double UT = unit_time_s(p);
  const double G = GRAVITY / std::pow(p->UnitLength_in_cm, 3) * p->UnitMass_in_g * std::pow(UT, 2);
  const double Hub = HUBBLE * UT;
  return 3 * Hub * Hub / (8 * M_PI * G);
}

// hubble_time (src/core/virial_properties.c:6-20, 79-96): 1 / (Hubble E(z)).
double hubble_time(const mxr_hii_params* p)
{
  const double opz = 1. + p->redshift;
  const double Ez = std::sqrt(p->OmegaM * opz * opz * opz + p->OmegaK * opz * opz + p->OmegaLambda);
  return 1.0 / (HUBBLE * unit_time_s(p) * Ez);
}

// cT_approx (src/core/XRayHeatingFunctions.c:566-569)
float cT_approx(float z) { return 0.58 - 0.006 * (z - 10.0); }

// Natural cubic spline through (x, y) evaluated at xq -- the interpolant GSL's
// gsl_interp_cspline builds (S'' = 0 at both ends), written from the textbook tridiagonal
// system for the knot second derivatives.
static double natural_spline(const double* x, const double* y, int n, double xq)
{
  std::vector<double> M(n, 0.0), cp(n, 0.0), dp(n, 0.0);
  for (int i = 1; i < n - 1; i++) {
    const double h0 = x[i] - x[i - 1], h1 = x[i + 1] - x[i];
    const double r = 6.0 * ((y[i + 1] - y[i]) / h1 - (y[i] - y[i - 1]) / h0);
    const double den = 2.0 * (h0 + h1) - (i > 1 ? h0 * cp[i - 1] : 0.0);
    cp[i] = h1 / den;
    dp[i] = (r - (i > 1 ? h0 * dp[i - 1] : 0.0)) / den;
  }
  for (int i = n - 2; i >= 1; i--)
    M[i] = dp[i] - (i < n - 2 ? cp[i] * M[i + 1] : 0.0);
  int lo = 0, hi = n - 1;
  while (hi > lo + 1) {
    const int mid = (lo + hi) / 2;
    if (x[mid] > xq)
      hi = mid;
    else
      lo = mid;
  }
  const double h = x[lo + 1] - x[lo], t = xq - x[lo];
  const double b = (y[lo + 1] - y[lo]) / h - h * (2.0 * M[lo] + M[lo + 1]) / 6.0;
  return y[lo] + t * (b + t * (M[lo] / 2.0 + t * (M[lo + 1] - M[lo]) / (6.0 * h)));
}

// T_RECFAST (src/core/XRayHeatingFunctions.c:134-192): IGM temperature from the 501-row
// recfast_LCDM.dat table (columns z, x_e, T_gas?, T_K; z descending), read as floats.
double T_RECFAST(const char* tables_dir, float z)
{
  if (!tables_dir) {
    set_error("T_RECFAST: no table directory");
    return -1;
  }
  // the table is read once per directory and kept
  struct Table {
    std::string dir;
    std::vector<double> zt, TK;
  };
  static std::mutex mu;
  static std::vector<Table> cache;
  std::vector<double> zt, TK;
  {
    std::lock_guard<std::mutex> lk(mu);
    for (const Table& t : cache)
      if (t.dir == tables_dir) {
        zt = t.zt;
        TK = t.TK;
      }
  }
  if (zt.empty()) {
    std::string fname = std::string(tables_dir) + "/recfast_LCDM.dat";
    FILE* F = std::fopen(fname.c_str(), "r");
    if (!F) {
      set_error("T_RECFAST: unable to open %s", fname.c_str());
      return -1;
    }
    zt.assign(RECFAST_NPTS, 0.0);
    TK.assign(RECFAST_NPTS, 0.0);
    for (int i = RECFAST_NPTS - 1; i >= 0; i--) {
      float currz, trash, currTK;
      if (std::fscanf(F, "%f %E %E %E", &currz, &trash, &trash, &currTK) != 4) {
        std::fclose(F);
        set_error("T_RECFAST: %s is truncated", fname.c_str());
        return -1;
      }
      zt[i] = currz;
      TK[i] = currTK;
    }
    std::fclose(F);
    std::lock_guard<std::mutex> lk(mu);
    cache.push_back(Table{tables_dir, zt, TK});
  }
  if (z > zt[RECFAST_NPTS - 1]) {
    set_error("T_RECFAST called with z=%f beyond the table", (double)z);
    return -1;
  }
  return natural_spline(zt.data(), TK.data(), RECFAST_NPTS, z);
}

int check_hii_params(const mxr_hii_params* p, int N)
{
  if (!p) {
    set_error("params is NULL");
    return MXR_ERR_INVALID_ARG;
  }
  if (p->ReionGridDim != N) {
    set_error("ReionGridDim %d does not match the context's %d", p->ReionGridDim, N);
    return MXR_ERR_INVALID_ARG;
  }
  if (!(p->BoxSize > 0) || !(p->ReionDeltaRFactor > 1.0) || !(p->UnitLength_in_cm > 0) ||
      !(p->UnitVelocity_in_cm_per_s > 0) || !(p->UnitMass_in_g > 0) || !(p->redshift >= 0) ||
      !(p->OmegaM > 0)) {
    set_error("non-physical parameter (BoxSize, ReionDeltaRFactor, units, redshift or OmegaM)");
    return MXR_ERR_INVALID_ARG;
  }
  if (p->ReionFilterType < 0 || p->ReionFilterType > 2) {
    set_error("ReionFilterType type %d is undefined", p->ReionFilterType);
    return MXR_ERR_FILTER;
  }
  if (p->ReionRtoMFilterType < 0 || p->ReionRtoMFilterType > 1) {
    set_error("Unrecognised RtoM filter (%d)", p->ReionRtoMFilterType);
    return MXR_ERR_INVALID_ARG;
  }
  if (p->Flag_IncludeRecombinations && !(p->Y_He > 0 && p->Y_He < 1)) {
    set_error("Flag_IncludeRecombinations needs 0 < Y_He < 1 (got %g)", p->Y_He);
    return MXR_ERR_INVALID_ARG;
  }
  return MXR_OK;
}

static const char* const kclass_names[K_NCLASS] = {
  "fwd_r2c_z", "fwd_fft_y", "fwd_fft_x_scale", "inv_fft_x_filter", "inv_fft_y", "inv_c2r_z_update",
  "plain_copy", "plain_filter", "plain_clamp", "plain_criterion", "plain_scale_fill", "postloop_sums",
  "ts_source", "ts_fwd_fft", "ts_inv_fft_x_filter", "ts_inv_fft_y", "ts_inv_c2r_z_store", "ts_reduce_r0",
  "ts_evolve"};

const char* kclass_name(int cls) { return (cls >= 0 && cls < K_NCLASS) ? kclass_names[cls] : "?"; }

int prof_begin(Ctx* c)
{
  if (!c->prof)
    return -1;
  while ((int)c->evpool.size() < c->evnext + 2) {
    hipEvent_t e;
    MXR_HIP(hipEventCreate(&e));
    c->evpool.push_back(e);
  }
  const int ev = c->evnext;
  c->evnext += 2;
  MXR_HIP(hipEventRecord(c->evpool[ev], c->stream));
  return ev;
}

int prof_end(Ctx* c, int ev, int cls, double bytes)
{
  if (ev < 0)
    return MXR_OK;
  MXR_HIP(hipEventRecord(c->evpool[ev + 1], c->stream));
  c->recs.push_back(Ctx::Rec{cls, ev, bytes});
  return MXR_OK;
}

int prof_resolve(Ctx* c)
{
  for (const auto& r : c->recs) {
    float ms = 0.f;
    MXR_HIP(hipEventElapsedTime(&ms, c->evpool[r.ev], c->evpool[r.ev + 1]));
    c->kstat[r.cls].launches++;
    c->kstat[r.cls].total_ms += ms;
    c->kstat[r.cls].alg_bytes += r.bytes;
  }
  c->recs.clear();
  c->evnext = 0;
  return MXR_OK;
}

}  // namespace mxr

using namespace mxr;

extern "C" {

const char* mxr_strerror(int code)
{
  switch (code) {
    case MXR_OK: return "success";
    case MXR_ERR_INVALID_ARG: return "invalid argument";
    case MXR_ERR_UNSUPPORTED: return "unsupported configuration";
    case MXR_ERR_HIP: return "HIP runtime error";
    case MXR_ERR_OOM: return "device out of memory";
    case MXR_ERR_TABLES: return "physics table unavailable";
    case MXR_ERR_NO_DEVICE: return "no usable gfx950 device";
    case MXR_ERR_FILTER: return "undefined ReionFilterType";
    case MXR_ERR_NO_WISDOM: return "no wisdom entry for this device and shape";
    default: return "unknown error";
  }
}

const char* mxr_last_error(void) { return g_err; }

int mxr_supported_dims(int* dims, int capacity)
{
  int n = 0;
#define X(d)                                                                                                         \
  if (dims && n < capacity)                                                                                          \
    dims[n] = d;                                                                                                     \
  n++;
  MXR_DIMS(X)
#undef X
  return n;
}

int mxr_create(mxr_ctx** out, int device, int grid_dim)
{
  if (!out) {
    set_error("ctx pointer is NULL");
    return MXR_ERR_INVALID_ARG;
  }
  *out = nullptr;
  if (!supported_dim(grid_dim)) {
    set_error("grid dimension %d is not supported (need one of 16,32,64,96,128,160,256,512,1024)",
              grid_dim);
    return MXR_ERR_INVALID_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) {
    set_error("HIP device %d not available (%d visible)", device, ndev);
    return MXR_ERR_NO_DEVICE;
  }
  hipDeviceProp_t prop;
  MXR_HIP(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    set_error("device %d is %s, this library is built for gfx950", device, prop.gcnArchName);
    return MXR_ERR_NO_DEVICE;
  }
  mxr_ctx* h = new (std::nothrow) mxr_ctx();
  if (!h) {
    set_error("host allocation failed");
    return MXR_ERR_OOM;
  }
  Ctx& c = h->c;
  c.device = device;
  c.N = grid_dim;
  int rc = MXR_OK;
  auto fail = [&](int code) {
    mxr_destroy(h);
    return code;
  };
  if (hipSetDevice(device) != hipSuccess)
    return fail(hip_fail(hipErrorInvalidDevice, "hipSetDevice", __FILE__, __LINE__));
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking)) != hipSuccess)
    return fail(hip_fail(e, "hipStreamCreate", __FILE__, __LINE__));
  if ((e = hipEventCreate(&c.ev0)) != hipSuccess || (e = hipEventCreate(&c.ev1)) != hipSuccess)
    return fail(hip_fail(e, "hipEventCreate", __FILE__, __LINE__));
  std::vector<float2> tw(grid_dim);
  for (int m = 0; m < grid_dim; m++) {
    const double a = -2.0 * M_PI * (double)m / (double)grid_dim;
    tw[m] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  if ((e = hipMalloc(&c.tw, sizeof(float2) * grid_dim)) != hipSuccess)
    return fail(hip_fail(e, "hipMalloc(twiddles)", __FILE__, __LINE__));
  if ((e = hipMemcpy(c.tw, tw.data(), sizeof(float2) * grid_dim, hipMemcpyHostToDevice)) != hipSuccess)
    return fail(hip_fail(e, "hipMemcpy(twiddles)", __FILE__, __LINE__));
  if ((e = hipMalloc(&c.sums, sizeof(double) * 32)) != hipSuccess)
    return fail(hip_fail(e, "hipMalloc(sums)", __FILE__, __LINE__));
  if ((e = hipHostMalloc((void**)&c.sums_host, sizeof(double) * 32, hipHostMallocDefault)) != hipSuccess)
    return fail(hip_fail(e, "hipHostMalloc(sums)", __FILE__, __LINE__));
  (void)rc;
  *out = h;
  return MXR_OK;
}

void mxr_destroy(mxr_ctx* h)
{
  if (!h)
    return;
  Ctx& c = h->c;
  (void)hipSetDevice(c.device);
  if (c.stream)
    (void)hipStreamSynchronize(c.stream);
  (void)hipFree(c.tw);
  for (int g = 0; g < 7; g++) {
    (void)hipFree(c.spec[g]);
    (void)hipFree(c.work[g]);
  }
  for (int g = 0; g < 7; g++)
    (void)hipFree(c.in_stage[g]);
  for (int g = 0; g < 10; g++)
    (void)hipFree(c.out_stage[g]);
  (void)hipFree(c.rec_tab);
  (void)hipFree(c.rec_coef);
  (void)hipFree(c.rec_x);
  (void)hipFree(c.ion);
  for (int g = 0; g < 4; g++)
    (void)hipFree(c.ts_src[g]);
  (void)hipFree(c.ts_store);
  (void)hipFree(c.ts_hist);
  (void)hipFree(c.ts_dev);
  if (c.ts_dev_host)
    (void)hipHostFree(c.ts_dev_host);
  if (c.ev_red)
    (void)hipEventDestroy(c.ev_red);
  (void)hipFree(c.partials);
  (void)hipFree(c.ps_part);
  (void)hipFree(c.bg_ws);
  if (c.bg_ev_in)
    (void)hipEventDestroy(c.bg_ev_in);
  for (int k = 0; k < BG_STREAMS; k++) {
    if (c.bg_ev_shift[k])
      (void)hipEventDestroy(c.bg_ev_shift[k]);
    if (c.bg_stream[k])
      (void)hipStreamDestroy(c.bg_stream[k]);
  }
  (void)hipFree(c.wtab);
  if (c.hires)
    mxr_destroy(c.hires);
  (void)hipFree(c.sums);
  if (c.sums_host)
    (void)hipHostFree(c.sums_host);
  for (auto e : c.evpool)
    (void)hipEventDestroy(e);
  if (c.ev0)
    (void)hipEventDestroy(c.ev0);
  if (c.ev1)
    (void)hipEventDestroy(c.ev1);
  if (c.stream)
    (void)hipStreamDestroy(c.stream);
  delete h;
}

int mxr_set_path(mxr_ctx* h, int path)
{
  if (!h || (path != MXR_PATH_FUSED && path != MXR_PATH_PLAIN && path != MXR_PATH_FUSED_DENSE)) {
    set_error("bad context or path %d", path);
    return MXR_ERR_INVALID_ARG;
  }
  h->c.path = path;
  return MXR_OK;
}

int mxr_get_tuning(const mxr_ctx* h, mxr_tuning* t)
{
  if (!h || !t) {
    set_error("null context or tuning");
    return MXR_ERR_INVALID_ARG;
  }
  tuning_get(h->c, t);
  return MXR_OK;
}

int mxr_set_tuning(mxr_ctx* h, const mxr_tuning* t)
{
  if (!h || !t) {
    set_error("null context or tuning");
    return MXR_ERR_INVALID_ARG;
  }
  RC(tuning_check(t));
  tuning_set(h->c, *t);
  return MXR_OK;
}

int mxr_wisdom_lookup(const char* path, const char* arch, int cus, int grid_dim, int filter_type, int spin,
                      mxr_tuning* t)
{
  if (!path || !arch || !t) {
    set_error("null wisdom path, architecture or tuning");
    return MXR_ERR_INVALID_ARG;
  }
  return wisdom_lookup(path, arch, cus, grid_dim, filter_type, spin ? 1 : 0, t);
}

int mxr_wisdom_store(const char* path, const char* arch, int cus, int grid_dim, int filter_type, int spin,
                     const mxr_tuning* t, double ms)
{
  if (!path || !arch || !t) {
    set_error("null wisdom path, architecture or tuning");
    return MXR_ERR_INVALID_ARG;
  }
  return wisdom_store(path, arch, cus, grid_dim, filter_type, spin ? 1 : 0, *t, ms);
}

int mxr_load_wisdom(mxr_ctx* h, const mxr_hii_params* p, const char* path)
{
  if (!h || !p || !path) {
    set_error("null context, params or wisdom path");
    return MXR_ERR_INVALID_ARG;
  }
  std::string arch;
  int cus = 0;
  RC(device_key(h->c, arch, cus));
  mxr_tuning t;
  RC(wisdom_lookup(path, arch.c_str(), cus, h->c.N, p->ReionFilterType, p->Flag_IncludeSpinTemp ? 1 : 0, &t));
  tuning_set(h->c, t);
  return MXR_OK;
}

int mxr_autotune(mxr_ctx* h, const mxr_hii_params* p, const mxr_hii_grids* g, const char* path, mxr_tuning* chosen)
{
  if (!h || !p || !g) {
    set_error("null context, params or grids");
    return MXR_ERR_INVALID_ARG;
  }
  RC(check_hii_params(p, h->c.N));
  if (!g->deltax || !g->stars || !g->weighted_sfr || (p->Flag_IncludeSpinTemp && (!g->x_e_box || !g->Tk_box)) ||
      (p->Flag_BHFeedback && (!g->effective_bhm || !g->effective_bhar)) ||
      (g->mem != MXR_MEM_HOST && g->mem != MXR_MEM_DEVICE) ||
      (g->storage != MXR_STORAGE_F32 && g->storage != MXR_STORAGE_BF16)) {
    set_error("missing input grid, bad memory kind or bad storage type");
    return MXR_ERR_INVALID_ARG;
  }
  if (hipSetDevice(h->c.device) != hipSuccess) {
    set_error("hipSetDevice(%d) failed", h->c.device);
    return MXR_ERR_HIP;
  }
  mxr_tuning best;
  double ms = 0;
  RC(autotune(&h->c, p, g, &best, &ms));
  if (chosen)
    *chosen = best;
  if (path) {
    std::string arch;
    int cus = 0;
    RC(device_key(h->c, arch, cus));
    RC(wisdom_store(path, arch.c_str(), cus, h->c.N, p->ReionFilterType, p->Flag_IncludeSpinTemp ? 1 : 0, best, ms));
  }
  return MXR_OK;
}

// find_HII_bubbles.c:27-46
double mxr_RtoM(const mxr_hii_params* p, double R)
{
  if (!p)
    return -1;
  const double OmegaM = p->OmegaM;
  const double RhoCrit = rho_crit(p);
  switch (p->ReionRtoMFilterType) {
    case 0: return (4.0 / 3.0) * M_PI * std::pow(R, 3) * (OmegaM * RhoCrit);
    case 1: return std::pow(2 * M_PI, 1.5) * OmegaM * RhoCrit * std::pow(R, 3);
    default: set_error("Unrecognised filter (%d)", p->ReionRtoMFilterType); return -1;
  }
}

// find_HII_bubbles.c:145-148, 256-285, 546
int mxr_hii_radii(const mxr_hii_params* p, double* radii, int capacity)
{
  if (!p || p->ReionGridDim <= 0 || !(p->ReionDeltaRFactor > 1.0) || !(p->BoxSize > 0)) {
    set_error("bad parameters for the radius schedule");
    return MXR_ERR_INVALID_ARG;
  }
  const int N = p->ReionGridDim;
  const double box_size = p->BoxSize;
  const double clf = (box_size / (double)N) < 1.0 ? 1.0 : L_FACTOR;
  double Rmax;
  if (p->Flag_EvolvingReionRBubbleMax)
    Rmax = (p->redshift > 6.) ? 25.483241248322766 : 112. * std::pow((1. + p->redshift) / 5., -4.4);
  else
    Rmax = p->Flag_IncludeRecombinations ? p->ReionRBubbleMaxRecomb : p->ReionRBubbleMax;
  double R = std::fmin(Rmax, L_FACTOR * box_size);
  int n = 0;
  bool last = false;
  while (!last) {
    if (((R / p->ReionDeltaRFactor) <= (clf * box_size / (double)N)) ||
        ((R / p->ReionDeltaRFactor) <= p->ReionRBubbleMin)) {
      last = true;
      R = clf * box_size / (double)N;
    }
    if (radii && n < capacity)
      radii[n] = R;
    n++;
    R /= p->ReionDeltaRFactor;
  }
  return n;
}

int mxr_find_hii_bubbles(mxr_ctx* h, const mxr_hii_params* params, const mxr_hii_grids* grids,
                         mxr_hii_globals* globals)
{
  if (!h || !grids || !globals) {
    set_error("NULL context, grids or globals");
    return MXR_ERR_INVALID_ARG;
  }
  int rc = check_hii_params(params, h->c.N);
  if (rc)
    return rc;
  if (!grids->deltax || !grids->stars || !grids->weighted_sfr || !grids->xH || !grids->z_at_ionization ||
      !grids->J_21_at_ionization || !grids->r_bubble || !grids->temp_kinetic_all_gas ||
      (params->Flag_IncludeSpinTemp && (!grids->x_e_box || !grids->Tk_box)) ||
      (params->Flag_BHFeedback && (!grids->effective_bhm || !grids->effective_bhar)) ||
      (grids->mem != MXR_MEM_HOST && grids->mem != MXR_MEM_DEVICE) ||
      (grids->storage != MXR_STORAGE_F32 && grids->storage != MXR_STORAGE_BF16)) {
    set_error("missing grid pointer, bad memory kind or bad storage type");
    return MXR_ERR_INVALID_ARG;
  }
  if (hipSetDevice(h->c.device) != hipSuccess) {
    set_error("hipSetDevice(%d) failed", h->c.device);
    return MXR_ERR_HIP;
  }
  return run_find_hii_bubbles(&h->c, params, grids, globals);
}

int mxr_recombination_tables(mxr_ctx* h, const mxr_hii_params* p, double* lnGamma, double* ln_rate,
                             double* clumping, double* ln_residual)
{
  if (!h || !p) {
    set_error("null context or params");
    return MXR_ERR_INVALID_ARG;
  }
  if (hipSetDevice(h->c.device) != hipSuccess) {
    set_error("hipSetDevice(%d) failed", h->c.device);
    return MXR_ERR_HIP;
  }
  return rec_tables_copy(&h->c, p, lnGamma, ln_rate, clumping, ln_residual);
}

int mxr_filter(mxr_ctx* h, void* box, int mem, int local_ix_start, int slab_nx, int grid_dim, double BoxSize,
               float R, int filter_type)
{
  if (!h || !box || slab_nx <= 0 || local_ix_start < 0 || local_ix_start + slab_nx > grid_dim ||
      grid_dim != h->c.N || !(BoxSize > 0) || (mem != MXR_MEM_HOST && mem != MXR_MEM_DEVICE)) {
    set_error("bad filter arguments");
    return MXR_ERR_INVALID_ARG;
  }
  if (filter_type < 0 || filter_type > 2) {
    set_error("ReionFilterType type %d is undefined", filter_type);
    return MXR_ERR_FILTER;
  }
  if (hipSetDevice(h->c.device) != hipSuccess) {
    set_error("hipSetDevice(%d) failed", h->c.device);
    return MXR_ERR_HIP;
  }
  // reionization.c:3376-3377: box size and delta_k in float
  const float box_size = (float)BoxSize;
  const float dk = (float)(2.0 * M_PI / box_size);
  Ctx& c = h->c;
  const long n = (long)slab_nx * grid_dim * (grid_dim / 2 + 1);
  if (mem == MXR_MEM_DEVICE)
    return run_filter(&c, (float2*)box, local_ix_start, slab_nx, grid_dim, dk, R, filter_type);
  float2* d = nullptr;
  MXR_HIP(hipMalloc(&d, sizeof(float2) * n));
  int rc = MXR_OK;
  hipError_t e = hipMemcpyAsync(d, box, sizeof(float2) * n, hipMemcpyHostToDevice, c.stream);
  if (e == hipSuccess)
    rc = run_filter(&c, d, local_ix_start, slab_nx, grid_dim, dk, R, filter_type);
  else
    rc = hip_fail(e, "hipMemcpyAsync", __FILE__, __LINE__);
  if (rc == MXR_OK) {
    e = hipMemcpyAsync(box, d, sizeof(float2) * n, hipMemcpyDeviceToHost, c.stream);
    if (e == hipSuccess)
      e = hipStreamSynchronize(c.stream);
    if (e != hipSuccess)
      rc = hip_fail(e, "hipMemcpyAsync", __FILE__, __LINE__);
  }
  (void)hipFree(d);
  return rc;
}

int mxr_velocity_gradient(mxr_ctx* h, void* box, int mem, int slab_nx, int grid_dim, double BoxSize)
{
  if (!h || !box || slab_nx <= 0 || slab_nx > grid_dim || grid_dim != h->c.N || !(BoxSize > 0) ||
      (mem != MXR_MEM_HOST && mem != MXR_MEM_DEVICE)) {
    set_error("bad velocity_gradient arguments");
    return MXR_ERR_INVALID_ARG;
  }
  if (hipSetDevice(h->c.device) != hipSuccess) {
    set_error("hipSetDevice(%d) failed", h->c.device);
    return MXR_ERR_HIP;
  }
  const float box_size = (float)BoxSize;  // reionization.c:3438-3439
  const float dk = (float)(2.0 * M_PI / box_size);
  Ctx& c = h->c;
  const long n = (long)slab_nx * grid_dim * (grid_dim / 2 + 1);
  if (mem == MXR_MEM_DEVICE)
    return run_velocity_gradient(&c, (float2*)box, slab_nx, grid_dim, dk);
  float2* d = nullptr;
  MXR_HIP(hipMalloc(&d, sizeof(float2) * n));
  int rc = MXR_OK;
  hipError_t e = hipMemcpyAsync(d, box, sizeof(float2) * n, hipMemcpyHostToDevice, c.stream);
  if (e == hipSuccess)
    rc = run_velocity_gradient(&c, d, slab_nx, grid_dim, dk);
  else
    rc = hip_fail(e, "hipMemcpyAsync", __FILE__, __LINE__);
  if (rc == MXR_OK) {
    e = hipMemcpyAsync(box, d, sizeof(float2) * n, hipMemcpyDeviceToHost, c.stream);
    if (e == hipSuccess)
      e = hipStreamSynchronize(c.stream);
    if (e != hipSuccess)
      rc = hip_fail(e, "hipMemcpyAsync", __FILE__, __LINE__);
  }
  (void)hipFree(d);
  return rc;
}

// ---- batches -----------------------------------------------------------------------------
}  // extern "C" (reopened below)

struct mxr_batch {
  int device = 0;
  std::vector<mxr_ctx*> ctx;
  hipStream_t cap = nullptr;  // origin stream of the capture and of the graph launch
  hipEvent_t fork = nullptr, t0 = nullptr, t1 = nullptr;
  std::vector<hipEvent_t> join;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<unsigned char> key;  // parameters and grids the instantiated graph was captured with
};

static void batch_drop_graph(mxr_batch* b)
{
  if (b->exec)
    (void)hipGraphExecDestroy(b->exec);
  if (b->graph)
    (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->key.clear();
}

extern "C" {

int mxr_batch_create(mxr_batch** out, int device, int n, const int* grid_dims)
{
  if (!out || n <= 0 || !grid_dims) {
    set_error("bad batch arguments");
    return MXR_ERR_INVALID_ARG;
  }
  *out = nullptr;
  mxr_batch* b = new (std::nothrow) mxr_batch();
  if (!b) {
    set_error("host allocation failed");
    return MXR_ERR_OOM;
  }
  b->device = device;
  for (int i = 0; i < n; i++) {
    mxr_ctx* c = nullptr;
    const int rc = mxr_create(&c, device, grid_dims[i]);
    if (rc) {
      mxr_batch_destroy(b);
      return rc;
    }
    b->ctx.push_back(c);
  }
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming)) != hipSuccess ||
      (e = hipEventCreate(&b->t0)) != hipSuccess || (e = hipEventCreate(&b->t1)) != hipSuccess) {
    const int rc = hip_fail(e, "batch stream/event creation", __FILE__, __LINE__);
    mxr_batch_destroy(b);
    return rc;
  }
  b->join.assign(n, nullptr);
  for (int i = 0; i < n; i++)
    if ((e = hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming)) != hipSuccess) {
      const int rc = hip_fail(e, "hipEventCreate", __FILE__, __LINE__);
      mxr_batch_destroy(b);
      return rc;
    }
  *out = b;
  return MXR_OK;
}

void mxr_batch_destroy(mxr_batch* b)
{
  if (!b)
    return;
  (void)hipSetDevice(b->device);
  if (b->cap)
    (void)hipStreamSynchronize(b->cap);
  batch_drop_graph(b);
  for (auto e : b->join)
    if (e)
      (void)hipEventDestroy(e);
  for (hipEvent_t e : {b->fork, b->t0, b->t1})
    if (e)
      (void)hipEventDestroy(e);
  if (b->cap)
    (void)hipStreamDestroy(b->cap);
  for (auto c : b->ctx)
    mxr_destroy(c);
  delete b;
}

// The bytes a captured realisation depends on, field by field (no struct padding, whose
// contents a caller's stack leaves unspecified): every parameter, the tables directory's text
// (it sets T_RECFAST when TK_recfast < 0), every grid pointer and the launch shapes.
extern "C++" {
template <class T>
static void key_put(std::vector<unsigned char>& k, const T& v)
{
  const unsigned char* p = reinterpret_cast<const unsigned char*>(&v);
  k.insert(k.end(), p, p + sizeof(T));
}

static void graph_key_append(std::vector<unsigned char>& k, const mxr_hii_params& p, const mxr_hii_grids& g,
                             const mxr_tuning& t)
{
  key_put(k, p.ReionGridDim);
  for (double v : {p.BoxSize, p.Hubble_h, p.OmegaM, p.OmegaLambda, p.OmegaK, p.BaryonFrac, p.ReionDeltaRFactor})
    key_put(k, v);
  for (int v : {p.ReionFilterType, p.ReionRtoMFilterType, p.ReionUVBFlag, p.Flag_IncludeSpinTemp,
                p.Flag_IncludeRecombinations, p.Flag_EvolvingReionRBubbleMax, p.Flag_BHFeedback,
                p.Flag_TemperatureDependentRec})
    key_put(k, v);
  for (double v : {p.ReionSfrTimescale, p.ReionEfficiency, p.ReionNionPhotPerBary, p.ReionGammaHaloBias,
                   p.ReionAlphaUV, p.ReionAlphaUVBH, p.ReionRBubbleMax, p.ReionRBubbleMaxRecomb, p.ReionRBubbleMin,
                   p.UnitLength_in_cm, p.UnitMass_in_g, p.UnitVelocity_in_cm_per_s, p.redshift, p.prev_redshift,
                   p.TK_recfast, p.Y_He})
    key_put(k, v);
  const char* dir = p.TablesForXHeatingDir ? p.TablesForXHeatingDir : "";
  const size_t len = std::strlen(dir);
  key_put(k, len);
  k.insert(k.end(), dir, dir + len);
  key_put(k, g.mem);
  key_put(k, g.storage);
  for (const void* q : {(const void*)g.deltax, (const void*)g.stars, (const void*)g.weighted_sfr,
                        (const void*)g.x_e_box, (const void*)g.Tk_box, (const void*)g.xH,
                        (const void*)g.z_at_ionization, (const void*)g.J_21_at_ionization, (const void*)g.r_bubble,
                        (const void*)g.temp_kinetic_all_gas, (const void*)g.effective_bhm,
                        (const void*)g.effective_bhar, (const void*)g.N_rec, (const void*)g.Gamma12,
                        (const void*)g.residual_xH, (const void*)g.clumping_factor, (const void*)g.t_resp})
    key_put(k, q);
  for (int v : {t.zdirect_kmax, t.zdirect_rpb, t.live_tiles, t.col_tc_x, t.col_tc_y})
    key_put(k, v);
}
}  // extern "C++"

int mxr_find_hii_bubbles_batched(mxr_batch* b, const mxr_hii_params* params, const mxr_hii_grids* grids,
                                 mxr_hii_globals* globals, int mode)
{
  if (!b || !params || !grids || !globals ||
      (mode != MXR_BATCH_GRAPH && mode != MXR_BATCH_STREAMS && mode != MXR_BATCH_SERIAL)) {
    set_error("bad batch call arguments (mode %d)", mode);
    return MXR_ERR_INVALID_ARG;
  }
  const int n = (int)b->ctx.size();
  for (int i = 0; i < n; i++) {
    const int rc = check_hii_params(&params[i], b->ctx[i]->c.N);
    if (rc)
      return rc;
    const mxr_hii_grids& g = grids[i];
    if (g.mem != MXR_MEM_DEVICE || !g.deltax || !g.stars || !g.weighted_sfr || !g.xH || !g.z_at_ionization ||
        !g.J_21_at_ionization || !g.r_bubble || !g.temp_kinetic_all_gas ||
        (params[i].Flag_IncludeSpinTemp && (!g.x_e_box || !g.Tk_box)) ||
        (params[i].Flag_BHFeedback && (!g.effective_bhm || !g.effective_bhar)) ||
        (params[i].Flag_IncludeRecombinations &&
         (!g.N_rec || !g.Gamma12 || !g.residual_xH || !g.clumping_factor || !g.t_resp))) {
      set_error("realisation %d: batched grids must all be device memory", i);
      return MXR_ERR_INVALID_ARG;
    }
  }
  if (hipSetDevice(b->device) != hipSuccess) {
    set_error("hipSetDevice(%d) failed", b->device);
    return MXR_ERR_HIP;
  }
  float ms = 0.f;
  if (mode == MXR_BATCH_SERIAL) {
    double total = 0;
    for (int i = 0; i < n; i++) {
      const int rc = run_find_hii_bubbles(&b->ctx[i]->c, &params[i], &grids[i], &globals[i]);
      if (rc)
        return rc;
      total += globals[i].device_ms;
    }
    for (int i = 0; i < n; i++)
      globals[i].device_ms = total;
    return MXR_OK;
  }
  for (int i = 0; i < n; i++) {
    Ctx& c = b->ctx[i]->c;
    // no allocation inside a capture: buffers and the recombination tables are made first
    int rc = ensure_buffers(&c, 3 + (params[i].Flag_IncludeSpinTemp ? 1 : 0) + (params[i].Flag_BHFeedback ? 2 : 0) +
                                  (params[i].Flag_IncludeRecombinations ? 1 : 0));
    if (rc)
      return rc;
    if (params[i].Flag_IncludeRecombinations) {
      RecTab rt;
      if ((rc = rec_tables(&c, &params[i], &rt)))
        return rc;
    }
    c.prof = false;
  }
  auto enqueue_all = [&]() -> int {
    MXR_HIP(hipEventRecord(b->fork, b->cap));
    for (int i = 0; i < n; i++) {
      Ctx& c = b->ctx[i]->c;
      MXR_HIP(hipStreamWaitEvent(c.stream, b->fork, 0));
      c.enqueue_only = true;
      const int rc = run_find_hii_bubbles(&c, &params[i], &grids[i], &globals[i]);
      c.enqueue_only = false;
      if (rc)
        return rc;
      MXR_HIP(hipEventRecord(b->join[i], c.stream));
      MXR_HIP(hipStreamWaitEvent(b->cap, b->join[i], 0));
    }
    return MXR_OK;
  };
  if (mode == MXR_BATCH_STREAMS) {
    MXR_HIP(hipEventRecord(b->t0, b->cap));
    const int rc = enqueue_all();
    if (rc)
      return rc;
    MXR_HIP(hipEventRecord(b->t1, b->cap));
  } else {
    // the captured kernels depend on the parameters, the grids and each context's launch shapes
    std::vector<unsigned char> key;
    for (int i = 0; i < n; i++) {
      mxr_tuning t;
      tuning_get(b->ctx[i]->c, &t);
      graph_key_append(key, params[i], grids[i], t);
    }
    if (!b->exec || key != b->key) {
      batch_drop_graph(b);
      MXR_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
      const int rc = enqueue_all();
      hipGraph_t gph = nullptr;
      const hipError_t e = hipStreamEndCapture(b->cap, &gph);
      if (rc || e != hipSuccess) {
        if (gph)
          (void)hipGraphDestroy(gph);
        return rc ? rc : hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
      }
      b->graph = gph;
      MXR_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      b->key = key;
      // the host-side finalisation data of each realisation was recorded while capturing
    }
    MXR_HIP(hipEventRecord(b->t0, b->cap));
    MXR_HIP(hipGraphLaunch(b->exec, b->cap));
    MXR_HIP(hipEventRecord(b->t1, b->cap));
  }
  MXR_HIP(hipStreamSynchronize(b->cap));
  MXR_HIP(hipEventElapsedTime(&ms, b->t0, b->t1));
  for (int i = 0; i < n; i++) {
    hii_globals(&b->ctx[i]->c, &globals[i]);
    globals[i].device_ms = ms;
  }
  return MXR_OK;
}

int mxr_batch_autotune(mxr_batch* b, const mxr_hii_params* params, const mxr_hii_grids* grids, const char* path)
{
  if (!b || !params || !grids) {
    set_error("null batch, params or grids");
    return MXR_ERR_INVALID_ARG;
  }
  struct Done {
    int N, filter, spin;
    mxr_tuning t;
  };
  std::vector<Done> done;
  for (size_t i = 0; i < b->ctx.size(); i++) {
    const int N = b->ctx[i]->c.N, f = params[i].ReionFilterType, sp = params[i].Flag_IncludeSpinTemp ? 1 : 0;
    const Done* hit = nullptr;
    for (const Done& d : done)
      if (d.N == N && d.filter == f && d.spin == sp)
        hit = &d;
    if (hit) {
      tuning_set(b->ctx[i]->c, hit->t);
      continue;
    }
    Done d{N, f, sp, {}};
    RC(mxr_autotune(b->ctx[i], &params[i], &grids[i], path, &d.t));
    done.push_back(d);
  }
  return MXR_OK;
}

int mxr_batch_load_wisdom(mxr_batch* b, const mxr_hii_params* params, const char* path)
{
  if (!b || !params || !path) {
    set_error("null batch, params or wisdom path");
    return MXR_ERR_INVALID_ARG;
  }
  int missing = 0;
  for (size_t i = 0; i < b->ctx.size(); i++) {
    const int rc = mxr_load_wisdom(b->ctx[i], &params[i], path);
    if (rc == MXR_ERR_NO_WISDOM)
      missing++;
    else if (rc)
      return rc;
  }
  if (missing) {
    set_error("%d realisations have no wisdom entry in %s", missing, path);
    return MXR_ERR_NO_WISDOM;
  }
  return MXR_OK;
}

int mxr_compute_ts(mxr_ctx* h, const mxr_ts_params* p, const mxr_ts_grids* g, mxr_ts_globals* globals,
                   mxr_ts_tables* tables)
{
  if (!h || !p || !g || !globals) {
    set_error("NULL context, params, grids or globals");
    return MXR_ERR_INVALID_ARG;
  }
  if (p->ReionGridDim != h->c.N) {
    set_error("ReionGridDim %d does not match the context's %d", p->ReionGridDim, h->c.N);
    return MXR_ERR_INVALID_ARG;
  }
  if (p->TsNumFilterSteps < 1 || p->TsNumFilterSteps > MXR_TS_MAX_STEPS || !p->ZZ || p->snapshot < 0 ||
      !(p->BoxSize > 0) || !(p->Hubble_h > 0) || !(p->OmegaM > 0) || !(p->UnitLength_in_cm > 0) ||
      !(p->UnitMass_in_g > 0) || !(p->UnitVelocity_in_cm_per_s > 0) || !(p->NuXrayThreshold > 0) ||
      p->ReionRtoMFilterType < 0 || p->ReionRtoMFilterType > 1) {
    set_error("bad spin-temperature parameters (TsNumFilterSteps 1..%d, ZZ, snapshot, units, RtoM type)",
              MXR_TS_MAX_STEPS);
    return MXR_ERR_INVALID_ARG;
  }
  if (p->TsHeatingFilterType < 0 || p->TsHeatingFilterType > 2) {
    set_error("TsHeatingFilterType type %d is undefined", p->TsHeatingFilterType);
    return MXR_ERR_FILTER;
  }
  if (!g->deltax || !g->stars || !g->x_e_box_prev || !g->x_e_box || !g->Tk_box || !g->TS_box ||
      !g->stored_fcoll || (g->mem != MXR_MEM_HOST && g->mem != MXR_MEM_DEVICE)) {
    set_error("missing grid pointer or bad memory kind");
    return MXR_ERR_INVALID_ARG;
  }
  if (hipSetDevice(h->c.device) != hipSuccess) {
    set_error("hipSetDevice(%d) failed", h->c.device);
    return MXR_ERR_HIP;
  }
  return run_compute_ts(&h->c, p, g, globals, tables);
}

int mxr_ts_build_tables(const mxr_ts_params* p, double collapse_fraction, double x_e_ave,
                        const double* stored_fcoll, mxr_ts_tables* tables)
{
  if (!p || !p->ZZ || !stored_fcoll || !tables || p->snapshot < 0 || p->TsNumFilterSteps < 1 ||
      p->TsNumFilterSteps > MXR_TS_MAX_STEPS || !(p->BoxSize > 0) || !(p->Hubble_h > 0) || !(p->OmegaM > 0) ||
      p->ReionGridDim <= 0 || !(p->NuXrayThreshold > 0)) {
    set_error("bad parameters for the spin-temperature tables");
    return MXR_ERR_INVALID_ARG;
  }
  TsPlan pl;
  int rc = ts_plan(p, &pl);
  if (rc)
    return rc;
  if (pl.homogeneous) {
    std::memset(tables, 0, sizeof(*tables));
    tables->homogeneous = 1;
    tables->zp = pl.zp;
    tables->dzp = pl.dzp;
    tables->dt_dzp = pl.dt_dzp;
    tables->hist_max = -1;
    return MXR_OK;
  }
  TsDev* dev = new (std::nothrow) TsDev;
  if (!dev) {
    set_error("out of host memory");
    return MXR_ERR_OOM;
  }
  rc = ts_tables(p, pl, collapse_fraction, x_e_ave, stored_fcoll, dev, tables);
  delete dev;
  return rc;
}

int mxr_ts_history_depth(const mxr_ts_params* p, int last_snapshot)
{
  if (!p || !p->ZZ || last_snapshot < 0 || p->TsNumFilterSteps < 1 || p->TsNumFilterSteps > MXR_TS_MAX_STEPS ||
      !(p->BoxSize > 0) || !(p->Hubble_h > 0) || !(p->OmegaM > 0) || p->ReionGridDim <= 0) {
    set_error("bad parameters for the history depth");
    return MXR_ERR_INVALID_ARG;
  }
  return ts_history_depth(p, last_snapshot);
}

int mxr_profile_enable(mxr_ctx* h, int on)
{
  if (!h) {
    set_error("NULL context");
    return MXR_ERR_INVALID_ARG;
  }
  h->c.prof = on != 0;
  h->c.recs.clear();
  h->c.evnext = 0;
  for (auto& k : h->c.kstat)
    k = KStat{};
  return MXR_OK;
}

int mxr_profile_read(mxr_ctx* h, mxr_kernel_stat* out, int capacity)
{
  if (!h || (capacity > 0 && !out)) {
    set_error("bad profile_read arguments");
    return MXR_ERR_INVALID_ARG;
  }
  for (int i = 0; i < K_NCLASS && i < capacity; i++) {
    std::snprintf(out[i].name, sizeof(out[i].name), "%s", kclass_name(i));
    out[i].launches = h->c.kstat[i].launches;
    out[i].total_ms = h->c.kstat[i].total_ms;
    out[i].alg_bytes = h->c.kstat[i].alg_bytes;
  }
  return K_NCLASS;
}

int mxr_brightness_temperature(mxr_ctx* h, const mxr_tb_params* p, const mxr_tb_grids* g, double* ave)
{
  if (!h || !p || !g || !ave || !g->xH || !g->deltax || !g->delta_T || p->ReionGridDim != h->c.N ||
      (p->Flag_IncludeSpinTemp && !g->TS_box) || (g->mem != MXR_MEM_HOST && g->mem != MXR_MEM_DEVICE)) {
    set_error("bad brightness-temperature arguments");
    return MXR_ERR_INVALID_ARG;
  }
  if (p->Flag_IncludePecVelsFor21cm < 0 || p->Flag_IncludePecVelsFor21cm > 3) {
    set_error("Flag_IncludePecVelsFor21cm %d is not in 0..3", p->Flag_IncludePecVelsFor21cm);
    return MXR_ERR_INVALID_ARG;
  }
  if (p->Flag_IncludeSpinTemp && p->Flag_IncludePecVelsFor21cm > 0 && (!g->vel || !(p->BoxSize > 0))) {
    set_error("peculiar-velocity terms need the velocity grid and BoxSize > 0");
    return MXR_ERR_INVALID_ARG;
  }
  if (hipSetDevice(h->c.device) != hipSuccess) {
    set_error("hipSetDevice(%d) failed", h->c.device);
    return MXR_ERR_HIP;
  }
  return run_brightness_temperature(&h->c, p, g, ave);
}

int mxr_brightness_temperature_noalloc(const mxr_tb_params* params, const mxr_tb_grids* grids,
                                       double* volume_ave_Tb, void* stream)
{
  if (!params || !grids) {
    set_error("null params or grids");
    return MXR_ERR_INVALID_ARG;
  }
  return run_brightness_temperature_noalloc(params, grids, volume_ave_Tb, static_cast<hipStream_t>(stream));
}

int mxr_ps_length(int ReionGridDim, double BoxSize, double Hubble_h)
{
  if (ReionGridDim <= 0 || !(BoxSize > 0) || !(Hubble_h > 0)) {
    set_error("bad ps_length arguments");
    return MXR_ERR_INVALID_ARG;
  }
  return ps_length(ReionGridDim, BoxSize, Hubble_h);
}

int mxr_power_spectrum(mxr_ctx* h, const float* delta_T, int mem, double BoxSize, double Hubble_h, int PS_Length,
                       float* PS_k, float* PS_data, float* PS_error)
{
  if (!h || !delta_T || !PS_k || !PS_data || !PS_error || !(BoxSize > 0) || !(Hubble_h > 0) ||
      (mem != MXR_MEM_HOST && mem != MXR_MEM_DEVICE)) {
    set_error("bad power_spectrum arguments");
    return MXR_ERR_INVALID_ARG;
  }
  const int nb = ps_length(h->c.N, BoxSize, Hubble_h);
  if (PS_Length != nb) {
    set_error("PS_Length %d differs from the %d bins of this grid and box", PS_Length, nb);
    return MXR_ERR_INVALID_ARG;
  }
  if (hipSetDevice(h->c.device) != hipSuccess) {
    set_error("hipSetDevice(%d) failed", h->c.device);
    return MXR_ERR_HIP;
  }
  return run_power_spectrum(&h->c, delta_T, mem, BoxSize, Hubble_h, PS_k, PS_data, PS_error, nb);
}

int mxr_grid_file_read_info(const char* fname, mxr_grid_file_info* info)
{
  if (!fname || !info) {
    set_error("bad grid_file_read_info arguments");
    return MXR_ERR_INVALID_ARG;
  }
  return read_grid_file_info(fname, info, nullptr);
}

static int grid_args_ok(mxr_ctx* h, int property, const mxr_grid_params* p, const float* slab, int mem_out)
{
  if (!h || !p || !slab || property < MXR_GRID_DENSITY || property > MXR_GRID_Z_VELOCITY || !(p->BoxSize > 0) ||
      (mem_out != MXR_MEM_HOST && mem_out != MXR_MEM_DEVICE) ||
      (property == MXR_GRID_DENSITY && (p->NPart <= 0 || !(p->PartMass > 0) || !(p->Hubble_h > 0)))) {
    set_error("bad grid arguments");
    return 0;
  }
  return 1;
}

}  // extern "C" (reopened below)

// the output grid on the device (staged for a host grid), then the work, then the copy back
template <class F>
static int with_device_slab(Ctx& c, float* slab, int mem_out, F&& work)
{
  if (mem_out == MXR_MEM_DEVICE)
    return work(slab);
  const size_t bytes = sizeof(float) * 2 * n_complex(c.N);
  float* d = nullptr;
  MXR_HIP(hipMalloc(&d, bytes));
  int rc = work(d);
  if (rc == MXR_OK) {
    hipError_t e = hipMemcpyAsync(slab, d, bytes, hipMemcpyDeviceToHost, c.stream);
    if (e == hipSuccess)
      e = hipStreamSynchronize(c.stream);
    if (e != hipSuccess)
      rc = hip_fail(e, "hipMemcpyAsync", __FILE__, __LINE__);
  }
  (void)hipFree(d);
  return rc;
}

extern "C" {

int mxr_prepare_grid(mxr_ctx* h, const float* hi, int n_hi, int mem_in, int property, double file_box,
                     const mxr_grid_params* p, float* slab, int mem_out)
{
  if (!grid_args_ok(h, property, p, slab, mem_out) || !hi || (mem_in != MXR_MEM_HOST && mem_in != MXR_MEM_DEVICE) ||
      !(file_box > 0)) {
    set_error("bad prepare_grid arguments");
    return MXR_ERR_INVALID_ARG;
  }
  if (n_hi < h->c.N || n_hi % h->c.N != 0) {
    set_error("n_cell %d is not a multiple of ReionGridDim %d (at least as large)", n_hi, h->c.N);
    return MXR_ERR_INVALID_ARG;
  }
  if (n_hi != h->c.N && !supported_dim(n_hi)) {
    set_error("smoothing a %d^3 grid is not supported (need one of 16,32,64,96,128,160,256,512,1024)", n_hi);
    return MXR_ERR_UNSUPPORTED;
  }
  if (hipSetDevice(h->c.device) != hipSuccess) {
    set_error("hipSetDevice(%d) failed", h->c.device);
    return MXR_ERR_HIP;
  }
  return with_device_slab(h->c, slab, mem_out, [&](float* d) {
    return run_prepare_grid(&h->c, hi, n_hi, mem_in, property, file_box, p, d);
  });
}

int mxr_read_grid_gbptrees(mxr_ctx* h, const char* fname, int property, const mxr_grid_params* p, float* slab,
                           int mem_out)
{
  if (!fname || !grid_args_ok(h, property, p, slab, mem_out))
    return MXR_ERR_INVALID_ARG;
  if (hipSetDevice(h->c.device) != hipSuccess) {
    set_error("hipSetDevice(%d) failed", h->c.device);
    return MXR_ERR_HIP;
  }
  return with_device_slab(h->c, slab, mem_out,
                          [&](float* d) { return run_read_grid_gbptrees(&h->c, fname, property, p, d); });
}

int mxr_construct_baryon_grids(mxr_ctx* h, const mxr_baryon_params* p, const mxr_galaxies* gal,
                               const mxr_baryon_grids* g, long* n_below_bh_limit)
{
  if (!h || !p || !gal || !g || p->ReionGridDim != h->c.N || !(p->BoxSize > 0) || gal->n < 0 ||
      gal->n >= 2147483647L || (gal->mem != MXR_MEM_HOST && gal->mem != MXR_MEM_DEVICE) ||
      (g->mem != MXR_MEM_HOST && g->mem != MXR_MEM_DEVICE)) {
    set_error("bad construct_baryon_grids arguments");
    return MXR_ERR_INVALID_ARG;
  }
  const int spin = p->Flag_IncludeSpinTemp, bh = p->Flag_BHFeedback, agn = p->Flag_IncludeAGNXray;
  if (agn < MXR_AGN_XRAY_NONE || agn > MXR_AGN_XRAY_SOFT || (spin && p->NstoreSnapshots_Heating < 1) ||
      (spin && !p->Flag_InstantaneousSFR && !(p->sfr_timescale > 0))) {
    set_error("construct_baryon_grids: Flag_IncludeAGNXray %d, NstoreSnapshots_Heating %d, sfr_timescale %g", agn,
              p->NstoreSnapshots_Heating, p->sfr_timescale);
    return MXR_ERR_INVALID_ARG;
  }
  const bool hard = spin && (agn == MXR_AGN_XRAY_BOTH || agn == MXR_AGN_XRAY_HARD);
  const bool soft = spin && (agn == MXR_AGN_XRAY_BOTH || agn == MXR_AGN_XRAY_SOFT);
  const bool want_sfr = spin && p->Flag_InstantaneousSFR, want_gsm = spin && !p->Flag_InstantaneousSFR;
  const bool gal_ok = gal->n == 0 ||
                      (gal->Pos && gal->Type && gal->FescWeightedGSM && gal->FescWeightedSfr &&
                       (!want_sfr || gal->Sfr) && (!want_gsm || gal->GrossStellarMass) &&
                       (!(spin && p->Flag_XrayMetallicity) || (gal->ColdGas && gal->MetalsColdGas)) &&
                       (!(bh || hard || soft) || gal->BlackHoleMass) &&
                       (!bh || (gal->EffectiveBHM && gal->EffectiveBHAR)) && (!hard || gal->BHXrayEmissivity_hard) &&
                       (!soft || gal->BHXrayEmissivity_soft));
  const bool grid_ok = g->stars && g->weighted_sfr && (!bh || (g->effective_bhm && g->effective_bhar)) &&
                       (!spin || (g->sfr && g->sfr_histories && g->xray_luminosity && g->xray_luminosity_histories)) &&
                       (!hard || (g->BHXrayEmissivity_hard && g->bh_xray_histories_hard)) &&
                       (!soft || (g->BHXrayEmissivity_soft && g->bh_xray_histories_soft));
  if (!gal_ok || !grid_ok) {
    set_error("construct_baryon_grids: a galaxy array or grid these flags need is NULL");
    return MXR_ERR_INVALID_ARG;
  }
  if (spin && g->history_head && (*g->history_head < 0 || *g->history_head >= p->NstoreSnapshots_Heating)) {
    set_error("construct_baryon_grids: history_head %d is not a slot of %d", *g->history_head,
              p->NstoreSnapshots_Heating);
    return MXR_ERR_INVALID_ARG;
  }
  if (hipSetDevice(h->c.device) != hipSuccess) {
    set_error("hipSetDevice(%d) failed", h->c.device);
    return MXR_ERR_HIP;
  }
  return run_construct_baryon_grids(&h->c, p, gal, g, n_below_bh_limit);
}

int mxr_ps_counts(mxr_ctx* h, double BoxSize, double Hubble_h, int PS_Length, double* counts)
{
  if (!h || !counts || !(BoxSize > 0) || !(Hubble_h > 0)) {
    set_error("bad ps_counts arguments");
    return MXR_ERR_INVALID_ARG;
  }
  const int nb = ps_length(h->c.N, BoxSize, Hubble_h);
  if (PS_Length != nb) {
    set_error("PS_Length %d differs from the %d bins of this grid and box", PS_Length, nb);
    return MXR_ERR_INVALID_ARG;
  }
  if (hipSetDevice(h->c.device) != hipSuccess) {
    set_error("hipSetDevice(%d) failed", h->c.device);
    return MXR_ERR_HIP;
  }
  return run_ps_counts(&h->c, BoxSize, Hubble_h, nb, counts);
}

}  // extern "C"
