// mxr_baryon.hip -- galaxies to grids (construct_baryon_grids, src/core/reionization.c:2025-2591)
// on one gfx950 GPU, for one rank holding the whole grid.
//
// The reference accumulates each cell's galaxies into a float buffer in galaxy order.  Float
// sums depend on that order, so instead of atomics: each live galaxy gets its NGP cell as a key
// (pos_to_ngp, misc_tools.c:207-217), a stable radix sort groups the galaxies by cell keeping
// their order, and one thread per occupied cell sums its run of galaxies in order with the
// reference's arithmetic and writes the finalised values.  Cells without galaxies keep the zero
// the preparation kernel wrote, which is what the reference's finalisation gives them.
#include <hipcub/hipcub.hpp>

#include <cmath>

#include "mxr_passes.hpp"

namespace mxr {

namespace {

constexpr double SOLAR_MASS_G = 1.989e33;     // src/meraxes.h:22
constexpr double SEC_PER_YEAR_S = 3.155e7;    // src/meraxes.h:34
constexpr double XRAY_LUM_UNIT = 1.0e40;      // src/meraxes.h:36
constexpr double REL_TOL_BAND = 1e-5;         // src/meraxes.h REL_TOL

struct BgArgs {
  int N;
  long n;
  double box;
  unsigned sentinel;  // N^3: the key of dead galaxies, sorted last
  int spin, bh, inst, hard, soft, metal;
  double bh_limit, sfr_timescale;
  double mass_conv, time_conv;  // UnitMass_in_g / SOLAR_MASS, SEC_PER_YEAR / UnitTime_in_s
  double lx_pow;                // pow(10, log10(LXrayGal)) (host libm), or 0 when LXrayGal <= 0
  double band_fraction;         // galaxy_xray_band_fraction (host libm)
  double XrayLogLNorm, XrayMetallicitySlope, XrayMetallicitySolar, XrayOxygenAbundanceFactor;
  const float* pos;
  const int* type;
  const double *gsm_w, *sfr_w, *sfr, *gsm, *cold, *metals, *bhm, *ebhm, *ebhar, *xh, *xs;
  float *stars, *wsfr, *g_bhm, *g_bhar, *g_sfr, *h_sfr, *g_x, *h_x, *g_h, *h_h, *g_s, *h_s;
  int nstore;
  int* err;
  unsigned long long* below;
};

// misc_tools.c:207-217; -1 for a position the reference's assert rejects
__device__ __forceinline__ int pos_to_ngp(float x, double side, int nx)
{
  const double v = rint((double)x / side * (double)nx);
  if (!(v > -1.0 && v < 2147483647.0))
    return -1;
  int ind = (int)v;
  if (ind > nx - 1)
    ind = 0;
  return ind;
}

__global__ void k_bg_keys(BgArgs a, unsigned* __restrict__ keys, unsigned* __restrict__ vals)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n)
    return;
  unsigned key = a.sentinel;
  if (a.type[i] <= 2) {  // Type > 2: dead galaxies are skipped (reionization.c:2226-2230)
    const int ix = pos_to_ngp(a.pos[3 * i + 0], a.box, a.N);
    const int iy = pos_to_ngp(a.pos[3 * i + 1], a.box, a.N);
    const int iz = pos_to_ngp(a.pos[3 * i + 2], a.box, a.N);
    if (ix < 0 || iy < 0 || iz < 0)
      atomicOr(a.err, 1);
    else
      key = ((unsigned)ix * a.N + (unsigned)iy) * a.N + (unsigned)iz;
  }
  keys[i] = key;
  vals[i] = (unsigned)i;
}

// reionization.c:2093-2141: clear the grids and shift each history by one slot; slot 0 keeps its
// padding (the reference's finalisation writes its real cells only), and so do the BH grids
__global__ void k_bg_prep(BgArgs a)
{
  const int N = a.N, H2 = 2 * (N / 2 + 1);
  const long npad = (long)N * N * H2;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < npad; p += (long)gridDim.x * blockDim.x) {
    const bool real = (int)(p % H2) < N;
    a.stars[p] = 0.f;
    a.wsfr[p] = 0.f;
    if (a.bh && real) {
      a.g_bhm[p] = 0.f;
      a.g_bhar[p] = 0.f;
    }
    if (a.spin) {
      // slot s moves to s + 1, from the top down, 8 slots per step with every load of a step
      // issued before its stores (the next step's slots lie below the ones just written)
      auto shift = [&](float* __restrict__ h) {
        for (int s = a.nstore - 2; s >= 0; s -= 8) {
          float t[8];
#pragma unroll
          for (int k = 0; k < 8; k++)
            if (k <= s)
              t[k] = h[(long)(s - k) * npad + p];
#pragma unroll
          for (int k = 0; k < 8; k++)
            if (k <= s)
              h[(long)(s - k + 1) * npad + p] = t[k];
        }
        if (real)
          h[p] = 0.f;
      };
      a.g_sfr[p] = 0.f;
      shift(a.h_sfr);
      a.g_x[p] = 0.f;
      shift(a.h_x);
      if (a.hard) {
        a.g_h[p] = 0.f;
        shift(a.h_h);
      }
      if (a.soft) {
        a.g_s[p] = 0.f;
        shift(a.h_s);
      }
    }
  }
}

// calculate_galaxy_xray_luminosity (reionization.c:1982-2023) without scatter
__device__ __forceinline__ double xray_luminosity(const BgArgs& a, unsigned g)
{
#pragma clang fp contract(off)
  double source_sfr = a.inst ? a.sfr[g] : a.gsm[g] / a.sfr_timescale;
  source_sfr *= a.mass_conv;
  source_sfr *= a.time_conv;
  if (source_sfr <= 0.0)
    return 0.0;
  double lx = a.lx_pow, bf = 1.0;
  if (a.metal) {
    // calc_metallicity (misc_tools.c:45-57)
    const double total_gas = a.cold[g], metals = a.metals[g];
    double Z = (total_gas > 0 && metals > 0) ? metals / total_gas : 0.0;
    Z = Z < 0 ? 0 : (Z > 1 ? 1 : Z);
    const double z_solar = Z / a.XrayMetallicitySolar;
    const double oxygen_abundance = a.XrayOxygenAbundanceFactor * z_solar;
    lx = pow(10.0, a.XrayLogLNorm + a.XrayMetallicitySlope * oxygen_abundance);
    bf = a.band_fraction;
  } else if (lx <= 0.0) {
    return 0.0;
  }
  return lx * bf * source_sfr / XRAY_LUM_UNIT;
}

// one galaxy's contributions, one 64-byte line: the terms the reference adds (0.0 where it
// skips the galaxy, since (float)((double)x + 0.0) == x), and whether it counts below the limit
struct __align__(64) GalRec {
  double st, ws, bm, br, sf, hh, hs;
  float xr;
  int below;
};

// the galaxy fields read once, coalesced, in galaxy order
__global__ void k_bg_pack(BgArgs a, GalRec* __restrict__ rec)
{
#pragma clang fp contract(off)
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n)
    return;
  GalRec r;
  r.st = a.gsm_w[i];
  r.ws = a.sfr_w[i];
  const bool ok = (a.bh || a.hard || a.soft) && a.bhm[i] >= a.bh_limit;
  r.bm = a.bh && ok ? a.ebhm[i] : 0.0;
  r.br = a.bh && ok ? a.ebhar[i] : 0.0;
  r.below = a.bh && !ok;
  r.xr = a.spin ? (float)xray_luminosity(a, (unsigned)i) : 0.f;
  r.sf = a.spin ? (a.inst ? a.sfr[i] : a.gsm[i]) : 0.0;
  r.hh = a.hard && ok ? a.xh[i] : 0.0;
  r.hs = a.soft && ok ? a.xs[i] : 0.0;
  rec[i] = r;
}

// the finalisation of one cell's sums into the padded grids (:2424-2569)
__device__ __forceinline__ void deposit_cell(const BgArgs& a, unsigned key, float st, float ws, float bm, float br,
                                             float xr, float sf, float hh, float hs)
{
#pragma clang fp contract(off)
  const int N = a.N;
  const long ip = (long)(key / N) * (2 * (N / 2 + 1)) + key % N;
  a.stars[ip] = st < 0 ? 0.f : st;
  {
    double v = (double)ws;
    if (v < 0)
      v = 0.0;
    a.wsfr[ip] = (float)v;
  }
  if (a.bh) {
    a.g_bhm[ip] = bm < 0 ? 0.f : bm;
    a.g_bhar[ip] = br < 0 ? 0.f : br;
  }
  if (a.spin) {
    a.g_x[ip] = xr;
    a.h_x[ip] = xr;
    double v = (double)sf;
    if (v < 0)
      v = 0.0;
    if (!a.inst)
      v /= a.sfr_timescale;
    a.g_sfr[ip] = (float)v;
    a.h_sfr[ip] = (float)v;
    if (a.hard) {
      const float h = hh < 0 ? 0.f : hh;
      a.g_h[ip] = h;
      a.h_h[ip] = h;
    }
    if (a.soft) {
      const float s = hs < 0 ? 0.f : hs;
      a.g_s[ip] = s;
      a.h_s[ip] = s;
    }
  }
}

// one thread per run of equal keys: the cell's galaxies in their given order (:2219-2411),
// then the finalisation of every property (:2424-2569); the grid-stride loop keeps the block
// count small so that the below-limit count needs one global atomic per block
__global__ void __launch_bounds__(256) k_bg_deposit(BgArgs a, const unsigned* __restrict__ keys,
                                                    const unsigned* __restrict__ vals,
                                                    const GalRec* __restrict__ rec)
{
#pragma clang fp contract(off)
  __shared__ unsigned long long blk_below;
  if (threadIdx.x == 0)
    blk_below = 0;
  __syncthreads();
  unsigned long long below = 0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += (long)gridDim.x * blockDim.x) {
    const unsigned key = keys[i];
    if (key >= a.sentinel || (i > 0 && keys[i - 1] == key))
      continue;
    float st = 0.f, ws = 0.f, bm = 0.f, br = 0.f, xr = 0.f, sf = 0.f, hh = 0.f, hs = 0.f;
    // the run in chunks of U: each chunk's keys and indices, then its records, are loaded
    // together (one round trip each); entries past the run load the chunk's first record again
    // (a cache hit) and add 0.0, which leaves every sum unchanged (no sum is -0.0)
    constexpr int U = 8;
    bool more = true;
    for (long j = i; more; j += U) {
      unsigned kk[U], vv[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const long jj = j + u < a.n ? j + u : a.n - 1;
        kk[u] = keys[jj];
        vv[u] = vals[jj];
      }
      bool in_run[U];
#pragma unroll
      for (int u = 0; u < U; u++)
        in_run[u] = j + u < a.n && kk[u] == key;
      GalRec rr[U];
#pragma unroll
      for (int u = 0; u < U; u++)
        rr[u] = rec[in_run[u] ? vv[u] : vv[0]];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const bool m = in_run[u];
        const GalRec& r = rr[u];
        st = (float)((double)st + (m ? r.st : 0.0));
        ws = (float)((double)ws + (m ? r.ws : 0.0));
        bm = (float)((double)bm + (m ? r.bm : 0.0));
        br = (float)((double)br + (m ? r.br : 0.0));
        below += m ? (unsigned)r.below : 0u;
        xr = xr + (m ? r.xr : 0.0f);
        sf = (float)((double)sf + (m ? r.sf : 0.0));
        hh = (float)((double)hh + (m ? r.hh : 0.0));
        hs = (float)((double)hs + (m ? r.hs : 0.0));
      }
      more = in_run[U - 1];
    }
    deposit_cell(a, key, st, ws, bm, br, xr, sf, hh, hs);
  }
  if (below)
    atomicAdd(&blk_below, below);
  __syncthreads();
  if (threadIdx.x == 0 && blk_below)
    atomicAdd(a.below, blk_below);
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

int run_construct_baryon_grids(Ctx* c, const mxr_baryon_params* p, const mxr_galaxies* gal,
                               const mxr_baryon_grids* grids, long* n_below)
{
  const int N = p->ReionGridDim;
  const long n = gal->n;
  const long npad = 2 * n_complex(N), nstore = p->Flag_IncludeSpinTemp ? p->NstoreSnapshots_Heating : 0;
  BgArgs a{};
  a.N = N;
  a.n = n;
  a.box = p->BoxSize;
  a.sentinel = (unsigned)((long)N * N * N);
  a.spin = p->Flag_IncludeSpinTemp != 0;
  a.bh = p->Flag_BHFeedback != 0;
  a.inst = p->Flag_InstantaneousSFR != 0;
  a.hard = a.spin && (p->Flag_IncludeAGNXray == MXR_AGN_XRAY_BOTH || p->Flag_IncludeAGNXray == MXR_AGN_XRAY_HARD);
  a.soft = a.spin && (p->Flag_IncludeAGNXray == MXR_AGN_XRAY_BOTH || p->Flag_IncludeAGNXray == MXR_AGN_XRAY_SOFT);
  a.metal = p->Flag_XrayMetallicity != 0;
  a.bh_limit = p->BlackHoleMassLimitReion;
  a.sfr_timescale = p->sfr_timescale;
  a.mass_conv = p->UnitMass_in_g / SOLAR_MASS_G;
  a.time_conv = SEC_PER_YEAR_S / p->UnitTime_in_s;
  a.lx_pow = p->LXrayGal > 0.0 ? std::pow(10.0, std::log10(p->LXrayGal)) : 0.0;
  {
    // galaxy_xray_band_fraction (reionization.c:1967-1980)
    const double e_min = p->NuXrayThreshold, e_max = p->NuXraySoftCut, alpha = p->SpecIndexXrayGal;
    if (std::fabs(alpha - 1.0) < REL_TOL_BAND) {
      a.band_fraction = std::log(e_max / e_min) / std::log(2000.0 / 500.0);
    } else {
      const double ex = 1.0 - alpha;
      a.band_fraction =
        (std::pow(e_max, ex) - std::pow(e_min, ex)) / (std::pow(2000.0, ex) - std::pow(500.0, ex));
    }
  }
  a.XrayLogLNorm = p->XrayLogLNorm;
  a.XrayMetallicitySlope = p->XrayMetallicitySlope;
  a.XrayMetallicitySolar = p->XrayMetallicitySolar;
  a.XrayOxygenAbundanceFactor = p->XrayOxygenAbundanceFactor;
  a.nstore = (int)nstore;
  // ring histories: the new grid's slot, written in place of the shift
  int ring_slot = 0;
  if (grids->history_head && a.spin) {
    ring_slot = (int)((*grids->history_head + nstore - 1) % nstore);
    a.nstore = 1;
  }

  // device workspace: keys and values twice, the sort's temporary storage, the error flag and
  // the below-limit counter, then (host galaxies) the staged galaxy arrays
  size_t sort_bytes = 0;
  int end_bit = 1;
  while (end_bit < 32 && ((unsigned long long)a.sentinel >> end_bit) != 0)
    end_bit++;
  if (n > 0)
    MXR_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (const unsigned*)nullptr, (unsigned*)nullptr,
                                               (const unsigned*)nullptr, (unsigned*)nullptr, (int)n, 0, end_bit,
                                               c->stream));
  const size_t kb = align256(sizeof(unsigned) * (size_t)n);
  const bool host_gal = gal->mem == MXR_MEM_HOST;
  const size_t stage = host_gal ? align256(sizeof(float) * 3 * n) + align256(sizeof(int) * n) +
                                    11 * align256(sizeof(double) * n)
                                : 0;
  const size_t rb = align256(sizeof(GalRec) * (size_t)n);
  const size_t need = 4 * kb + align256(sort_bytes) + 256 + rb + stage;
  if (c->bg_ws_cap < need) {
    (void)hipFree(c->bg_ws);
    c->bg_ws = nullptr;
    c->bg_ws_cap = 0;
    MXR_HIP(hipMalloc(&c->bg_ws, need));
    c->bg_ws_cap = need;
  }
  char* w = static_cast<char*>(c->bg_ws);
  unsigned* keys_in = reinterpret_cast<unsigned*>(w);
  unsigned* keys_out = reinterpret_cast<unsigned*>(w + kb);
  unsigned* vals_in = reinterpret_cast<unsigned*>(w + 2 * kb);
  unsigned* vals_out = reinterpret_cast<unsigned*>(w + 3 * kb);
  void* sort_tmp = w + 4 * kb;
  char* flags = w + 4 * kb + align256(sort_bytes);
  a.err = reinterpret_cast<int*>(flags);
  a.below = reinterpret_cast<unsigned long long*>(flags + 128);
  GalRec* recs = reinterpret_cast<GalRec*>(flags + 256);
  char* st = flags + 256 + rb;

  // galaxy arrays on the device
  auto dev_in = [&](const void* src, size_t bytes, const void** dst) -> int {
    if (!host_gal || !src) {
      *dst = src;
      return MXR_OK;
    }
    MXR_HIP(hipMemcpyAsync(st, src, bytes, hipMemcpyHostToDevice, c->stream));
    *dst = st;
    st += align256(bytes);
    return MXR_OK;
  };
  const void* q = nullptr;
  RC(dev_in(gal->Pos, sizeof(float) * 3 * n, &q));
  a.pos = static_cast<const float*>(q);
  RC(dev_in(gal->Type, sizeof(int) * n, &q));
  a.type = static_cast<const int*>(q);
  const double* const* dsrc[11] = {&gal->FescWeightedGSM, &gal->FescWeightedSfr, &gal->Sfr,
                                   &gal->GrossStellarMass, &gal->ColdGas, &gal->MetalsColdGas,
                                   &gal->BlackHoleMass, &gal->EffectiveBHM, &gal->EffectiveBHAR,
                                   &gal->BHXrayEmissivity_hard, &gal->BHXrayEmissivity_soft};
  const double** ddst[11] = {&a.gsm_w, &a.sfr_w, &a.sfr, &a.gsm, &a.cold, &a.metals,
                             &a.bhm, &a.ebhm, &a.ebhar, &a.xh, &a.xs};
  for (int k = 0; k < 11; k++) {
    RC(dev_in(*dsrc[k], sizeof(double) * n, &q));
    *ddst[k] = static_cast<const double*>(q);
  }

  // output grids on the device (host grids are staged in and out: the histories and the BH
  // grids' padding keep their contents)
  struct G {
    float* user;
    float** arg;
    long len;
  };
  G gl[12] = {{grids->stars, &a.stars, npad},
              {grids->weighted_sfr, &a.wsfr, npad},
              {a.bh ? grids->effective_bhm : nullptr, &a.g_bhm, npad},
              {a.bh ? grids->effective_bhar : nullptr, &a.g_bhar, npad},
              {a.spin ? grids->sfr : nullptr, &a.g_sfr, npad},
              {a.spin ? grids->sfr_histories : nullptr, &a.h_sfr, npad * nstore},
              {a.spin ? grids->xray_luminosity : nullptr, &a.g_x, npad},
              {a.spin ? grids->xray_luminosity_histories : nullptr, &a.h_x, npad * nstore},
              {a.hard ? grids->BHXrayEmissivity_hard : nullptr, &a.g_h, npad},
              {a.hard ? grids->bh_xray_histories_hard : nullptr, &a.h_h, npad * nstore},
              {a.soft ? grids->BHXrayEmissivity_soft : nullptr, &a.g_s, npad},
              {a.soft ? grids->bh_xray_histories_soft : nullptr, &a.h_s, npad * nstore}};
  const bool host_grids = grids->mem == MXR_MEM_HOST;
  float* tmp[12] = {};
  int rc = MXR_OK;
  for (int k = 0; k < 12 && rc == MXR_OK; k++) {
    *gl[k].arg = gl[k].user;
    if (!host_grids || !gl[k].user)
      continue;
    hipError_t e = hipMalloc(&tmp[k], sizeof(float) * gl[k].len);
    if (e == hipSuccess)
      e = hipMemcpyAsync(tmp[k], gl[k].user, sizeof(float) * gl[k].len, hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess)
      rc = hip_fail(e, "staging a baryon grid", __FILE__, __LINE__);
    *gl[k].arg = tmp[k];
  }

  if (ring_slot) {
    a.h_sfr += (long)ring_slot * npad;
    a.h_x += (long)ring_slot * npad;
    if (a.hard)
      a.h_h += (long)ring_slot * npad;
    if (a.soft)
      a.h_s += (long)ring_slot * npad;
  }
  int err = 0;
  if (rc == MXR_OK) {
    hipError_t e = hipMemsetAsync(flags, 0, 256, c->stream);
    if (e == hipSuccess && n > 0) {
      hipLaunchKernelGGL(k_bg_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, a, keys_in, vals_in);
      e = hipGetLastError();
    }
    if (e == hipSuccess)
      e = hipMemcpyAsync(&err, a.err, sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
      e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess)
      rc = hip_fail(e, "galaxy keys", __FILE__, __LINE__);
    else if (err) {
      set_error("a galaxy position maps below cell 0 (pos_to_ngp); the reference asserts");
      rc = MXR_ERR_INVALID_ARG;
    }
  }
  if (rc == MXR_OK && n > 0) {
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(sort_tmp, sort_bytes, keys_in, keys_out, vals_in, vals_out,
                                                      (int)n, 0, end_bit, c->stream);
    if (e != hipSuccess)
      rc = hip_fail(e, "hipcub::DeviceRadixSort::SortPairs", __FILE__, __LINE__);
  }
  if (rc == MXR_OK) {
    hipLaunchKernelGGL(k_bg_prep, dim3(ew_blocks(npad)), dim3(256), 0, c->stream, a);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && n > 0) {
      hipLaunchKernelGGL(k_bg_pack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, a, recs);
      e = hipGetLastError();
    }
    if (e == hipSuccess && n > 0) {
      hipLaunchKernelGGL(k_bg_deposit, dim3(ew_blocks(n)), dim3(256), 0, c->stream, a, keys_out, vals_out, recs);
      e = hipGetLastError();
    }
    unsigned long long below = 0;
    if (e == hipSuccess)
      e = hipMemcpyAsync(&below, a.below, sizeof(below), hipMemcpyDeviceToHost, c->stream);
    for (int k = 0; k < 12 && e == hipSuccess; k++)
      if (tmp[k])
        e = hipMemcpyAsync(gl[k].user, tmp[k], sizeof(float) * gl[k].len, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
      e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess)
      rc = hip_fail(e, "construct_baryon_grids", __FILE__, __LINE__);
    else {
      if (n_below)
        *n_below = (long)below;
      if (grids->history_head && a.spin)
        *grids->history_head = ring_slot;
    }
  }
  for (int k = 0; k < 12; k++)
    (void)hipFree(tmp[k]);
  return rc;
}

}  // namespace mxr
