// mxr_baryon.hip -- galaxies to grids (construct_baryon_grids, src/core/reionization.c:2025-2591)
// on one gfx950 GPU, for one rank holding the whole grid.
//
// The reference accumulates each cell's galaxies into a float buffer in galaxy order.  Float
// sums depend on that order, so instead of atomics: each live galaxy gets its NGP cell as a key
// (pos_to_ngp, misc_tools.c:207-217), a stable radix sort groups the galaxies by cell keeping
// their order, a run-length encoding lists the occupied cells, and one thread per occupied cell
// sums its run of galaxies in order with the reference's arithmetic and finalises the cell's
// values.  One wavefront per grid row then writes the row of every grid once, coalesced: the
// occupied cells' values and zero elsewhere, which is what the reference's clear and
// finalisation give a cell without galaxies.  The histories move up by device copies meanwhile,
// each history array on a stream of its own.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_run_length_encode.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
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

// one occupied cell's finalised values (:2424-2569): the clamps and the sfr conversion
struct __align__(32) CellVal {
  float st, ws, bm, br, xr, sf, hh, hs;
};

// one thread per occupied cell (run of equal keys): the cell's galaxies in their given order
// (:2219-2411), then the finalisation of every property
__global__ void __launch_bounds__(256) k_bg_runs(BgArgs a, const unsigned* __restrict__ run_keys,
                                                 const unsigned* __restrict__ run_len,
                                                 const unsigned* __restrict__ run_off, const unsigned* __restrict__ nrun,
                                                 const unsigned* __restrict__ vals, const GalRec* __restrict__ rec,
                                                 CellVal* __restrict__ out)
{
#pragma clang fp contract(off)
  __shared__ unsigned long long blk_below;
  if (threadIdx.x == 0)
    blk_below = 0;
  __syncthreads();
  unsigned long long below = 0;
  const long nr = *nrun;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nr; i += (long)gridDim.x * blockDim.x) {
    if (run_keys[i] >= a.sentinel)  // the dead galaxies' run (sorted last)
      continue;
    const long off = run_off[i], len = run_len[i];
    float st = 0.f, ws = 0.f, bm = 0.f, br = 0.f, xr = 0.f, sf = 0.f, hh = 0.f, hs = 0.f;
    // the run in chunks of U: a chunk's indices, then its records, are loaded together (one
    // round trip each); slots past the run reload the run's last record (a cache hit) and add
    // 0.0, which leaves every sum unchanged (no sum is -0.0)
    constexpr int U = 8;
    for (long j = 0; j < len; j += U) {
      unsigned vv[U];
#pragma unroll
      for (int u = 0; u < U; u++)
        vv[u] = vals[off + (j + u < len ? j + u : len - 1)];
      GalRec rr[U];
#pragma unroll
      for (int u = 0; u < U; u++)
        rr[u] = rec[vv[u]];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const bool m = j + u < len;
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
    }
    CellVal v;
    v.st = st < 0 ? 0.f : st;
    v.ws = (float)((double)ws < 0 ? 0.0 : (double)ws);
    v.bm = bm < 0 ? 0.f : bm;
    v.br = br < 0 ? 0.f : br;
    v.xr = xr;
    double s = (double)sf;
    if (s < 0)
      s = 0.0;
    if (!a.inst)
      s /= a.sfr_timescale;
    v.sf = (float)s;
    v.hh = hh < 0 ? 0.f : hh;
    v.hs = hs < 0 ? 0.f : hs;
    out[i] = v;
  }
  if (below)
    atomicAdd(&blk_below, below);
  __syncthreads();
  if (threadIdx.x == 0 && blk_below)
    atomicAdd(a.below, blk_below);
}

// first occupied-cell index of each grid row (and of the end, row N^2): lower bound of the row's
// first key in the sorted keys of the occupied cells
__global__ void k_bg_row_start(const unsigned* __restrict__ run_keys, const unsigned* __restrict__ nrun, int N,
                               unsigned* __restrict__ row_run)
{
  const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r > (long)N * N)
    return;
  const unsigned target = (unsigned)(r * N);
  unsigned lo = 0, hi = *nrun;
  while (lo < hi) {
    const unsigned mid = lo + (hi - lo) / 2;
    if (run_keys[mid] < target)
      lo = mid + 1;
    else
      hi = mid;
  }
  row_run[r] = lo;
}

// One wavefront per grid row (rows of at most 64 x DENSE_CPL cells): the row's occupied cells
// are placed by an LDS map from cell to lane, their values pulled from that lane, and every grid's
// row is written once with coalesced stores: the values, zero elsewhere.  The padding of stars,
// weighted_sfr and the current spin-temperature grids is zeroed as the reference's clear does;
// the BH grids' and the new history slot's padding are left as they are.
constexpr int DENSE_CPL = 8;
__global__ void __launch_bounds__(256) k_bg_dense(BgArgs a, const unsigned* __restrict__ run_keys,
                                                  const CellVal* __restrict__ vals,
                                                  const unsigned* __restrict__ row_run)
{
  __shared__ short map[4][64 * DENSE_CPL];
  const int N = a.N, H2 = 2 * (N / 2 + 1);
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  const long row = (long)blockIdx.x * 4 + w;
  if (row >= (long)N * N)
    return;  // no block barrier below: each wavefront works on its own map
  const int rs = (int)row_run[row], re = (int)row_run[row + 1];
  const unsigned key0 = (unsigned)(row * N);
  float v[8][DENSE_CPL];
#pragma unroll
  for (int k = 0; k < 8; k++)
#pragma unroll
    for (int u = 0; u < DENSE_CPL; u++)
      v[k][u] = 0.f;
  for (int c0 = rs; c0 < re; c0 += 64) {
#pragma unroll
    for (int u = 0; u < DENSE_CPL; u++)
      map[w][lane + 64 * u] = -1;
    __builtin_amdgcn_wave_barrier();
    const bool have = c0 + lane < re;
    const int src = have ? c0 + lane : c0;
    const unsigned key = run_keys[src];
    const CellVal cv = vals[src];
    if (have)
      map[w][key - key0] = (short)lane;
    __builtin_amdgcn_wave_barrier();
    const float f[8] = {cv.st, cv.ws, cv.bm, cv.br, cv.xr, cv.sf, cv.hh, cv.hs};
#pragma unroll
    for (int u = 0; u < DENSE_CPL; u++) {
      const int j = map[w][lane + 64 * u];
      const int from = j < 0 ? lane : j;
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const float x = __shfl(f[k], from, 64);
        v[k][u] = j < 0 ? v[k][u] : x;
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  const long base = row * H2;
#pragma unroll
  for (int u = 0; u < DENSE_CPL; u++) {
    const int z = lane + 64 * u;
    if (z >= N)
      continue;
    const long p = base + z;
    a.stars[p] = v[0][u];
    a.wsfr[p] = v[1][u];
    if (a.bh) {
      a.g_bhm[p] = v[2][u];
      a.g_bhar[p] = v[3][u];
    }
    if (a.spin) {
      a.g_x[p] = v[4][u];
      a.h_x[p] = v[4][u];
      a.g_sfr[p] = v[5][u];
      a.h_sfr[p] = v[5][u];
      if (a.hard) {
        a.g_h[p] = v[6][u];
        a.h_h[p] = v[6][u];
      }
      if (a.soft) {
        a.g_s[p] = v[7][u];
        a.h_s[p] = v[7][u];
      }
    }
  }
  if (lane < H2 - N) {
    const long p = base + N + lane;
    a.stars[p] = 0.f;
    a.wsfr[p] = 0.f;
    if (a.spin) {
      a.g_x[p] = 0.f;
      a.g_sfr[p] = 0.f;
      if (a.hard)
        a.g_h[p] = 0.f;
      if (a.soft)
        a.g_s[p] = 0.f;
    }
  }
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

int run_construct_baryon_grids(Ctx* c, const mxr_baryon_params* p, const mxr_galaxies* gal,
                               const mxr_baryon_grids* grids, long* n_below)
{
  const int N = p->ReionGridDim;
  const long n = gal->n;
  if (N > 64 * DENSE_CPL) {
    set_error("construct_baryon_grids: ReionGridDim %d exceeds %d", N, 64 * DENSE_CPL);
    return MXR_ERR_UNSUPPORTED;
  }
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

  // device workspace: keys and values twice (after the sort, the first pair holds the occupied
  // cells' keys and run lengths), the run offsets, the library calls' temporary storage, the
  // error flag, the below-limit counter and the run count, the galaxy records, the cells'
  // values, the rows' first cells, then (host galaxies) the staged galaxy arrays
  size_t tmp_bytes = 0;
  int end_bit = 1;
  while (end_bit < 32 && ((unsigned long long)a.sentinel >> end_bit) != 0)
    end_bit++;
  if (n > 0) {
    size_t b = 0;
    MXR_HIP(rocprim::radix_sort_pairs(nullptr, b, (const unsigned*)nullptr, (unsigned*)nullptr,
                                      (const unsigned*)nullptr, (unsigned*)nullptr, (unsigned)n, 0u,
                                      (unsigned)end_bit, c->stream));
    tmp_bytes = std::max(tmp_bytes, b);
    MXR_HIP(rocprim::run_length_encode(nullptr, b, (const unsigned*)nullptr, (unsigned)n, (unsigned*)nullptr,
                                       (unsigned*)nullptr, (unsigned*)nullptr, c->stream));
    tmp_bytes = std::max(tmp_bytes, b);
    MXR_HIP(rocprim::exclusive_scan(nullptr, b, (const unsigned*)nullptr, (unsigned*)nullptr, 0u, (size_t)n,
                                    rocprim::plus<unsigned>(), c->stream));
    tmp_bytes = std::max(tmp_bytes, b);
  }
  const size_t kb = align256(sizeof(unsigned) * (size_t)n);
  const bool host_gal = gal->mem == MXR_MEM_HOST;
  const size_t stage = host_gal ? align256(sizeof(float) * 3 * n) + align256(sizeof(int) * n) +
                                    11 * align256(sizeof(double) * n)
                                : 0;
  const size_t rb = align256(sizeof(GalRec) * (size_t)n);
  const size_t cvb = align256(sizeof(CellVal) * (size_t)n);
  const size_t rrb = align256(sizeof(unsigned) * ((size_t)N * N + 1));
  const size_t need = 5 * kb + align256(tmp_bytes) + 256 + rb + cvb + rrb + stage;
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
  unsigned* run_keys = keys_in;  // free once sorted
  unsigned* run_len = vals_in;
  unsigned* run_off = reinterpret_cast<unsigned*>(w + 4 * kb);
  void* lib_tmp = w + 5 * kb;
  char* flags = w + 5 * kb + align256(tmp_bytes);
  a.err = reinterpret_cast<int*>(flags);
  a.below = reinterpret_cast<unsigned long long*>(flags + 128);
  unsigned* nrun = reinterpret_cast<unsigned*>(flags + 192);  // stays 0 without galaxies
  GalRec* recs = reinterpret_cast<GalRec*>(flags + 256);
  CellVal* cell_vals = reinterpret_cast<CellVal*>(flags + 256 + rb);
  unsigned* row_run = reinterpret_cast<unsigned*>(flags + 256 + rb + cvb);
  char* st = flags + 256 + rb + cvb + rrb;

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
  // reionization.c:2093-2141, the history part: each history moves up by one slot, whole
  // slots (padding included) from the top down, one device copy per slot (which the memory
  // system serves faster than all slots moved together by one kernel).  Each history array's
  // copies run on a stream of its own, beside each other and beside the keys, sort and cell sums,
  // which do not touch the grids; the row writer waits for all of them.  They start once the positions are known to be valid, so a rejected
  // call leaves the grids as they were.  Slot 0's real cells are then written by the row writer, so its
  // padding keeps the old slot 0's (the reference's finalisation writes real cells only).
  int shifting = 0;  // streams the shift was started on
  if (rc == MXR_OK && a.spin && a.nstore > 1) {
    hipError_t e = hipSuccess;
    float* hs[4] = {a.h_sfr, a.h_x, a.hard ? a.h_h : nullptr, a.soft ? a.h_s : nullptr};
    static_assert(BG_STREAMS >= 4, "one stream per history array");
    if (!c->bg_ev_in)
      e = hipEventCreateWithFlags(&c->bg_ev_in, hipEventDisableTiming);
    if (e == hipSuccess)
      e = hipEventRecord(c->bg_ev_in, c->stream);  // staged grids in place
    for (int k = 0; k < 4 && e == hipSuccess; k++) {
      if (!hs[k])
        continue;
      if (!c->bg_stream[k])
        e = hipStreamCreateWithFlags(&c->bg_stream[k], hipStreamNonBlocking);
      if (e == hipSuccess && !c->bg_ev_shift[k])
        e = hipEventCreateWithFlags(&c->bg_ev_shift[k], hipEventDisableTiming);
      if (e == hipSuccess)
        e = hipStreamWaitEvent(c->bg_stream[k], c->bg_ev_in, 0);
      if (e == hipSuccess)
        shifting |= 1 << k;
      for (int sl = a.nstore - 2; sl >= 0 && e == hipSuccess; sl--)
        e = hipMemcpyAsync(hs[k] + (long)(sl + 1) * npad, hs[k] + (long)sl * npad, sizeof(float) * npad,
                           hipMemcpyDeviceToDevice, c->bg_stream[k]);
      if (e == hipSuccess)
        e = hipEventRecord(c->bg_ev_shift[k], c->bg_stream[k]);
    }
    if (e != hipSuccess)
      rc = hip_fail(e, "history shift", __FILE__, __LINE__);
  }
  if (rc == MXR_OK && n > 0) {
    size_t b = tmp_bytes;
    hipError_t e = rocprim::radix_sort_pairs(lib_tmp, b, keys_in, keys_out, vals_in, vals_out, (unsigned)n, 0u,
                                             (unsigned)end_bit, c->stream);
    if (e != hipSuccess)
      rc = hip_fail(e, "rocprim::radix_sort_pairs", __FILE__, __LINE__);
    // the occupied cells: their keys, run lengths and count, and the runs' offsets
    b = tmp_bytes;
    if (rc == MXR_OK && (e = rocprim::run_length_encode(lib_tmp, b, keys_out, (unsigned)n, run_keys, run_len, nrun,
                                                        c->stream)) != hipSuccess)
      rc = hip_fail(e, "rocprim::run_length_encode", __FILE__, __LINE__);
    b = tmp_bytes;
    if (rc == MXR_OK && (e = rocprim::exclusive_scan(lib_tmp, b, run_len, run_off, 0u, (size_t)n,
                                                     rocprim::plus<unsigned>(), c->stream)) != hipSuccess)
      rc = hip_fail(e, "rocprim::exclusive_scan", __FILE__, __LINE__);
  }
  if (rc == MXR_OK) {
    hipError_t e = hipSuccess;
    if (e == hipSuccess && n > 0) {
      hipLaunchKernelGGL(k_bg_pack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, a, recs);
      e = hipGetLastError();
    }
    if (e == hipSuccess && n > 0) {
      hipLaunchKernelGGL(k_bg_runs, dim3(ew_blocks(n)), dim3(256), 0, c->stream, a, run_keys, run_len, run_off, nrun,
                         vals_out, recs, cell_vals);
      e = hipGetLastError();
    }
    const long nrows = (long)N * N;
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_bg_row_start, dim3((unsigned)((nrows + 1 + 255) / 256)), dim3(256), 0, c->stream, run_keys,
                         nrun, N, row_run);
      e = hipGetLastError();
    }
    for (int k = 0; k < 4 && e == hipSuccess; k++)
      if (shifting >> k & 1)
        e = hipStreamWaitEvent(c->stream, c->bg_ev_shift[k], 0);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_bg_dense, dim3((unsigned)((nrows + 3) / 4)), dim3(256), 0, c->stream, a, run_keys,
                         cell_vals, row_run);
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
  for (int k = 0; k < 4; k++)  // the copies may still run when an earlier step failed
    if (shifting >> k & 1)
      (void)hipStreamSynchronize(c->bg_stream[k]);
  for (int k = 0; k < 12; k++)
    (void)hipFree(tmp[k]);
  return rc;
}

}  // namespace mxr
