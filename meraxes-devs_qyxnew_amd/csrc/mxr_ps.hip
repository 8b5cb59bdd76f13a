// mxr_ps.hip -- the dimensional 21-cm power spectrum of a real grid (Compute_PS,
// src/core/ComputePowerSpectrum.c:18-255) on one gfx950 GPU.
//
// delta_T (real layout) -> <dT> (FP64, fixed order) -> r2c whose row loads form
// dd = (dT/<dT> - 1) V/N^3 <dT> with the reference's float/double mix -> y and x passes ->
// per-bin sums of k^3 |X|^2 over the half spectrum, bins with float ceilings delta_k 1.35^i.
//
// Binning: wavefronts take runs of PS_U consecutive 64-mode chunks round-robin and keep, per
// lane, FP64 accumulators for a window of PS_WIN consecutive bins; they reduce across lanes only
// when a chunk leaves the window (a ballot decides), which along the half-spectrum rows is rare.  The
// mode counts and |k| sums per bin depend on (N, delta_k) only: they are computed by the same
// kernel once and kept in the context.  Every sum has a fixed order, so results repeat bit for
// bit.
#include <climits>
#include <cmath>

#include "mxr_passes.hpp"

namespace mxr {

constexpr int PS_MAXB = MXR_PS_MAXB;
constexpr int PS_WIN = 4;     // bins held per lane between cross-lane reductions
constexpr int PS_U = 4;       // chunk loads in flight per wavefront
constexpr int PS_NT = 256;
constexpr int PS_MAXBLK = 2048;

int ps_length(int N, double BoxSize, double Hubble_h)
{
  const double box_size = BoxSize / Hubble_h;  // Mpc (ComputePowerSpectrum.c:21)
  const float k_factor = 1.35f;
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

struct PsArgs {
  int nb;
  float delta_k;
  float rdelta_k;            // 1 / delta_k, for the first guess of the bin
  float inv_log2_factor;     // 1 / log2(1.35)
  float edge[PS_MAXB + 2];   // edge[0] = 0, edge[i + 1] = bin i's ceiling (bin i: edge[i] <= k < edge[i + 1]),
                             // edge[nb + 1] = +inf
  const float2* spec;
  double* partials;          // [block][bin][NQ]
};

__device__ __forceinline__ double wave_sum(double x)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
    x += __shfl_xor(x, o, 64);
  return x;
}

__device__ __forceinline__ int wave_min(int x)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
    x = min(x, __shfl_xor(x, o, 64));
  return x;
}

__device__ __forceinline__ int wave_max(int x)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
    x = max(x, __shfl_xor(x, o, 64));
  return x;
}

// GEOM: per bin the |k| sum and the mode count (ComputePowerSpectrum.c:186, 176);
// otherwise the sum of k^3 |X|^2 (:177-179, the 1 / (2 pi^2 V) applied on the host)
template <int N, bool GEOM>
__global__ void __launch_bounds__(PS_NT) k_ps_bin(PsArgs a)
{
#pragma clang fp contract(off)
  // the spectrum's rows are spec_pitch(N) values apart: the index space includes the padding,
  // whose entries are skipped
  constexpr int NW = PS_NT / 64, H = N / 2 + 1, HP = spec_pitch(N), NQ = GEOM ? 2 : 1;
  constexpr unsigned n = (unsigned)N * N * HP, nchunk = (n + 63) / 64;
  __shared__ float2 epair[PS_MAXB + 1];  // (edge[g], edge[g + 1])
  __shared__ double acc[NW][PS_MAXB][NQ];
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64, nb = a.nb;
  for (int i = threadIdx.x; i <= nb; i += PS_NT)
    epair[i] = make_float2(a.edge[i], a.edge[i + 1]);
  for (int i = threadIdx.x; i < NW * PS_MAXB * NQ; i += PS_NT)
    (&acc[0][0][0])[i] = 0.0;
  __syncthreads();

  // wavefronts take segments of PS_U consecutive chunks round-robin, which spreads the chunks
  // near k = 0 (many bins each) over all wavefronts
  const unsigned nwaves = gridDim.x * NW, gw = blockIdx.x * NW + w;
  constexpr unsigned nseg = (nchunk + PS_U - 1) / PS_U;
  double r0[PS_WIN], r1[PS_WIN];
#pragma unroll
  for (int q = 0; q < PS_WIN; q++)
    r0[q] = r1[q] = 0.0;
  int base = 0;  // wave-uniform: the window holds bins base .. base + PS_WIN - 1

  auto flush = [&] {
#pragma unroll
    for (int q = 0; q < PS_WIN; q++) {
      if (base + q < nb && __ballot(r0[q] != 0.0 || r1[q] != 0.0)) {
        const double s0 = wave_sum(r0[q]);
        const double s1 = GEOM ? wave_sum(r1[q]) : 0.0;
        if (lane == 0) {
          acc[w][base + q][0] += s0;
          if constexpr (GEOM)
            acc[w][base + q][1] += s1;
        }
      }
      r0[q] = r1[q] = 0.0;
    }
  };

  // this lane's mode (n_x, n_y, n_z), stepped by 64 per chunk; k_x^2 + k_y^2 per row (:149-159)
  int nz, ny, nx;
  float kxy2;
  auto row_k = [&] {
    const float kx = (float)(nx > N / 2 ? nx - N : nx) * a.delta_k;
    const float ky = (float)(ny > N / 2 ? ny - N : ny) * a.delta_k;
    kxy2 = add_rn(mul_rn(kx, kx), mul_rn(ky, ky));
  };
  auto step = [&] {
    nz += 64;
    while (nz >= HP) {
      nz -= HP;
      if (++ny == N) {
        ny = 0;
        nx++;
      }
      row_k();
    }
  };
  for (unsigned sg = gw; sg < nseg; sg += nwaves) {
    const unsigned c = sg * PS_U, c1 = min(nchunk, c + PS_U);
    {
      const unsigned idx0 = c * 64 + lane;
      nz = (int)(idx0 % HP);
      ny = (int)((idx0 / HP) % N);
      nx = (int)(idx0 / (HP * (unsigned)N));
      row_k();
    }
    float2 xs[PS_U];
    if constexpr (!GEOM) {
      // unconditional loads (clamped index) so that all PS_U are in flight together
#pragma unroll
      for (int u = 0; u < PS_U; u++)
        xs[u] = a.spec[min((c + u) * 64 + lane, n - 1)];
    }
#pragma unroll
    for (int u = 0; u < PS_U; u++) {
      if (c + u >= c1)
        break;
      const unsigned idx = (c + u) * 64 + lane;
      int b = -1;
      double val = 0.0;
      if (idx < n && nz < H) {
        // (:160-166) |k| = (float)sqrt((double)(k_x^2 + k_y^2 + k_z^2)) in the reference, which is
        // the correctly rounded float square root: rounding sqrt to double (53 bits) and then to
        // float (24 bits) equals rounding it to float once, as 53 >= 2 x 24 + 2
        const float kz = (float)nz * a.delta_k;
        const float k2 = add_rn(kxy2, mul_rn(kz, kz));
        const float k_mag = sqrt_rn(k2);
        // bin i: edge[i] <= k < edge[i + 1].  The guess 1 + floor(log_1.35(k / delta_k)) is within
        // one bin of i (the float ceilings drift from delta_k 1.35^j by under 3e-6 relative), so
        // one look at edge[g] and edge[g + 1] decides.
        int g = k_mag > 0.f ? 1 + (int)floorf(__log2f(k_mag * a.rdelta_k) * a.inv_log2_factor) : 0;
        g = min(max(g, 0), nb);
        const float2 e = epair[g];
        const int i = g - (k_mag < e.x ? 1 : 0) + (k_mag >= e.y ? 1 : 0);
        if (i < nb) {
          b = i;
          const double kd = (double)k_mag;
          if constexpr (GEOM) {
            val = kd;
          } else {
            const float2 x = xs[u];
            // a float cubed is exact after one rounding; |X|^2 from the exact squares
            val = kd * kd * kd * ((double)x.x * x.x + (double)x.y * x.y);
          }
        }
      }
      if (__ballot(b >= 0 && (b < base || b >= base + PS_WIN))) {
        flush();
        base = wave_min(b >= 0 ? b : INT_MAX);
        const int bmax = wave_max(b);
        if (bmax >= base + PS_WIN) {
          // a chunk spanning more bins than the window (near k = 0): reduce it directly
          for (int bb = base; bb <= bmax; bb++) {
            const double s0 = wave_sum(b == bb ? val : 0.0);
            const double s1 = GEOM ? wave_sum(b == bb ? 1.0 : 0.0) : 0.0;
            if (lane == 0) {
              acc[w][bb][0] += s0;
              if constexpr (GEOM)
                acc[w][bb][1] += s1;
            }
          }
          step();
          continue;
        }
      }
#pragma unroll
      for (int q = 0; q < PS_WIN; q++) {
        const bool hit = b == base + q;
        r0[q] += hit ? val : 0.0;
        if constexpr (GEOM)
          r1[q] += hit ? 1.0 : 0.0;
      }
      step();
    }
  }
  flush();
  __syncthreads();
  for (int i = threadIdx.x; i < nb * NQ; i += PS_NT) {
    const int bb = i / NQ, q = i % NQ;
    double s = 0.0;
    for (int ww = 0; ww < NW; ww++)
      s += acc[ww][bb][q];
    a.partials[((long)blockIdx.x * nb + bb) * NQ + q] = s;
  }
}

// one block per (bin, quantity): the block partials in a fixed order
__global__ void __launch_bounds__(256) k_ps_reduce(const double* __restrict__ partials, int nblk, int nb, int nq,
                                                  double* __restrict__ out)
{
  __shared__ double red[256];
  const int bb = blockIdx.x / nq, q = blockIdx.x % nq;
  double s = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 256)
    s += partials[((long)b * nb + bb) * nq + q];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o)
      red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    out[bb * nq + q] = red[0];
}

template <int N, bool GEOM>
static int ps_bins(Ctx* c, const PsArgs& a0, double* host)
{
  constexpr int NQ = GEOM ? 2 : 1;
  const long nchunk = (n_spec(N) + 63) / 64;
  const int nblk = (int)std::min<long>(PS_MAXBLK, std::max<long>(1, (nchunk + 8 * PS_U - 1) / (8 * PS_U)));
  PsArgs a = a0;
  a.partials = c->ps_part;
  double* out = c->ps_part + (size_t)PS_MAXBLK * PS_MAXB * 2;
  hipLaunchKernelGGL((k_ps_bin<N, GEOM>), dim3(nblk), dim3(PS_NT), 0, c->stream, a);
  MXR_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_ps_reduce, dim3(a.nb * NQ), dim3(256), 0, c->stream, c->ps_part, nblk, a.nb, NQ, out);
  MXR_HIP(hipGetLastError());
  MXR_HIP(hipMemcpyAsync(host, out, sizeof(double) * a.nb * NQ, hipMemcpyDeviceToHost, c->stream));
  MXR_HIP(hipStreamSynchronize(c->stream));
  return MXR_OK;
}

static PsArgs ps_args(double BoxSize, double Hubble_h, int nb)
{
  PsArgs a{};
  a.nb = nb;
  a.delta_k = (float)(2. * M_PI / (BoxSize / Hubble_h));  // (ComputePowerSpectrum.c:127)
  a.rdelta_k = 1.0f / a.delta_k;
  a.inv_log2_factor = (float)(1.0 / std::log2(1.35));
  // ceilings exactly as the reference steps them: *= 1.35f in float (:162-192)
  const float k_factor = 1.35f;
  float k_ceil = a.delta_k;
  a.edge[0] = 0.0f;
  for (int i = 0; i < nb; i++) {
    a.edge[i + 1] = k_ceil;
    k_ceil *= k_factor;
  }
  a.edge[nb + 1] = INFINITY;
  return a;
}

static int ps_workspace(Ctx* c)
{
  const size_t need = (size_t)PS_MAXBLK * PS_MAXB * 2 + 2 * PS_MAXB;
  if (c->ps_part_cap < need) {
    (void)hipFree(c->ps_part);
    c->ps_part = nullptr;
    MXR_HIP(hipMalloc(&c->ps_part, sizeof(double) * need));
    c->ps_part_cap = need;
  }
  return MXR_OK;
}

// per-bin |k| sums and mode counts: data-independent, computed once per (N, delta_k)
template <int N>
static int ps_geometry(Ctx* c, const PsArgs& a)
{
  if (c->ps_geom_nb == a.nb && c->ps_geom_dk == a.delta_k)
    return MXR_OK;
  RC(ps_workspace(c));
  c->ps_geom_nb = 0;
  RC((ps_bins<N, true>(c, a, c->ps_geom)));
  c->ps_geom_nb = a.nb;
  c->ps_geom_dk = a.delta_k;
  return MXR_OK;
}

template <int N>
static int ps_run(Ctx* c, double BoxSize, double Hubble_h, const float* dT, float* PS_k, float* PS_data,
                  float* PS_error, int nb)
{
  using K = Passes<N>;
  const long nreal = n_real(N);
  const double box_size = BoxSize / Hubble_h;  // Mpc (ComputePowerSpectrum.c:55)
  const float volume = powf((float)box_size, 3);
  RC(ensure_buffers(c, 1));
  RC(ps_workspace(c));
  PsArgs a = ps_args(BoxSize, Hubble_h, nb);
  a.spec = c->spec[0];
  RC(ps_geometry<N>(c, a));
  // <dT> in FP64, fixed order (:76-91)
  const int nblk = (int)ew_blocks(nreal);
  RC(launch_sum_f(c, dT, nreal, nblk));
  RC(launch_final_sum(c, c->partials, nblk, 1, c->sums));
  // unnormalised r2c of dd, formed as the rows load (:97-110)
  CRealPtrs rin{};
  rin.p[0] = dT;
  rin.dd_sum = c->sums;
  rin.dd_volume = (double)volume;
  rin.dd_total = (double)(float)std::pow((double)N, 3);
  rin.dd_rtotal = 1.0 / rin.dd_total;
  rin.dd_total_pow2 = (N & (N - 1)) == 0;
  GridPtrs spec{};
  spec.p[0] = c->spec[0];
  CGridPtrs cspec{};
  cspec.p[0] = c->spec[0];
  FiltArgs none{};
  none.N = N;
  RC(K::r2c(c, rin, spec, 1));
  RC((K::template col<-1, COL_PLAIN>(c, cspec, spec, 1, K::ygeo(), none, 1.0)));
  RC((K::template col<-1, COL_PLAIN>(c, cspec, spec, 1, K::xgeo(), none, 1.0)));
  double p_sum[PS_MAXB];
  RC((ps_bins<N, false>(c, a, p_sum)));
  // (:212-222) with p_box = sum k^3 |X|^2 / (2 pi^2 V)
  const double norm = 2.0 * M_PI * M_PI * (double)volume;
  for (int i = 0; i < nb; i++) {
    const double k_ave = c->ps_geom[2 * i], ct = c->ps_geom[2 * i + 1];
    const double p_box = p_sum[i] / norm;
    PS_k[i] = (float)(k_ave / ct);
    PS_data[i] = (float)(p_box / ct);
    PS_error[i] = (float)((p_box / ct) / std::sqrt(ct));
  }
  return MXR_OK;
}

template <int N>
static int ps_counts(Ctx* c, double BoxSize, double Hubble_h, int nb, double* counts)
{
  const PsArgs a = ps_args(BoxSize, Hubble_h, nb);
  RC(ps_geometry<N>(c, a));
  for (int i = 0; i < nb; i++)
    counts[i] = c->ps_geom[2 * i + 1];
  return MXR_OK;
}

int run_ps_counts(Ctx* c, double BoxSize, double Hubble_h, int nb, double* counts)
{
  if (nb < 1 || nb > PS_MAXB) {
    set_error("power spectrum with %d bins: the library holds 1..%d", nb, PS_MAXB);
    return MXR_ERR_UNSUPPORTED;
  }
  switch (c->N) {
#define X(n)                                                                                                         \
  case n:                                                                                                            \
    return ps_counts<n>(c, BoxSize, Hubble_h, nb, counts);
    MXR_DIMS(X)
#undef X
    default: set_error("unsupported grid dimension %d", c->N); return MXR_ERR_INVALID_ARG;
  }
}

int run_power_spectrum(Ctx* c, const float* dT, int mem, double BoxSize, double Hubble_h, float* PS_k,
                       float* PS_data, float* PS_error, int nb)
{
  if (nb < 1 || nb > PS_MAXB) {
    set_error("power spectrum with %d bins: the library holds 1..%d", nb, PS_MAXB);
    return MXR_ERR_UNSUPPORTED;
  }
  // host grids, and device grids not 16-byte aligned (the row loads read float2), are staged
  if (mem == MXR_MEM_HOST || reinterpret_cast<uintptr_t>(dT) % 16 != 0) {
    if (!c->in_stage[0])
      MXR_HIP(hipMalloc(&c->in_stage[0], sizeof(float) * 2 * n_complex(c->N)));
    MXR_HIP(hipMemcpyAsync(c->in_stage[0], dT, sizeof(float) * n_real(c->N),
                           mem == MXR_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, c->stream));
    dT = c->in_stage[0];
  }
  switch (c->N) {
#define X(n)                                                                                                         \
  case n:                                                                                                            \
    return ps_run<n>(c, BoxSize, Hubble_h, dT, PS_k, PS_data, PS_error, nb);
    MXR_DIMS(X)
#undef X
    default: set_error("unsupported grid dimension %d", c->N); return MXR_ERR_INVALID_ARG;
  }
}

}  // namespace mxr
