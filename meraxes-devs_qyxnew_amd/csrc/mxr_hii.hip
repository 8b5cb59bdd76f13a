// mxr_hii.hip -- find_HII_bubbles on one gfx950 GPU.
//
// Data layout in HBM (all per grid g in {deltax, stars, weighted_sfr[, x_e]}):
//   spec[g]  unfiltered half spectrum I = k + HP (j + N i), k < H = N/2 + 1, HP = spec_pitch(N)
//            (rows padded to 128 bytes), divided by N^3 (find_HII_bubbles.c:235-253); written
//            once per call.
//   work[g]  same shape: per radius the x-pass writes filtered+transformed columns here, the
//            y-pass transforms them in place, the z-pass (c2r) consumes them.
//   ion      one byte per cell: ionised during this call (replaces the xH = 1 reset and the
//            per-radius xH reads of find_HII_bubbles.c:150-153, 490-543).
// FUSED path per radius (3 launches for all grids):
//   x-pass  : W(kR) filter (reionization.c:3373-3433) fused into the loads, inverse FFT over i
//   y-pass  : inverse FFT over j, in place
//   z-pass  : per-row c2r (half-length complex FFT + Hermitian pre-twiddle), then the clamp
//             (find_HII_bubbles.c:372-422), criterion and state update (:451-544) and, on the
//             last radius, the temperatures and FP64 global sums (:549-645).
// For the sharp-k filter every pass skips the modes the window zeroes (exact zeros in the
// reference too): a column or row segment whose smallest |k| is cut is neither read nor
// written, so the work per radius shrinks with the kept k-ball.  The skip tests evaluate the
// window with the reference's float arithmetic, so the skipped set is exactly the zeroed set.
// PLAIN path: the reference's one-kernel-per-step form (find_HII_bubbles_gpu.cu:256-521):
//   per radius copy, filter, 3-pass c2r writing the real field, clamp, criterion kernel.
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "mxr_passes.hpp"
#include "mxr_recomb.hpp"

namespace mxr {

// z-pass epilogues: store the real field, or the find_HII update with the x_e grid (SPIN)
// and / or the BH grids (effective_bhm, effective_bhar)
enum { C2R_STORE = 0, C2R_HII = 1, C2R_HII_SPIN = 2, C2R_HII_BH = 3, C2R_HII_SPIN_BH = 4 };
constexpr bool epi_spin(int e) { return e == C2R_HII_SPIN || e == C2R_HII_SPIN_BH; }
constexpr bool epi_bh(int e) { return e == C2R_HII_BH || e == C2R_HII_SPIN_BH; }

// ---- fused epilogue state -------------------------------------------------------------
struct HiiEpi {
  double f_coll_prefactor;
  float J21c;     // J_21_aux_constant (float, find_HII_bubbles.c:434-435)
  float J21c_bh;  // J_21_aux_constant_BH (:436-437)
  float z;        // (float)redshift
  float TK;       // T_RECFAST(z)
  float cT_ad;    // cT_approx(z)
  // the call's constant powers in the fully ionised temperature, from the host's libm as the
  // reference computes them: ((1+z)/7.1)^2.5, T_RE^1.7, (1e4 (1+z)/4)^1.7
  double fi_z7, fi_tre, fi_he;
  int first, last, uvb, spin, bh;
  uint8_t* ion;
  float* xH;
  float* z_in;
  float* J21;
  float* r_bubble;
  float* temp;
  const float* deltax;  // unfiltered input (padded)
  const float* sfr;     // unfiltered input (padded)
  int sfr_bf16;         // sfr holds bfloat16 bit patterns
  const float* Tk_box;  // real (spin only)
  const float* bhar;    // unfiltered effective_bhar (padded; BH only)
  double* partials;     // up to 17 per block
  // recombinations (find_HII_bubbles.c:226-230, 467-506, 613-636)
  int rec, temp_dep;
  float R;               // (float)R: r_bubble at the first crossing
  float gpref, gpref_bh; // (float)Gamma_R_prefactor, (float)Gamma_R_prefactor_BH (with sigma_HI)
  double redshift, hubble_h;
  float fabs_dtdz, zstep;
  float* Gamma12;        // real, in/out
  float* N_rec;          // padded, in/out (updated after the radius loop)
  float* residual_xH;    // real, out
  float* clumping;       // real, out
  float* t_resp;         // real, out
  RecTab rt;
};

// x^y for x > 0 (or +0, +inf) in double as exp(y log x): the relative error is about
// |y log x| 2^-52 (below 2^-47 for the arguments here), far below the float that each power is
// rounded to, at half the cost of the full-accuracy pow
__device__ __forceinline__ double pow_pos(double x, double y) { return exp(y * log(x)); }

// find_HII_bubbles.c:48-74
__device__ __noinline__ float fully_ionized_T(float z_re, float z, float delta, double fi_z7, double fi_tre,
                                              double fi_he)
{
#pragma clang fp contract(off)
  float result, delta_re;
  if (fabs((double)(z - z_re)) < 1e-4)
    result = 1;
  else {
    delta_re = (float)((double)delta * (1. + (double)z) / (1. + (double)z_re));
    if (delta_re <= -1)
      delta_re = (float)(-1. + (double)ABS_TOL);
    if (delta <= -1)
      delta = (float)(-1. + (double)ABS_TOL);
    result = (float)pow_pos((1. + (double)delta) / (1. + (double)delta_re), 1.1333);
    result = (float)((double)result * pow_pos((1. + (double)z) / (1. + (double)z_re), 3.4));
    result = (float)((double)result * (double)expf((float)(fi_z7 - pow_pos((1. + (double)z_re) / 7.1, 2.5))));
  }
  result = (float)((double)result * fi_tre);
  result = (float)((double)result + fi_he * (double)(1 + delta));
  result = (float)pow_pos((double)result, 0.5882);
  return result;
}

// T_HI * (1 + cT delta) with the reference's promotions: float product, double sum
__device__ __forceinline__ float thermal_T(float TK, float cT, float dx)
{
  return (float)((double)TK * (1. + (double)mul_rn(cT, dx)));
}


// Clamp (find_HII_bubbles.c:377-411) and criterion (:457-485) of one cell, in the reference's
// types: returns whether the cell ionises at this radius and, when `need_lhs`, the exact
// fl(fl(sf / dom) pref) that the last step's partial xH uses.  The quotient is only formed
// when the division-free comparison sf pref vs nf dom is within 1e-12 of a tie; elsewhere
// both orderings agree exactly.
__device__ __forceinline__ bool hii_criterion(const HiiEpi& ea, float df, float& sf, float xf, float& bf,
                                              bool need_lhs, double& lhs, double& nf)
{
#pragma clang fp contract(off)
  df = fmaxf(df, -1.0f + REL_TOL);
  sf = fmaxf(sf, 0.0f);
  if (sf < ABS_TOL)
    sf = 0.0f;
  if (ea.bh) {
    bf = fmaxf(bf, 0.0f);
    if (bf < ABS_TOL)
      bf = 0.0f;
  }
  if (ea.spin) {
    xf = fmaxf(xf, 0.0f);
    if (xf < ABS_TOL)
      xf = 0.0f;
    xf = fminf(xf, 0.999f);
  }
  const double density_over_mean = 1.0 + (double)df;
  nf = ea.spin ? 1.0 - (double)xf : 1.0;
  // (bhm + stars) pref vs nf dom: the two quotients' rounding is far inside the 1e-12 band
  const double a = ((double)sf + (double)(ea.bh ? bf : 0.0f)) * ea.f_coll_prefactor;
  const double b = nf * density_over_mean;
  if (!need_lhs && a > b * (1.0 + 1e-12))
    return true;
  if (!need_lhs && a < b * (1.0 - 1e-12))
    return false;
  const double f_coll_stars = (double)sf / density_over_mean;
  const double f_coll_effective_bhm = ea.bh ? (double)bf / density_over_mean : 0.0;
  lhs = (f_coll_effective_bhm + f_coll_stars) * ea.f_coll_prefactor;
  return lhs > nf * (1. + 0.0);
}

// J_21 at a first crossing with UVBFlag 1 (:531-542): the stars term, plus the BH term
__device__ __forceinline__ float j21_at_crossing(const HiiEpi& ea, float sf, float bf)
{
  // two rounded products and a rounded sum: with contraction on, the compiler may fuse either
  // product into the add, differently in each kernel, and the paths' values would differ
#pragma clang fp contract(off)
  float j = mul_rn(sf, ea.J21c);
  if (ea.bh)
    j = add_rn(j, mul_rn(bf, ea.J21c_bh));
  return j;
}

// The fused update of two adjacent cells ir, ir+1 (ip, ip+1 in the padded layout).
// Per radius (:486-543): a cell not yet ionised in this call that passes the criterion is
// flagged; on its first crossing (z_at_ionization < 0) z_at_ionization and, for UVBFlag 1,
// J_21_at_ionization are set.  On the last (unfiltered) radius every cell also gets its final
// xH (0, or the clamped partial 1 - f_coll) and the not-ionised ones their partial-ionisation
// temperature T_HI (ComputePartiallyIoinizedTemperature(T_HI, xH = 1) = T_HI); the post-loop
// temperatures and the FP64 sums follow in k_postloop.
// `was` holds the two cells' flags, loaded by the caller ahead of the transform (ignored on the
// first radius, where no flag is read).
__device__ __forceinline__ void hii_pair(const HiiEpi& ea, int ir, int ip, float2 df, float2 sf, float2 xf,
                                         float2 bf, uchar2 was)
{
#pragma clang fp contract(off)
  if (ea.first)
    was = make_uchar2(0, 0);
  const float dfs[2] = {df.x, df.y};
  float sfs[2] = {sf.x, sf.y};
  const float xfs[2] = {xf.x, xf.y};
  float bfs[2] = {bf.x, bf.y};
  const bool w[2] = {was.x != 0, was.y != 0};
  bool ion[2], cross[2];
  float xh[2];
#pragma unroll
  for (int e = 0; e < 2; e++) {
    double lhs = 0.0, nf = 1.0;
    ion[e] = w[e];
    xh[e] = w[e] ? 0.0f : 1.0f;
    cross[e] = false;
    if (!w[e]) {
      const bool i = hii_criterion(ea, dfs[e], sfs[e], xfs[e], bfs[e], ea.last != 0, lhs, nf);
      ion[e] = i;
      if (i) {
        xh[e] = 0.0f;
      } else if (ea.last) {
        float v = (float)(nf - lhs);
        xh[e] = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
      }
      cross[e] = xh[e] < REL_TOL;
    }
  }
  if (ea.first || ion[0] != w[0] || ion[1] != w[1])
    *reinterpret_cast<uchar2*>(ea.ion + ir) = make_uchar2(ion[0] ? 1 : 0, ion[1] ? 1 : 0);
  if (cross[0] || cross[1]) {  // a new crossing in this call: z_in read only here
    float2 z = *reinterpret_cast<const float2*>(ea.z_in + ir);
    const float zs[2] = {z.x, z.y};
#pragma unroll
    for (int e = 0; e < 2; e++)
      if (cross[e] && zs[e] < 0) {
        ea.z_in[ir + e] = ea.z;
        if (ea.uvb == 1)
          ea.J21[ir + e] = j21_at_crossing(ea, sfs[e], bfs[e]);
      }
  }
  if (ea.last) {
    *reinterpret_cast<float2*>(ea.xH + ir) = make_float2(xh[0], xh[1]);
    *reinterpret_cast<float2*>(ea.r_bubble + ir) = make_float2(0.0f, 0.0f);
    if (!ion[0] || !ion[1]) {
      float2 T = *reinterpret_cast<const float2*>(ea.temp + ir);
      float Ts[2] = {T.x, T.y};
#pragma unroll
      for (int e = 0; e < 2; e++)
        if (!ion[e])
          Ts[e] = ea.spin ? ea.Tk_box[ir + e] : thermal_T(ea.TK, ea.cT_ad, ea.deltax[ip + e]);
      *reinterpret_cast<float2*>(ea.temp + ir) = make_float2(Ts[0], Ts[1]);
    }
  }
}


// ---- z-pass inverse: half-spectrum rows -> real rows, + store or fused update -----------
// Z[n] = (X[n] + conj X[M-n]) + i exp(+2 pi i n / N) (X[n] - conj X[M-n]), n < M;
// an M-point inverse FFT of Z gives y[2n] + i y[2n+1] (unnormalised c2r).
// C2R_HII keeps the filtered deltax and stars (and x_e) in registers for the update;
// weighted_sfr is transformed and clamped like the reference does (find_HII_bubbles.c:382-385)
// but, with recombinations off, nothing reads it afterwards.
// LAST: this launch is the last (unfiltered) radius; the update's last-radius work (partial xH,
// T_HI, r_bubble) is compiled only into that instance.
template <int N, int NT, int TR, int EPI, int LAST>
__device__ __forceinline__ void row_c2r_body(CGridPtrs in, int ngrid, long nrows, const float2* __restrict__ twg,
                                             FiltArgs fa, const HiiEpi& ea)
{
  constexpr int M = N / 2, H = N / 2 + 1, T = NT / TR, P = M + M / 8 + 1;
  using SL = Slots<M, T>;
  constexpr int NS = SL::QL * SL::RL;
  constexpr bool SPIN = epi_spin(EPI), BH = epi_bh(EPI);
  // grids kept in registers: deltax, stars, [x_e], [effective_bhm]
  constexpr int NK = (EPI == C2R_STORE) ? 1 : 2 + (SPIN ? 1 : 0) + (BH ? 1 : 0);
  constexpr int IXE = 3, IBHM = 3 + (SPIN ? 1 : 0), IBHAR = 4 + (SPIN ? 1 : 0);  // positions in `in`
  __shared__ float2 lds[TR * P];
  __shared__ float2 tw[N];
  const int c = threadIdx.x / T, t = threadIdx.x % T;
  const long row = (long)blockIdx.x * TR + c;
  const bool valid = row < nrows;
  for (int m = threadIdx.x; m < N; m += NT)
    tw[m] = twg[m];
  // number of k_z the window keeps on the k_x = k_y = 0 line (all of them when not pruning)
  int kc = H;
  if (fa.apply && fa.prune) {
    kc = 0;
    for (int k0 = 0; k0 < H; k0 += NT) {
      const int k = k0 + (int)threadIdx.x;
      kc += __syncthreads_count(k < H && sharpk_keep(0, 0, k, fa));
    }
  }
  __syncthreads();
  // the flags of this thread's cells, loaded now so that their latency hides behind the
  // transforms (the epilogue reads them only after all three grids)
  // The loads are unconditional (in-bounds addresses for every lane; the first radius ignores
  // the values), so they issue back to back; two cells' flags per 16 bits, packed two slots per
  // register after the first transform.
  unsigned short fraw[EPI == C2R_STORE ? 1 : NS];
  unsigned flags[EPI == C2R_STORE ? 1 : (NS + 1) / 2];
  if constexpr (EPI != C2R_STORE) {
    const int rr = valid ? (int)row : 0;
#pragma unroll
    for (int q = 0; q < SL::QL; q++) {
      const int j = t + T * q;
#pragma unroll
      for (int r = 0; r < SL::RL; r++) {
        const int n = (SL::BL % T == 0 || j < SL::BL) ? j + r * SL::BL : 0;
        fraw[q * SL::RL + r] = *reinterpret_cast<const unsigned short*>(ea.ion + rr * N + 2 * n);
      }
    }
  }
  // grid order keeps few fields live: weighted_sfr first (transformed, clamped, dropped), then
  // deltax (and x_e) kept, stars last and consumed straight from the transform's registers
  float2 res[NK][NS];
  const int g0 = (EPI == C2R_STORE) ? blockIdx.y : 0;
  if constexpr (EPI != C2R_STORE) {
    float2 wsfr[NS];
    row_c2r_grid<N, T, TR>(in.p[2], valid, row, lds, tw, kc, c, t, wsfr);  // weighted_sfr
    // pack the flags here: their loads stay ahead of the first transform instead of being sunk
    // into the epilogue (they complete with that transform's own loads)
#pragma unroll
    for (int i = 0; i < (NS + 1) / 2; i++) {
      flags[i] = (unsigned)fraw[2 * i] | (2 * i + 1 < NS ? (unsigned)fraw[2 * i + 1] << 16 : 0u);
      asm volatile("" : "+v"(flags[i]));
    }
#pragma unroll
    for (int s = 0; s < NS; s++) {
      float wx = fmaxf(wsfr[s].x, 0.0f), wy = fmaxf(wsfr[s].y, 0.0f);
      wx = wx < ABS_TOL ? 0.0f : wx;
      wy = wy < ABS_TOL ? 0.0f : wy;
      asm volatile("" ::"v"(wx), "v"(wy));  // the clamped field is computed, then dropped
    }
    if constexpr (BH) {
      float2 wbh[NS];
      row_c2r_grid<N, T, TR>(in.p[IBHAR], valid, row, lds, tw, kc, c, t, wbh);  // effective_bhar
#pragma unroll
      for (int s = 0; s < NS; s++) {  // clamped like the reference; Gamma12 alone reads it
        float wx = fmaxf(wbh[s].x, 0.0f), wy = fmaxf(wbh[s].y, 0.0f);
        wx = wx < ABS_TOL ? 0.0f : wx;
        wy = wy < ABS_TOL ? 0.0f : wy;
        asm volatile("" ::"v"(wx), "v"(wy));
      }
    }
    if constexpr (SPIN)
      row_c2r_grid<N, T, TR>(in.p[IXE], valid, row, lds, tw, kc, c, t, res[2]);  // x_e
    if constexpr (BH)
      row_c2r_grid<N, T, TR>(in.p[IBHM], valid, row, lds, tw, kc, c, t, res[NK - 1]);  // effective_bhm
    row_c2r_grid<N, T, TR>(in.p[1], valid, row, lds, tw, kc, c, t, res[1]);  // stars
  }
  row_c2r_grid<N, T, TR>(pick(in, g0), valid, row, lds, tw, kc, c, t, res[0]);  // deltax / the grid

  if constexpr (EPI == C2R_STORE) {
    if (!valid)
      return;
    float* __restrict__ dst =
        const_cast<float*>(reinterpret_cast<const float*>(pick(in, g0))) + row * 2 * spec_pitch(N);
#pragma unroll
    for (int q = 0; q < SL::QL; q++) {
      const int j = t + T * q;
      if (SL::BL % T == 0 || j < SL::BL)
#pragma unroll
        for (int r = 0; r < SL::RL; r++) {
          const int n = j + r * SL::BL;
          *reinterpret_cast<float2*>(dst + 2 * n) = res[0][q * SL::RL + r];
        }
    }
  } else {
    if (!valid)
      return;
    HiiEpi ek = ea;  // the grid set fixes x_e and the BH terms, the launch the last radius
    ek.spin = SPIN;
    ek.bh = BH;
    ek.last = LAST;
#pragma unroll
    for (int q = 0; q < SL::QL; q++) {
      const int j = t + T * q;
      if (SL::BL % T == 0 || j < SL::BL)
#pragma unroll
        for (int r = 0; r < SL::RL; r++) {
          const int n = j + r * SL::BL;
          const int s = q * SL::RL + r;
          const float2 xe = SPIN ? res[2][s] : make_float2(0.f, 0.f);
          const float2 bf = BH ? res[NK - 1][s] : make_float2(0.f, 0.f);
          const int ir = (int)row * N + 2 * n;  // N^3 < 2^31 for N <= 1024
          const int ip = (int)row * 2 * H + 2 * n;
          const unsigned f = flags[s / 2] >> (16 * (s % 2));
          hii_pair(ek, ir, ip, res[0][s], res[1][s], xe, bf, make_uchar2(f & 0xffu, (f >> 8) & 0xffu));
        }
    }
  }
}

// Three waves per SIMD for the store and find_HII epilogues (their registers fit without
// spilling); the epilogues with x_e or the BH grids keep their registers.
template <int N, int NT, int TR, int EPI, int LAST>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(3)))
k_row_c2r(CGridPtrs in, int ngrid, long nrows, const float2* __restrict__ twg, FiltArgs fa, HiiEpi ea)
{
  row_c2r_body<N, NT, TR, EPI, LAST>(in, ngrid, nrows, twg, fa, ea);
}
template <int N, int NT, int TR, int EPI, int LAST>
__global__ void __launch_bounds__(NT) k_row_c2r_spin(CGridPtrs in, int ngrid, long nrows,
                                                     const float2* __restrict__ twg, FiltArgs fa, HiiEpi ea)
{
  row_c2r_body<N, NT, TR, EPI, LAST>(in, ngrid, nrows, twg, fa, ea);
}

// One cell of the fused update (the per-cell form of hii_pair, same order of operations).
__device__ __forceinline__ void hii_cell(const HiiEpi& ea, int ir, int ip, float df, float sf, float xf, float bf,
                                         unsigned char flag)
{
#pragma clang fp contract(off)
  const bool was = !ea.first && flag != 0;
  bool ion = was, cross = false;
  float xh = was ? 0.0f : 1.0f;
  if (!was) {
    double lhs = 0.0, nf = 1.0;
    ion = hii_criterion(ea, df, sf, xf, bf, ea.last != 0, lhs, nf);
    if (ion) {
      xh = 0.0f;
    } else if (ea.last) {
      const float v = (float)(nf - lhs);
      xh = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
    }
    cross = xh < REL_TOL;
  }
  if (ea.first || ion != was)
    ea.ion[ir] = ion ? 1 : 0;
  if (cross && ea.z_in[ir] < 0) {
    ea.z_in[ir] = ea.z;
    if (ea.uvb == 1)
      ea.J21[ir] = j21_at_crossing(ea, sf, bf);
  }
  if (ea.last) {
    ea.xH[ir] = xh;
    ea.r_bubble[ir] = 0.0f;
    if (!ion)
      ea.temp[ir] = ea.spin ? ea.Tk_box[ir] : thermal_T(ea.TK, ea.cT_ad, ea.deltax[ip]);
  }
}

// ---- z-pass inverse by direct summation, for sharp-k radii that keep few k_z -------------
// When only the kc < N/2 leading modes of a row are nonzero, the c2r output is
//   y[n] = sum_{k<kc} w_k Re(X_k exp(+2 pi i k n / N)),  w_0 = 1, w_k = 2 (Im X_0 ignored),
// which costs kc multiply-adds per output instead of a full-length transform.  One lane owns
// p in [0, N/4) and one set of twiddles (C_k, S_k) = w_k (cos, sin)(2 pi k p / N); with
// A_j = sum_{k = j mod 4} a_k C_k and B_j = sum_{k = j mod 4} b_k S_k (X_k = a_k + i b_k):
//   y[p] = sA - sB,  y[N-p] = sA + sB,  y[N/2+p] = dA - dB,  y[N/2-p] = dA + dB,
//   sA = (A0+A2)+(A1+A3), dA = (A0+A2)-(A1+A3) (same for B).
// Lane p = 0 takes (C_k, S_k) = w_k (1, 1) and forms y[0] = sA, y[N/2] = dA,
// y[N/4] = (A0-A2) - (B1-B3), y[3N/4] = (A0-A2) + (B1-B3).  A block is one wavefront owning 64
// consecutive p of `rpb` rows; the twiddles stay in registers across the rows.  A row's
// coefficients (k < kc of every grid; zero above) sit in LDS, where every lane reads the same
// entry (a broadcast); the next row's are loaded, one or a few per lane, while the current row
// is summed, and written to LDS after it.
template <int KMAX>
__device__ __forceinline__ void direct_sum(const float2* cf, const float* C, const float* S, bool lane0, float* o)
{
  float A[4] = {0.f, 0.f, 0.f, 0.f}, B[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k = 0; k < KMAX; k++) {
    const float2 ab = cf[k];
    A[k & 3] = fmaf(ab.x, C[k], A[k & 3]);
    B[k & 3] = fmaf(ab.y, S[k], B[k & 3]);
  }
  const float e02 = A[0] + A[2], e13 = A[1] + A[3], f02 = B[0] + B[2], f13 = B[1] + B[3];
  const float sA = e02 + e13, dA = e02 - e13, sB = f02 + f13, dB = f02 - f13;
  const float a02 = A[0] - A[2], b13 = B[1] - B[3];
  o[0] = lane0 ? sA : sA - sB;
  o[1] = lane0 ? dA : sA + sB;
  o[2] = lane0 ? a02 - b13 : dA - dB;
  o[3] = lane0 ? a02 + b13 : dA + dB;
  // the sums are complete here, in program order.  Without this the compiler sinks a grid's
  // multiply-adds past the next row's LDS commit while its coefficient reads must stay ahead of
  // that store, so every read of the row stayed live across it (KMAX 24 needed 183 VGPRs and
  // spilled; with the sums pinned it takes 111, with nothing spilled up to KMAX 64)
  asm volatile("" ::"v"(o[0]), "v"(o[1]), "v"(o[2]), "v"(o[3]) : "memory");
}

template <int N, int KMAX, int EPI>
// waves per SIMD by the registers each form needs with its sums pinned (at most 128 VGPRs for
// four waves, 168 for three): the find_HII epilogue fits four up to 32 modes and three up to 48;
// the x_e / BH epilogues, a few registers larger, four up to 24 and three up to 40
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(KMAX <= (EPI == C2R_HII ? 32 : 24) ? 4 : KMAX <= (EPI == C2R_HII ? 48 : 40) ? 3 : 2)))
k_row_direct(CGridPtrs in, long nrows, int rpb, const float2* __restrict__ twg, int kc, HiiEpi ea)
{
  constexpr int H = N / 2 + 1, Q = N / 4, SEGS = Q / 64;
  static_assert(Q % 64 == 0, "N/4 must be a multiple of the wavefront");
  static_assert(KMAX < H, "direct summation covers filtered radii only");
  constexpr bool SPIN = epi_spin(EPI), BH = epi_bh(EPI);
  constexpr int IXE = 3, IBHM = 3 + (SPIN ? 1 : 0), IBHAR = 4 + (SPIN ? 1 : 0);
  // grids in summation order: weighted_sfr, [effective_bhar], [x_e], [effective_bhm], stars, deltax
  constexpr int NG = 3 + (SPIN ? 1 : 0) + (BH ? 2 : 0);
  constexpr int G_BHAR = 1, G_XE = 1 + (BH ? 1 : 0), G_BHM = G_XE + (SPIN ? 1 : 0), G_ST = NG - 2, G_DX = NG - 1;
  constexpr int NL = (NG * KMAX + 63) / 64;  // coefficient loads per lane and row
  __shared__ float2 cf[NG * KMAX];
  const float2* gp[NG];
  {
    int s = 0;
    gp[s++] = in.p[2];
    if constexpr (BH)
      gp[s++] = in.p[IBHAR];
    if constexpr (SPIN)
      gp[s++] = in.p[IXE];
    if constexpr (BH)
      gp[s++] = in.p[IBHM];
    gp[s++] = in.p[1];
    gp[s++] = in.p[0];
  }
  // XCD-aware order: the GPU deals blocks to its 8 XCDs round-robin by index, so the SEGS blocks
  // of a row group take indices 8 apart and run on one XCD, whose L2 serves the group's
  // coefficient rows to all of them (row groups past the last are empty)
  const unsigned xcd = blockIdx.x % 8, kq = blockIdx.x / 8;
  const int seg = (int)(kq % SEGS);
  const long row0 = ((long)(kq / SEGS) * 8 + xcd) * rpb;
  const int p = seg * 64 + (int)threadIdx.x;
  const bool lane0 = p == 0;
  float C[KMAX], S[KMAX];
#pragma unroll
  for (int k = 0; k < KMAX; k++) {
    const float2 w = twg[(k * p) % N];  // exp(-2 pi i k p / N)
    const float wk = k == 0 ? 1.f : 2.f;
    C[k] = lane0 ? wk : wk * w.x;
    S[k] = lane0 ? wk : -wk * w.y;
  }
  const int n[4] = {lane0 ? 0 : p, lane0 ? N / 2 : N - p, lane0 ? N / 4 : N / 2 + p, lane0 ? 3 * N / 4 : N / 2 - p};
  const long row1 = row0 + rpb < nrows ? row0 + rpb : nrows;
  // this lane's coefficient slots e = lane + 64 i: grid e / KMAX, mode e % KMAX; slots past the
  // last grid and modes k >= kc hold zero (the y pass leaves stale values there), the loads of
  // such slots read the row's first entry (in bounds) and are discarded
  auto fetch = [&](long row, float2* pre) {
#pragma unroll
    for (int i = 0; i < NL; i++) {
      const int e = (int)threadIdx.x + 64 * i;
      const int g = e / KMAX, k = e % KMAX;
      const bool ok = e < NG * KMAX && k < kc;
      const float2* base = gp[0];
#pragma unroll
      for (int t = 1; t < NG; t++)
        base = g == t ? gp[t] : base;
      const float2 x = base[row * spec_pitch(N) + (ok ? k : 0)];
      pre[i] = ok ? x : make_float2(0.f, 0.f);
    }
  };
  auto commit = [&](const float2* pre) {
#pragma unroll
    for (int i = 0; i < NL; i++) {
      const int e = (int)threadIdx.x + 64 * i;
      if (e < NG * KMAX)
        cf[e] = pre[i];
    }
    SyncWave::sync();
  };
  // the epilogue's flags as compile-time facts: this form serves filtered radii only, never the
  // last (unfiltered) one, and its grid set fixes x_e and the BH terms
  HiiEpi ek = ea;
  ek.last = 0;
  ek.spin = SPIN;
  ek.bh = BH;
  float2 pre[NL];
  if (row0 < row1) {
    fetch(row0, pre);
    commit(pre);
  }
  // flags one row ahead, so their latency hides behind the row's sums
  unsigned char fl[4] = {0, 0, 0, 0};
  if (!ea.first && row0 < row1)
#pragma unroll
    for (int e = 0; e < 4; e++)
      fl[e] = ea.ion[(int)row0 * N + n[e]];
  for (long row = row0; row < row1; row++) {
    const bool more = row + 1 < row1;
    if (more)
      fetch(row + 1, pre);
    unsigned char nfl[4] = {0, 0, 0, 0};
    if (!ea.first && more)
#pragma unroll
      for (int e = 0; e < 4; e++)
        nfl[e] = ea.ion[(int)(row + 1) * N + n[e]];
    float sfr[4], xe[4] = {0.f, 0.f, 0.f, 0.f}, bf[4] = {0.f, 0.f, 0.f, 0.f}, st[4], dx[4];
    // weighted_sfr and effective_bhar: transformed, clamped, dropped (Gamma12 alone reads them)
    direct_sum<KMAX>(cf, C, S, lane0, sfr);
#pragma unroll
    for (int e = 0; e < 4; e++) {
      float w = fmaxf(sfr[e], 0.0f);
      w = w < ABS_TOL ? 0.0f : w;
      asm volatile("" ::"v"(w));
    }
    if constexpr (BH) {
      float bar[4];
      direct_sum<KMAX>(cf + G_BHAR * KMAX, C, S, lane0, bar);
#pragma unroll
      for (int e = 0; e < 4; e++) {
        float w = fmaxf(bar[e], 0.0f);
        w = w < ABS_TOL ? 0.0f : w;
        asm volatile("" ::"v"(w));
      }
    }
    if constexpr (SPIN)
      direct_sum<KMAX>(cf + G_XE * KMAX, C, S, lane0, xe);
    if constexpr (BH)
      direct_sum<KMAX>(cf + G_BHM * KMAX, C, S, lane0, bf);
    direct_sum<KMAX>(cf + G_ST * KMAX, C, S, lane0, st);
    direct_sum<KMAX>(cf + G_DX * KMAX, C, S, lane0, dx);
    // every lane has read this row's coefficients (a wavefront's LDS accesses run in order)
    if (more) {
      SyncWave::sync();
      commit(pre);
    }
#pragma unroll
    for (int e = 0; e < 4; e++)
      hii_cell(ek, (int)row * N + n[e], (int)row * 2 * H + n[e], dx[e], st[e], xe[e], bf[e], fl[e]);
#pragma unroll
    for (int e = 0; e < 4; e++)
      fl[e] = nfl[e];
  }
}

// ---- plain-path kernels (one kernel per reference step) -------------------------------
__global__ void k_fill(float* __restrict__ p, float v, long n)
{
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    p[i] = v;
}

// complex_vector_times_scalar: spectrum /= N^3 (double divisor)
__global__ void k_scale(float2* __restrict__ p, long n, double div)
{
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    float2 x = p[i];
    p[i] = make_float2((float)((double)x.x / div), (float)((double)x.y / div));
  }
}

// filter(): the window on a slab of a half spectrum
// pitch: complex values per row (N/2 + 1 for the caller's layout, spec_pitch(N) internally)
__global__ void k_filter(float2* __restrict__ box, int local_ix_start, int slab_nx, FiltArgs fa, int pitch)
{
  const int N = fa.N, H = N / 2 + 1;
  const long n = (long)slab_nx * N * pitch;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += (long)gridDim.x * blockDim.x) {
    const int nz = (int)(idx % pitch);
    if (nz >= H)
      continue;
    const int ny = (int)((idx / pitch) % N);
    const int nx = (int)(idx / ((long)pitch * N)) + local_ix_start;
    if (fa.ftype == 1) {
      if (!sharpk_keep(nx, ny, nz, fa))
        box[idx] = make_float2(0.f, 0.f);
    } else {
      const float w = window_mul(nx, ny, nz, fa);
      const float2 x = box[idx];
      box[idx] = make_float2(mul_rn(x.x, w), mul_rn(x.y, w));
    }
  }
}

// sanity_check_aliasing on one real grid in a spectrum buffer (rows of 2 spec_pitch(N) floats):
// kind 0 deltax, 1 stars / sfr / BH grids, 2 x_e
__global__ void k_clamp(float* __restrict__ f, int N, int kind)
{
  const long n = (long)N * N * N;
  for (long ir = (long)blockIdx.x * blockDim.x + threadIdx.x; ir < n; ir += (long)gridDim.x * blockDim.x) {
    const long row = ir / N;
    const long ip = row * 2 * spec_pitch(N) + ir % N;
    float x = f[ip];
    if (kind == 0) {
      x = fmaxf(x, -1.0f + REL_TOL);
    } else {
      x = fmaxf(x, 0.0f);
      if (x < ABS_TOL)
        x = 0.0f;
      if (kind == 2)
        x = fminf(x, 0.999f);
    }
    f[ip] = x;
  }
}

// find_HII_bubbles_gpu_main_loop shape, with the CPU path's semantics (:451-544)
struct PlainArgs {
  int N;
  // filtered fields in the spectrum buffers: real rows of 2 spec_pitch(N) floats
  const float* dfl;
  const float* sfl;
  const float* xfl;
  const float* bfl;   // filtered effective_bhm (BH only)
  const float* wfl;   // filtered weighted_sfr (recombinations only)
  const float* bafl;  // filtered effective_bhar (recombinations with BH only)
  const float* nfl;   // filtered N_rec (recombinations only)
};

__global__ void k_criterion(PlainArgs pa, HiiEpi ea)
{
#pragma clang fp contract(off)
  const int N = pa.N, H = N / 2 + 1;
  const long n = (long)N * N * N;
  for (long ir = (long)blockIdx.x * blockDim.x + threadIdx.x; ir < n; ir += (long)gridDim.x * blockDim.x) {
    const long ip = (ir / N) * 2 * H + ir % N;                  // the caller's padded grids
    const long iw = (ir / N) * 2 * spec_pitch(N) + ir % N;      // the filtered fields
    const double density_over_mean = 1.0 + (double)pa.dfl[iw];
    const float sf = pa.sfl[iw];
    const double f_coll_stars = (double)sf / density_over_mean;
    const float bf = ea.bh ? pa.bfl[iw] : 0.0f;
    const double f_coll_effective_bhm = ea.bh ? (double)bf / density_over_mean : 0.0;
    const double neutral_fraction = ea.spin ? 1.0 - (double)pa.xfl[iw] : 1.0;
    const double rec = ea.rec ? (double)pa.nfl[iw] / density_over_mean : 0.0;
    const double lhs = (f_coll_effective_bhm + f_coll_stars) * ea.f_coll_prefactor;
    float xh = ea.xH[ir];
    if (lhs > neutral_fraction * (1. + rec)) {
      // first crossing in this call (largest R) with recombinations: the ionising background,
      // the radius and, for UVBFlag 2, J_21 (:488-506)
      if (ea.rec && xh > REL_TOL) {
        float g12 = mul_rn(pa.wfl[iw], ea.gpref);
        if (ea.bh)
          g12 = add_rn(g12, mul_rn(pa.bafl[iw], ea.gpref_bh));
        ea.Gamma12[ir] = g12;
        ea.r_bubble[ir] = ea.R;
        if (ea.uvb == 2)
          ea.J21[ir] = j21_at_crossing(ea, sf, bf);
      }
      xh = 0.0f;
    } else if (ea.last && (xh > REL_TOL)) {
      const float T_HI = ea.spin ? ea.Tk_box[ir] : thermal_T(ea.TK, ea.cT_ad, ea.deltax[ip]);
      // ComputePartiallyIoinizedTemperature (find_HII_bubbles.c:76-84)
      float tp;
      if (xh <= 0.f)
        tp = (float)T_RE;
      else if (xh >= 1.f)
        tp = T_HI;
      else
        tp = (float)((double)mul_rn(T_HI, xh) + T_RE * (1. - (double)xh));
      ea.temp[ir] = tp;
      xh = (float)(neutral_fraction - lhs);
      if (xh < 0.0f)
        xh = 0.0f;
      else if (xh > 1.0f)
        xh = 1.0f;
    }
    ea.xH[ir] = xh;
    if ((xh < REL_TOL) && (ea.z_in[ir] < 0)) {
      ea.z_in[ir] = ea.z;
      if (ea.uvb == 1)
        ea.J21[ir] = j21_at_crossing(ea, sf, bf);
    }
  }
}

// post-loop temperatures and FP64 partial sums (find_HII_bubbles.c:577-645)
template <int NT, bool REC>
__global__ void __launch_bounds__(NT) k_postloop(int N, HiiEpi ea)
{
#pragma clang fp contract(off)
  const int H = N / 2 + 1;
  const long n = (long)N * N * N;
  constexpr int NV = REC ? 17 : 7;
  double acc[NV];
#pragma unroll
  for (int q = 0; q < NV; q++)
    acc[q] = 0.0;
  for (long ir = (long)blockIdx.x * NT + threadIdx.x; ir < n; ir += (long)gridDim.x * NT) {
    const long ip = (ir / N) * 2 * H + ir % N;
    const float xh = ea.xH[ir];
    const float dx_in = ea.deltax[ip];
    const double dom = 1.0 + (double)dx_in;
    const float zin = ea.z_in[ir];
    float temp = ea.temp[ir];
    if (zin > 0 && xh < REL_TOL)
      temp = fully_ionized_T(zin, ea.z, dx_in, ea.fi_z7, ea.fi_tre, ea.fi_he);
    const float floor_T = ea.spin ? ea.Tk_box[ir] : thermal_T(ea.TK, ea.cT_ad, dx_in);
    if (temp < floor_T)
      temp = floor_T;
    ea.temp[ir] = temp;
    acc[0] += (double)xh;
    acc[1] += (double)xh * dom;
    acc[2] += dom;
    acc[3] += (double)(ea.sfr_bf16 ? bf16_to_f32(reinterpret_cast<const unsigned short*>(ea.sfr)[ip]) : ea.sfr[ip]);
    acc[4] += (double)temp;
    acc[5] += (double)temp * dom;
    if (ea.bh)
      acc[6] += (double)ea.bhar[ip];
    if constexpr (REC) {
      // :612-636: the cell's recombinations over the snapshot step from the tables at its
      // effective redshift, background Gamma_12 and temperature
      const double z_eff = (1. + ea.redshift) * pow(dom, 1.0 / 3.0) - 1;
      const double T = ea.temp_dep ? (double)temp : 1e4;
      const double h = ea.hubble_h;
      const float g12 = ea.Gamma12[ir];
      double rate, rnh, cf;
      splined_recombination(ea.rt, z_eff, (double)g12 * h * h, T, &rate, &rnh, &cf);
      const float step = mul_rn(mul_rn((float)rate, ea.fabs_dtdz), ea.zstep);
      const float nrec = (float)((double)ea.N_rec[ip] + (double)step * (1. - (double)(float)xh));
      ea.N_rec[ip] = nrec;
      const float res = (float)rnh, clump = (float)cf;
      ea.residual_xH[ir] = res;
      ea.clumping[ir] = clump;
      float tr = (float)((double)g12 * h * h * 1e-12 + rate);
      tr = tr > 0 ? (float)(1.0 / (double)tr / SEC_PER_MEGAYEAR) : 1e30f;
      ea.t_resp[ir] = tr;
      acc[7] += (double)nrec;
      acc[8] += (double)nrec * dom;
      acc[9] += (double)res;
      acc[10] += (double)res * dom;
      acc[11] += (double)clump;
      acc[12] += (double)clump * dom;
      acc[13] += (double)g12;
      acc[14] += (double)g12 * dom;
      const float rb = ea.r_bubble[ir];  // zero unless set at a first crossing with recombinations
      acc[15] += (double)rb;
      acc[16] += (double)rb * dom;
    }
  }
  block_sum<NT, NV>(acc, ea.partials + (long)blockIdx.x * NV);
}

// fixed-order final sum of per-block partials (deterministic)
__global__ void k_final_sum(const double* __restrict__ partials, int nblk, int nv, double* __restrict__ out)
{
  __shared__ double red[256];
  for (int q = 0; q < nv; q++) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 256)
      s += partials[(long)b * nv + q];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if (threadIdx.x < o)
        red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0)
      out[q] = red[0];
    __syncthreads();
  }
}

int launch_final_sum(Ctx* c, const double* partials, int nblk, int nv, double* out)
{
  hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(256), 0, c->stream, partials, nblk, nv, out);
  MXR_HIP(hipGetLastError());
  return MXR_OK;
}

// ---- brightness temperature (BrightnessTemperature.c:44-124, 512-545) ------------------
struct TbArgs {
  int N;
  float const_factor;
  double T_rad;
  int spin;
  int tau;   // PecVels > 1: optical-depth form (1 + z) / (1000 T_S)
  float z;
  const float* xH;
  const float* deltax;
  const float* TS;
  float* dT;
  double* partials;
};

template <int NT>
__global__ void __launch_bounds__(NT) k_tb(TbArgs a)
{
#pragma clang fp contract(off)
  const int N = a.N, H = N / 2 + 1;
  const long n = (long)N * N * N;
  double acc[1] = {0.0};
  for (long ir = (long)blockIdx.x * NT + threadIdx.x; ir < n; ir += (long)gridDim.x * NT) {
    const long ip = (ir / N) * 2 * H + ir % N;
    const double pixel_deltax = a.deltax[ip];
    const double pixel_x_HI = a.xH[ir];
    float dT = (float)((double)a.const_factor * pixel_x_HI * (1.0 + pixel_deltax));
    if (a.spin) {
      const double f = a.tau ? (1. + (double)a.z) / (1000. * (double)a.TS[ir]) : (1 - a.T_rad / (double)a.TS[ir]);
      dT = (float)((double)dT * f);
    }
    a.dT[ir] = dT;
    acc[0] += (double)dT;
  }
  block_sum<NT, 1>(acc, a.partials + blockIdx.x);
}

// ---- the brightness box without allocation ---------------------------------------------
// Scratch of mxr_brightness_temperature_noalloc: the per-block FP64 partials of k_tb (at most
// 8192 blocks, ew_blocks' cap) and the final sum.  Each GPU holds its own copy for the life of
// the process; calls on one GPU share it, so calls that overlap on one GPU race.
constexpr int TB_SCRATCH_BLOCKS = 8192;
static __device__ __align__(256) double scratch[TB_SCRATCH_BLOCKS + 8];

__global__ void k_mean_out(const double* __restrict__ sum, double total, double* __restrict__ out)
{
  if (threadIdx.x == 0 && blockIdx.x == 0)
    out[0] = sum[0] / total;
}

// the device address of `scratch` on the current device, looked up once per device
static int scratch_ptr(double** out)
{
  constexpr int MAXDEV = 64;
  static double* addr[MAXDEV] = {};
  static std::mutex mu;
  int dev = 0;
  MXR_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAXDEV) {
    set_error("device %d beyond the %d this library tracks", dev, MAXDEV);
    return MXR_ERR_INVALID_ARG;
  }
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    MXR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    addr[dev] = static_cast<double*>(p);
  }
  *out = addr[dev];
  return MXR_OK;
}

int run_brightness_temperature_noalloc(const mxr_tb_params* p, const mxr_tb_grids* g, double* ave_dev,
                                       hipStream_t stream)
{
  const int N = p->ReionGridDim;
  if (N < 2 || N % 2 || N > 2048) {
    set_error("ReionGridDim %d: must be even and at most 2048", N);
    return MXR_ERR_INVALID_ARG;
  }
  if (g->mem != MXR_MEM_DEVICE || !g->xH || !g->deltax || !g->delta_T || !ave_dev ||
      (p->Flag_IncludeSpinTemp && !g->TS_box)) {
    set_error("the no-allocation brightness box takes device pointers: xH, deltax, delta_T, the mean%s",
              p->Flag_IncludeSpinTemp ? ", TS_box" : "");
    return MXR_ERR_INVALID_ARG;
  }
  if (p->Flag_IncludeSpinTemp && p->Flag_IncludePecVelsFor21cm) {
    set_error("peculiar-velocity terms need the FFT workspace of a context (mxr_brightness_temperature)");
    return MXR_ERR_UNSUPPORTED;
  }
  const long nreal = (long)N * N * N;
  const int nblk = (int)ew_blocks(nreal);
  if (nblk > TB_SCRATCH_BLOCKS) {
    set_error("%d blocks of partial sums exceed the scratch of %d", nblk, TB_SCRATCH_BLOCKS);
    return MXR_ERR_INVALID_ARG;
  }
  double* sc = nullptr;
  RC(scratch_ptr(&sc));
  const float redshift = (float)p->redshift;
  const float Hubble_h = (float)p->Hubble_h;
  const float OmegaM = (float)p->OmegaM;
  const float OmegaB = (float)(OmegaM * p->BaryonFrac);
  TbArgs a{};
  a.N = N;
  a.const_factor = (float)(27.0 * (OmegaB * Hubble_h * Hubble_h / 0.023) *
                           std::sqrt((0.15 / OmegaM / Hubble_h / Hubble_h) * (1.0 + redshift) / 10.0));
  a.T_rad = TCMB * (1. + redshift);
  a.spin = p->Flag_IncludeSpinTemp ? 1 : 0;
  a.tau = 0;
  a.z = redshift;
  a.xH = g->xH;
  a.deltax = g->deltax;
  a.TS = g->TS_box;
  a.dT = g->delta_T;
  a.partials = sc;
  hipLaunchKernelGGL((k_tb<256>), dim3(nblk), dim3(256), 0, stream, a);
  MXR_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(256), 0, stream, sc, nblk, 1, sc + TB_SCRATCH_BLOCKS);
  MXR_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_mean_out, dim3(1), dim3(64), 0, stream, sc + TB_SCRATCH_BLOCKS, std::pow((double)N, 3),
                     ave_dev);
  MXR_HIP(hipGetLastError());
  return MXR_OK;
}

// ---- peculiar-velocity terms (BrightnessTemperature.c:158-510), each in the reference's types

// velocity grid to comoving velocity, in place on the N^3 cells of the padded grid (:169-198)
__global__ void k_tb_velconv(float* __restrict__ vel, int N, int trees, float z)
{
#pragma clang fp contract(off)
  const int H = N / 2 + 1;
  const long n = (long)N * N * N;
  const double opz = 1. + (double)z;
  for (long ir = (long)blockIdx.x * blockDim.x + threadIdx.x; ir < n; ir += (long)gridDim.x * blockDim.x) {
    const long ip = (ir / N) * 2 * H + ir % N;
    float v = vel[ip];
    if (trees == MXR_TREES_GBPTREES)
      v = (float)(sqrt(1. / opz) * (double)v / 1000.);
    else if (trees == MXR_TREES_VELOCIRAPTOR || trees == MXR_TREES_VELOCIRAPTOR_AUG)
      v = (float)(opz * (double)v);
    v = (float)((double)v * (1000. * 100. / MPC));
    v = (float)((double)v / (1. / opz));
    vel[ip] = v;
  }
}

// optionally x Hubble_h (double product, float result), then velocity_gradient
// (reionization.c:3435-3450): (a + ib)(0 + i k_z), k_z = n_z dk in float; pitch = values per row
// (N/2 + 1 in the caller's layout; the internal pitch's padding entries are scaled and unused)
__global__ void k_vgrad(float2* __restrict__ s, long n, int pitch, float dk, double h, int scale_h)
{
#pragma clang fp contract(off)
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    float2 x = s[i];
    if (scale_h)
      x = make_float2((float)((double)x.x * h), (float)((double)x.y * h));
    const float kz = (float)(int)(i % pitch) * dk;
    s[i] = make_float2(x.x * 0.0f - x.y * kz, x.x * kz + x.y * 0.0f);
  }
}

struct TbVelArgs {
  int N, pecvel;
  float H_z, z;
  double T_rad, max_v_deriv;
  const float* grad;  // dv/dz, real rows of 2 spec_pitch(N) floats (a spectrum buffer)
  const float* TS;
  float* dT;
};

// PecVels 1: capped gradient (:217-248); 2 and 3: optical depth with the uncapped gradient (:250-309)
__global__ void k_tb_pecvel(TbVelArgs a)
{
#pragma clang fp contract(off)
  const int N = a.N;
  const long n = (long)N * N * N;
  const double Hz = (double)a.H_z, opz = 1. + (double)a.z;
  for (long ir = (long)blockIdx.x * blockDim.x + threadIdx.x; ir < n; ir += (long)gridDim.x * blockDim.x) {
    const double g = (double)a.grad[(ir / N) * 2 * spec_pitch(N) + ir % N];
    if (a.pecvel == 1) {
      double dvdx = g;
      if (fabs(dvdx) > a.max_v_deriv)
        dvdx = dvdx < 0 ? -a.max_v_deriv : a.max_v_deriv;
      a.dT[ir] = (float)((double)a.dT[ir] / (dvdx / Hz + 1.0));
    } else {
      const double gc = fabs(g / Hz + 1.0);
      const double TS = (double)a.TS[ir];
      if (gc < 1e-7)
        a.dT[ir] = (float)(1000. * (TS - a.T_rad) / opz);
      else
        a.dT[ir] = (float)((1. - exp(-(double)a.dT[ir] / gc)) * 1000. * (TS - a.T_rad) / opz);
    }
  }
}

constexpr int N_RSD_STEPS = 50;  // reionization.h:34
struct TbRsdArgs {
  int N;
  float H_z;
  double sw;                        // sub-cell width (BoxSize/h / N / 50)
  double cw;                        // BoxSize/h / (float)N
  float x_pos[N_RSD_STEPS], x_off[N_RSD_STEPS];
  float* dT;
  const float* vel;                 // comoving velocity, padded
  const float* xH;
  float* los;                       // N^3 scratch
};

__device__ __forceinline__ int wrap_cell(long k, int N) { return (int)(((k % N) + N) % N); }

// PecVels 3: line-of-sight redshift-space redistribution (:311-510).  One thread per (i, j)
// row walks the cells and sub-cells in the reference's order, so every destination sums its
// contributions in the same order; destinations wrap modulo N (identical to the reference's
// periodic index wherever that stays inside the row).
__global__ void __launch_bounds__(64) k_tb_rsd(TbRsdArgs a)
{
#pragma clang fp contract(off)
  const int N = a.N, H = N / 2 + 1;
  const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= (long)N * N)
    return;
  float* los = a.los + r * N;
  float* dT = a.dT + r * N;
  const float* v = a.vel + r * 2 * H;
  const float* xH = a.xH + r * N;
  for (int k = 0; k < N; k++)
    los[k] = 0.0f;
  const double fw_unit = a.sw / a.cw;
  const double half = (a.sw / 2.) / a.cw;
  for (int kk = 0; kk < N; kk++) {
    const float d = dT[kk];
    if (!(fabs((double)d) >= 1e-7 && (double)xH[kk] >= 1e-7))
      continue;
    const double d1_low = (double)__fdiv_rn(v[kk == 0 ? N - 1 : kk - 1], a.H_z);
    const double d2_low = (double)__fdiv_rn(v[kk], a.H_z);
    const double d1_high = d2_low;
    const double d2_high = (double)__fdiv_rn(v[kk == N - 1 ? 0 : kk + 1], a.H_z);
    const float d50 = __fdiv_rn(d, (float)N_RSD_STEPS);
    for (int s = 0; s < N_RSD_STEPS; s++) {
      const double xp = (double)a.x_pos[s];
      const double disp = xp <= 0.5 ? d1_low + ((xp + 0.5) - 0.) * (d2_low - d1_low) / (1. - 0.)
                                    : d1_high + ((xp - 0.5) - 0.) * (d2_high - d1_high) / (1. - 0.);
      const double pos = ((double)a.x_off[s] + disp) / a.cw;
      const double lo = pos - half, hi = pos + half;
      auto add = [&](long k, double f) { los[wrap_cell(k, N)] = (float)((double)los[wrap_cell(k, N)] + f * (double)d / (double)N_RSD_STEPS); };
      if (lo >= 0.0 && hi < 1.0) {
        los[kk] = los[kk] + d50;
      } else if (lo < 0.0 && hi < 0.0) {
        const double cd = ceil(fabs(lo)) - 1.;
        if (fabs(hi) > cd) {
          const int k = wrap_cell(kk - ((long)cd + 1), N);
          los[k] = los[k] + d50;
        } else {
          const double fo = (fabs(lo) - cd) / fw_unit;
          add(kk - (long)cd, 1. - fo);
          add(kk - ((long)cd + 1), fo);
        }
      } else if (lo < 0.0 && (hi > 0.0 && hi < 1.0)) {
        const double fwi = hi / fw_unit;
        add(kk - 1, 1. - fwi);
        add(kk, fwi);
      } else if ((lo >= 0.0 && lo < 1.0) && (hi >= 1.0)) {
        const double fo = (hi - 1.) / fw_unit;
        add(kk, 1. - fo);
        add(kk + 1, fo);
      } else {
        const double cd = floor(fabs(hi));
        if (lo >= cd) {
          const int k = wrap_cell(kk + (long)cd, N);
          los[k] = los[k] + d50;
        } else {
          const double fo = (hi - cd) / fw_unit;
          add(kk + (long)cd - 1, 1. - fo);
          add(kk + (long)cd, fo);
        }
      }
    }
  }
  for (int k = 0; k < N; k++)
    dT[k] = los[k];
}

// FP64 block sums of a real grid (scalar loads)
template <int NT>
__global__ void __launch_bounds__(NT) k_sum_f(const float* __restrict__ x, long n, double* partials)
{
  double acc[1] = {0.0};
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < n; i += (long)gridDim.x * NT)
    acc[0] += (double)x[i];
  block_sum<NT, 1>(acc, partials + blockIdx.x);
}

// the same with 16-byte loads, four in flight per thread (x 16-byte aligned, n % 4 == 0)
template <int NT>
__global__ void __launch_bounds__(NT) k_sum_f4(const float4* __restrict__ x, long n4, double* partials)
{
  constexpr int U = 4;
  double acc[1] = {0.0};
  const long stride = (long)gridDim.x * NT;
  long i = (long)blockIdx.x * NT + threadIdx.x;
  for (; i + (U - 1) * stride < n4; i += U * stride) {
    float4 v[U];
#pragma unroll
    for (int u = 0; u < U; u++)
      v[u] = x[i + u * stride];
#pragma unroll
    for (int u = 0; u < U; u++)
      acc[0] += ((double)v[u].x + (double)v[u].y) + ((double)v[u].z + (double)v[u].w);
  }
  for (; i < n4; i += stride) {
    const float4 v = x[i];
    acc[0] += ((double)v.x + (double)v.y) + ((double)v.z + (double)v.w);
  }
  block_sum<NT, 1>(acc, partials + blockIdx.x);
}

int launch_sum_f(Ctx* c, const float* x, long n, int nblk)
{
  if (n % 4 == 0 && reinterpret_cast<uintptr_t>(x) % 16 == 0)
    hipLaunchKernelGGL((k_sum_f4<256>), dim3(nblk), dim3(256), 0, c->stream, reinterpret_cast<const float4*>(x), n / 4,
                       c->partials);
  else
    hipLaunchKernelGGL((k_sum_f<256>), dim3(nblk), dim3(256), 0, c->stream, x, n, c->partials);
  MXR_HIP(hipGetLastError());
  return MXR_OK;
}

template <int N>
struct Hii : Passes<N> {
  using C = Cfg<N>;

  template <int EPI>
  static int c2r(Ctx* c, CGridPtrs in, int ngrid, FiltArgs fa, const HiiEpi& ea)
  {
    const long nrows = (long)N * N;
    dim3 grid((unsigned)((nrows + C::TR - 1) / C::TR), (unsigned)(EPI == C2R_STORE ? ngrid : 1));
    if constexpr (epi_spin(EPI) || epi_bh(EPI)) {
      if (ea.last)
        hipLaunchKernelGGL((k_row_c2r_spin<N, C::ROW_NT, C::TR, EPI, 1>), grid, dim3(C::ROW_NT), 0, c->stream, in,
                           ngrid, nrows, c->tw, fa, ea);
      else
        hipLaunchKernelGGL((k_row_c2r_spin<N, C::ROW_NT, C::TR, EPI, 0>), grid, dim3(C::ROW_NT), 0, c->stream, in,
                           ngrid, nrows, c->tw, fa, ea);
    } else if constexpr (EPI == C2R_STORE) {
      hipLaunchKernelGGL((k_row_c2r<N, C::ROW_NT, C::TR, EPI, 0>), grid, dim3(C::ROW_NT), 0, c->stream, in, ngrid,
                         nrows, c->tw, fa, ea);
    } else {
      if (ea.last)
        hipLaunchKernelGGL((k_row_c2r<N, C::ROW_NT, C::TR, EPI, 1>), grid, dim3(C::ROW_NT), 0, c->stream, in, ngrid,
                           nrows, c->tw, fa, ea);
      else
        hipLaunchKernelGGL((k_row_c2r<N, C::ROW_NT, C::TR, EPI, 0>), grid, dim3(C::ROW_NT), 0, c->stream, in, ngrid,
                           nrows, c->tw, fa, ea);
    }
    MXR_HIP(hipGetLastError());
    return MXR_OK;
  }

  // direct-summation z pass for a radius keeping kc leading k_z (kc <= 64, N a multiple of 256)
  template <int EPI>
  static int c2r_direct(Ctx* c, CGridPtrs in, int kc, int rpb, const HiiEpi& ea)
  {
    if constexpr (N % 256 != 0 || N / 2 + 1 <= 64) {
      return MXR_ERR_UNSUPPORTED;
    } else {
      if (ea.last) {  // the kernel compiles the last radius's epilogue out
        set_error("direct z pass called for the last (unfiltered) radius");
        return MXR_ERR_UNSUPPORTED;
      }
      const long nrows = (long)N * N;
      constexpr int SEGS = N / 4 / 64;
      const long groups = (nrows + rpb - 1) / rpb;  // rounded up to whole sets of 8 below
      const dim3 grid((unsigned)(SEGS * 8 * ((groups + 7) / 8)));
      switch ((kc + 7) / 8) {
#define L(m)                                                                                                         \
  case m:                                                                                                            \
    hipLaunchKernelGGL((k_row_direct<N, 8 * m, EPI>), grid, dim3(64), 0, c->stream, in, nrows, rpb, c->tw, kc, ea);  \
    break;
        L(1) L(2) L(3) L(4) L(5) L(6) L(7) L(8)
#undef L
        default: return MXR_ERR_UNSUPPORTED;
      }
      MXR_HIP(hipGetLastError());
      return MXR_OK;
    }
  }
};



int ensure_buffers(Ctx* c, int ngrid)
{
  const long Ns = n_spec(c->N);
  for (int g = 0; g < ngrid; g++) {
    if (!c->spec[g])
      MXR_HIP(hipMalloc(&c->spec[g], sizeof(float2) * Ns));
    if (!c->work[g])
      MXR_HIP(hipMalloc(&c->work[g], sizeof(float2) * Ns));
  }
  if (!c->ion)
    MXR_HIP(hipMalloc(&c->ion, (size_t)n_real(c->N)));
  // per-block FP64 partials: up to 17 values for each of at most 8192 blocks (ew_blocks)
  const size_t need = 17 * 8192 + 6 * (size_t)(((long)c->N * c->N + 15) / 16 + 16);
  if (c->partials_cap < need) {
    (void)hipFree(c->partials);
    c->partials = nullptr;
    MXR_HIP(hipMalloc(&c->partials, sizeof(double) * need));
    c->partials_cap = need;
  }
  return MXR_OK;
}

static int stage_in(Ctx* c, int slot, const float* host, size_t n, const float** dev)
{
  if (!c->in_stage[slot])
    MXR_HIP(hipMalloc(&c->in_stage[slot], sizeof(float) * 2 * n_complex(c->N)));
  MXR_HIP(hipMemcpyAsync(c->in_stage[slot], host, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
  *dev = c->in_stage[slot];
  return MXR_OK;
}

static int stage_out(Ctx* c, int slot, const float* host, size_t n, float** dev)
{
  if (!c->out_stage[slot])
    MXR_HIP(hipMalloc(&c->out_stage[slot], sizeof(float) * 2 * n_complex(c->N)));
  MXR_HIP(hipMemcpyAsync(c->out_stage[slot], host, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
  *dev = c->out_stage[slot];
  return MXR_OK;
}



// dtdz(float z) (XRayHeatingFunctions.c:426-440): dt/dz of a flat matter + Lambda universe, in s;
// 1 + z is a float sum, as the reference's argument is a float
static double hii_dtdz(const mxr_hii_params* p, float z)
{
  const double OMl = p->OmegaLambda, OMm = p->OmegaM;
  const double x = std::sqrt(OMl / OMm) * std::pow(1 + z, -3.0 / 2.0);
  const double dxdz = std::sqrt(OMl / OMm) * std::pow(1 + z, -5.0 / 2.0) * (-3.0 / 2.0);
  const double const1 = 2 * std::sqrt(1 + OMm / OMl) / (3.0 * HUBBLE * p->Hubble_h);
  const double numer = dxdz * (1 + x * std::pow(std::pow(x, 2) + 1, -0.5));
  const double denom = x + std::sqrt(std::pow(x, 2) + 1);
  return const1 * numer / denom;
}

template <int N>
static int hii_run(Ctx* c, const mxr_hii_params* p, const mxr_hii_grids* g, mxr_hii_globals* out)
{
  using K = Hii<N>;
  constexpr int H = N / 2 + 1;
  const long Nc = n_complex(N), nreal = n_real(N);
  const int spin = p->Flag_IncludeSpinTemp ? 1 : 0;
  const bool bf16 = g->storage == MXR_STORAGE_BF16;  // stars and weighted_sfr as bfloat16
  if (bf16 && c->path == MXR_PATH_PLAIN) {
    set_error("BF16 storage is implemented on the fused paths only");
    return MXR_ERR_UNSUPPORTED;
  }
  const int bh = p->Flag_BHFeedback ? 1 : 0;
  const int rec = p->Flag_IncludeRecombinations ? 1 : 0;
  // grids: deltax, stars, weighted_sfr, [x_e], [effective_bhm, effective_bhar], [N_rec]
  const int ngrid = 3 + spin + 2 * bh + rec;
  const int ixe = 3, ibhm = 3 + spin, ibhar = 4 + spin, inrec = ngrid - 1;
  if (rec && (!g->N_rec || !g->Gamma12 || !g->residual_xH || !g->clumping_factor || !g->t_resp)) {
    set_error("Flag_IncludeRecombinations needs N_rec, Gamma12, residual_xH, clumping_factor and t_resp");
    return MXR_ERR_INVALID_ARG;
  }
  RecTab rt{};
  if (rec)
    RC(rec_tables(c, p, &rt));
  RC(ensure_buffers(c, ngrid));

  // ---- inputs and state on the device
  const float* din[MAX_GRIDS] = {g->deltax, g->stars, g->weighted_sfr, nullptr, nullptr, nullptr};
  if (spin)
    din[ixe] = g->x_e_box;
  if (bh) {
    din[ibhm] = g->effective_bhm;
    din[ibhar] = g->effective_bhar;
  }
  const float* dTk = spin ? g->Tk_box : nullptr;
  // state: xH, z_at_ionization, J_21, r_bubble, temp, and with recombinations N_rec (padded, in/out),
  // Gamma12 (in/out), residual_xH, clumping_factor, t_resp
  const int nst = rec ? 10 : 5;
  float* dst[10] = {g->xH, g->z_at_ionization, g->J_21_at_ionization, g->r_bubble, g->temp_kinetic_all_gas,
                    g->N_rec, g->Gamma12, g->residual_xH, g->clumping_factor, g->t_resp};
  if (g->mem == MXR_MEM_HOST) {
    for (int i = 0; i < ngrid - rec; i++)
      RC(stage_in(c, i, din[i], (bf16 && (i == 1 || i == 2)) ? Nc : 2 * Nc, &din[i]));  // bf16: half the bytes
    if (spin)
      RC(stage_in(c, 6, dTk, nreal, &dTk));
    for (int i = 0; i < nst; i++)
      RC(stage_out(c, i, dst[i], i == 5 ? 2 * Nc : nreal, &dst[i]));
  }
  if (rec)
    din[inrec] = dst[5];

  // ---- scalars (find_HII_bubbles.c:90-148)
  const double box_size = p->BoxSize;
  const double pixel_volume = std::pow(box_size / (double)N, 3);
  const double total_n_cells = std::pow((double)N, 3);
  const double UL = p->UnitLength_in_cm, UM = p->UnitMass_in_g;
  const double UT = p->UnitLength_in_cm / p->UnitVelocity_in_cm_per_s;
  const double sfr_timescale = p->ReionSfrTimescale * hubble_time(p);
  const double redshift = p->redshift;
  double TKd = p->TK_recfast;
  if (TKd < 0)
    TKd = T_RECFAST(p->TablesForXHeatingDir, (float)redshift);
  if (TKd < 0)
    return MXR_ERR_TABLES;
  const int nR = mxr_hii_radii(p, nullptr, 0);
  std::vector<double> radii(nR);
  mxr_hii_radii(p, radii.data(), nR);

  HiiEpi ea{};
  ea.z = (float)redshift;
  ea.TK = (float)TKd;
  ea.cT_ad = cT_approx((float)redshift);
  ea.fi_z7 = std::pow((1. + (double)ea.z) / 7.1, 2.5);
  ea.fi_tre = std::pow(T_RE, 1.7);
  ea.fi_he = std::pow(1e4 * ((1. + (double)ea.z) / 4.), 1.7);
  ea.uvb = p->ReionUVBFlag;
  ea.spin = spin;
  ea.bh = bh;
  ea.bhar = bh ? din[ibhar] : nullptr;
  ea.ion = c->ion;
  ea.xH = dst[0];
  ea.z_in = dst[1];
  ea.J21 = dst[2];
  ea.r_bubble = dst[3];
  ea.temp = dst[4];
  ea.deltax = din[0];
  ea.sfr = din[2];
  ea.sfr_bf16 = bf16 ? 1 : 0;
  ea.Tk_box = dTk;
  ea.partials = c->partials;
  ea.rec = rec;
  if (rec) {
    ea.temp_dep = p->Flag_TemperatureDependentRec ? 1 : 0;
    ea.redshift = redshift;
    ea.hubble_h = p->Hubble_h;
    // find_HII_bubbles.c:128-130: the snapshot step and |dt/dz| / h in float
    ea.zstep = (float)(p->prev_redshift - redshift);
    ea.fabs_dtdz = (float)std::fabs(hii_dtdz(p, (float)redshift) / p->Hubble_h);
    ea.Gamma12 = dst[6];
    ea.N_rec = dst[5];
    ea.residual_xH = dst[7];
    ea.clumping = dst[8];
    ea.t_resp = dst[9];
    ea.rt = rt;
  }

  FiltArgs fa{};
  fa.N = N;
  {
    const float bs = (float)box_size;  // reionization.c:3376-3377
    fa.dk = (float)(2.0 * M_PI / bs);
  }
  fa.ftype = p->ReionFilterType;
  fa.prune = (c->path == MXR_PATH_FUSED && p->ReionFilterType == 1) ? 1 : 0;  // only sharp-k zeroes modes

  CRealPtrs rin{};
  GridPtrs spec{}, work{};
  CGridPtrs cspec{}, cwork{};
  for (int i = 0; i < ngrid; i++) {
    rin.p[i] = din[i];
  rin.bf16 = bf16 ? (1u << 1) | (1u << 2) : 0u;
    spec.p[i] = c->spec[i];
    work.p[i] = c->work[i];
    cspec.p[i] = c->spec[i];
    cwork.p[i] = c->work[i];
  }
  const bool plain = (c->path == MXR_PATH_PLAIN);

  // timed launch: event pair around fn() when profiling, plus its algorithmic bytes
  auto T = [&](int cls, double bytes, auto&& fn) -> int {
    const int ev = prof_begin(c);
    if (ev < -1)
      return ev;
    const int rc = fn();
    if (rc)
      return rc;
    return prof_end(c, ev, cls, bytes);
  };
  const double cplx = 8.0 * (double)Nc;  // bytes of one half spectrum
  const double G = (double)ngrid;

  if (!c->enqueue_only)
    MXR_HIP(hipEventRecord(c->ev0, c->stream));
  // ---- forward transforms (:163-253)
  FiltArgs none{};
  none.N = N;
  RC(T(K_FWD_R2C, G * 2 * cplx, [&] { return K::r2c(c, rin, spec, ngrid); }));
  RC(T(K_FWD_Y, G * 2 * cplx, [&] { return K::template col<-1, COL_PLAIN>(c, cspec, spec, ngrid, K::ygeo(), none, 1.0); }));
  if (plain) {
    RC(T(K_FWD_X, G * 2 * cplx,
         [&] { return K::template col<-1, COL_PLAIN>(c, cspec, spec, ngrid, K::xgeo(), none, 1.0); }));
    for (int i = 0; i < ngrid; i++)
      RC(T(K_SCALE, 2 * cplx, [&] {
        hipLaunchKernelGGL(k_scale, dim3(ew_blocks(n_spec(N))), dim3(256), 0, c->stream, c->spec[i], n_spec(N),
                           total_n_cells);
        MXR_HIP(hipGetLastError());
        return MXR_OK;
      }));
  } else {
    RC(T(K_FWD_X, G * 2 * cplx,
         [&] { return K::template col<-1, COL_SCALE>(c, cspec, spec, ngrid, K::xgeo(), none, total_n_cells); }));
  }
  // the criterion runs as its own kernel on the stored fields (the plain path, and every path with
  // recombinations, whose first-crossing update reads weighted_sfr and N_rec too): xH and r_bubble
  // start at 1 and 0 as the reference resets them (find_HII_bubbles.c:150-158)
  const bool split = plain || rec;
  if (split)
    for (int f = 0; f < 2; f++)
      RC(T(K_SCALE, 4.0 * nreal, [&] {
        hipLaunchKernelGGL(k_fill, dim3(ew_blocks(nreal)), dim3(256), 0, c->stream, dst[f == 0 ? 0 : 3],
                           f == 0 ? 1.0f : 0.0f, nreal);
        MXR_HIP(hipGetLastError());
        return MXR_OK;
      }));

  // ---- radius loop (:279-547)
  for (int ri = 0; ri < nR; ri++) {
    const double R = radii[ri];
    const bool last = (ri == nR - 1);
    const double M_mean = mxr_RtoM(p, R);
    double Gamma_R_prefactor = (1.0 + redshift) * (1.0 + redshift) * R * UL / pixel_volume * std::pow(UL, -3.);
    Gamma_R_prefactor *= (UM / UT) / PROTONMASS * p->ReionNionPhotPerBary;
    ea.J21c = (float)(Gamma_R_prefactor * PLANCK * 1e21 / (4.0 * M_PI) * p->ReionGammaHaloBias / sfr_timescale *
                      p->ReionAlphaUV);
    double Gamma_R_prefactor_BH = Gamma_R_prefactor;  // :429-433
    ea.J21c_bh = (float)(Gamma_R_prefactor_BH * PLANCK * 1e21 / (4.0 * M_PI) * p->ReionGammaHaloBias /
                         sfr_timescale * p->ReionAlphaUVBH);
    // :438-441: the ionising-background prefactors, used at first crossings with recombinations
    Gamma_R_prefactor *= SIGMA_HI * p->ReionAlphaUV / (p->ReionAlphaUV + 2.75) * 1e12;
    Gamma_R_prefactor_BH *= SIGMA_HI * p->ReionAlphaUVBH / (p->ReionAlphaUVBH + 2.75) * 1e12;
    ea.gpref = (float)Gamma_R_prefactor;
    ea.gpref_bh = (float)Gamma_R_prefactor_BH;
    ea.R = (float)R;
    ea.f_coll_prefactor = p->ReionEfficiency * (4.0 / 3.0) * M_PI / pixel_volume / M_mean * (R * R * R);
    ea.first = (ri == 0);
    ea.last = last;
    fa.R = (float)R;
    fa.apply = last ? 0 : 1;

    if (!plain) {
      // kept-mode counts of this radius, for the profiler's byte counts only
      KeptModes km = (c->prof && fa.apply && fa.prune && fa.ftype == 1) ? kept_modes(N, fa.dk, fa.R)
                                                                        : KeptModes::all(N);
      // flags read (+ written on the first radius); on the last radius xH, r_bubble and temp
      // written and the deltax / temp (or Tk_box) needed for T_HI read
      const double state_b = last ? (double)nreal * (1 + 4 + 4 + 4 + 4 + 4) : (double)nreal * 1.0;
      const int kc = kept_kz(fa);
      const bool direct = !rec && (N % 256 == 0) && kc < H && kc <= c->zdirect_kmax && kc <= 64;
      // live-tile grids for the x and y passes when the window cuts k_z (the kept columns then
      // have |n_y| < kc and k_z < kc)
      fa.live = (kc < H && c->live_tiles) ? kc : 0;
      RC(T(K_INV_X, G * 8.0 * ((double)km.k3 + (double)N * km.c2),
           [&] { return K::template col<+1, COL_FILTER>(c, cspec, work, ngrid, K::xgeo(), fa, 1.0, c->col_tc_x); }));
      RC(T(K_INV_Y, G * 8.0 * ((double)N * km.c2 + (double)N * N * km.k1),
           [&] { return K::template col<+1, COL_YPRUNE>(c, cwork, work, ngrid, K::ygeo(), fa, 1.0, c->col_tc_y); }));
      if (rec) {
        // every filtered field stored as a real grid (kept k_z only), then the criterion kernel
        RC(T(K_INV_Z, G * 8.0 * (double)N * N * km.k1 + G * 4.0 * nreal,
             [&] { return K::template c2r<C2R_STORE>(c, cwork, ngrid, fa, ea); }));
      } else {
        RC(T(K_INV_Z, G * 8.0 * (double)N * N * km.k1 + state_b, [&] {
          if (direct)
            return bh ? (spin ? K::template c2r_direct<C2R_HII_SPIN_BH>(c, cwork, kc, c->zdirect_rpb, ea)
                              : K::template c2r_direct<C2R_HII_BH>(c, cwork, kc, c->zdirect_rpb, ea))
                      : (spin ? K::template c2r_direct<C2R_HII_SPIN>(c, cwork, kc, c->zdirect_rpb, ea)
                              : K::template c2r_direct<C2R_HII>(c, cwork, kc, c->zdirect_rpb, ea));
          return bh ? (spin ? K::template c2r<C2R_HII_SPIN_BH>(c, cwork, ngrid, fa, ea)
                            : K::template c2r<C2R_HII_BH>(c, cwork, ngrid, fa, ea))
                    : (spin ? K::template c2r<C2R_HII_SPIN>(c, cwork, ngrid, fa, ea)
                            : K::template c2r<C2R_HII>(c, cwork, ngrid, fa, ea));
        }));
      }
    } else {
      for (int i = 0; i < ngrid; i++) {
        RC(T(K_COPY, 2 * cplx, [&] {
          MXR_HIP(hipMemcpyAsync(c->work[i], c->spec[i], sizeof(float2) * n_spec(N), hipMemcpyDeviceToDevice,
                                 c->stream));
          return MXR_OK;
        }));
        if (!last)
          RC(T(K_FILTER, 2 * cplx, [&] {
            hipLaunchKernelGGL(k_filter, dim3(ew_blocks(n_spec(N))), dim3(256), 0, c->stream, c->work[i], 0, N, fa,
                               spec_pitch(N));
            MXR_HIP(hipGetLastError());
            return MXR_OK;
          }));
      }
      for (int i = 0; i < ngrid; i++) {
        CGridPtrs ci{};
        GridPtrs wi{};
        ci.p[0] = c->work[i];
        wi.p[0] = c->work[i];
        RC(T(K_INV_X, 2 * cplx, [&] { return K::template col<+1, COL_PLAIN>(c, ci, wi, 1, K::xgeo(), none, 1.0); }));
        RC(T(K_INV_Y, 2 * cplx, [&] { return K::template col<+1, COL_PLAIN>(c, ci, wi, 1, K::ygeo(), none, 1.0); }));
        RC(T(K_INV_Z, 2 * cplx, [&] { return K::template c2r<C2R_STORE>(c, ci, 1, none, ea); }));
      }
    }
    if (split) {
      // the aliasing clamp of each stored field (:372-422), then the criterion and state update
      for (int i = 0; i < ngrid; i++)
        RC(T(K_CLAMP, 8.0 * nreal, [&] {
          hipLaunchKernelGGL(k_clamp, dim3(ew_blocks(nreal)), dim3(256), 0, c->stream, (float*)c->work[i], N,
                             i == 0 ? 0 : ((spin && i == ixe) ? 2 : 1));
          MXR_HIP(hipGetLastError());
          return MXR_OK;
        }));
      PlainArgs pa{N, (const float*)c->work[0], (const float*)c->work[1],
                   spin ? (const float*)c->work[ixe] : nullptr, bh ? (const float*)c->work[ibhm] : nullptr,
                   rec ? (const float*)c->work[2] : nullptr, (rec && bh) ? (const float*)c->work[ibhar] : nullptr,
                   rec ? (const float*)c->work[inrec] : nullptr};
      RC(T(K_CRIT, (double)nreal * (4 + 4 + (spin ? 4 : 0) + 4 + 4 + 4 + (rec ? 8 : 0)), [&] {
        hipLaunchKernelGGL(k_criterion, dim3(ew_blocks(nreal)), dim3(256), 0, c->stream, pa, ea);
        MXR_HIP(hipGetLastError());
        return MXR_OK;
      }));
    }
  }

  // ---- reductions (:549-664)
  // post-loop temperatures and FP64 sums: xH, deltax, z_in, temp, weighted_sfr read, temp written
  const int nblk = (int)ew_blocks(nreal);
  // with recombinations also Gamma12, N_rec and r_bubble read, N_rec, residual_xH,
  // clumping_factor and t_resp written
  const int nsum = rec ? 17 : 7;
  RC(T(K_POST, (double)nreal * (4 + 4 + 4 + 4 + 4 + 4 + (spin ? 4 : 0) + (bh ? 4 : 0) + (rec ? 7 * 4 : 0)), [&] {
    if (rec)
      hipLaunchKernelGGL((k_postloop<256, true>), dim3(nblk), dim3(256), 0, c->stream, N, ea);
    else
      hipLaunchKernelGGL((k_postloop<256, false>), dim3(nblk), dim3(256), 0, c->stream, N, ea);
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(256), 0, c->stream, c->partials, nblk, nsum, c->sums);
    MXR_HIP(hipGetLastError());
    return MXR_OK;
  }));
  if (!c->enqueue_only)
    MXR_HIP(hipEventRecord(c->ev1, c->stream));
  MXR_HIP(hipMemcpyAsync(c->sums_host, c->sums, sizeof(double) * nsum, hipMemcpyDeviceToHost, c->stream));
  c->fin_total_n_cells = total_n_cells;
  c->fin_sfr_conv = UM / UT * SEC_PER_YEAR / SOLAR_MASS;
  c->fin_n_radii = nR;
  c->fin_bh = bh;
  c->fin_rec = rec;
  if (c->enqueue_only)  // batched: the caller synchronises and calls hii_globals
    return MXR_OK;
  if (g->mem == MXR_MEM_HOST) {
    float* hdst[10] = {g->xH,     g->z_at_ionization, g->J_21_at_ionization, g->r_bubble,        g->temp_kinetic_all_gas,
                       g->N_rec, g->Gamma12,         g->residual_xH,        g->clumping_factor, g->t_resp};
    for (int i = 0; i < nst; i++)
      MXR_HIP(hipMemcpyAsync(hdst[i], dst[i], sizeof(float) * (i == 5 ? 2 * Nc : nreal), hipMemcpyDeviceToHost,
                             c->stream));
  }
  MXR_HIP(hipStreamSynchronize(c->stream));
  float ms = 0.f;
  MXR_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  RC(prof_resolve(c));
  hii_globals(c, out);
  out->device_ms = ms;
  return MXR_OK;
}

// globals of the last find_HII_bubbles call from its FP64 sums (find_HII_bubbles.c:661-708)
void hii_globals(const Ctx* c, mxr_hii_globals* out)
{
  const double* s = c->sums_host;
  std::memset(out, 0, sizeof(*out));
  out->volume_weighted_global_xH = s[0] / c->fin_total_n_cells;
  out->mass_weighted_global_xH = s[1] / s[2];
  out->volume_weighted_global_weighted_sfr = s[3] / c->fin_total_n_cells * c->fin_sfr_conv;
  out->volume_weighted_global_temp_kinetic_all_gas = s[4] / c->fin_total_n_cells;
  out->mass_weighted_global_temp_kinetic_all_gas = s[5] / s[2];
  out->volume_weighted_global_effective_bhar = c->fin_bh ? s[6] / c->fin_total_n_cells * c->fin_sfr_conv : 0.0;
  if (c->fin_rec) {  // :683-700
    out->volume_weighted_global_N_rec = s[7] / c->fin_total_n_cells;
    out->mass_weighted_global_N_rec = s[8] / s[2];
    out->volume_weighted_global_residual_xH = s[9] / c->fin_total_n_cells;
    out->mass_weighted_global_residual_xH = s[10] / s[2];
    out->volume_weighted_global_clumping_factor = s[11] / c->fin_total_n_cells;
    out->mass_weighted_global_clumping_factor = s[12] / s[2];
    out->volume_weighted_global_Gamma12 = s[13] / c->fin_total_n_cells;
    out->mass_weighted_global_Gamma12 = s[14] / s[2];
    out->volume_weighted_global_r_bubble = s[15] / c->fin_total_n_cells;
    out->mass_weighted_global_r_bubble = s[16] / s[2];
  }
  out->n_radii = c->fin_n_radii;
}

bool supported_dim(int N)
{
#define X(n) if (N == n) return true;
  MXR_DIMS(X)
#undef X
  return false;
}

int run_find_hii_bubbles(Ctx* c, const mxr_hii_params* p, const mxr_hii_grids* g, mxr_hii_globals* out)
{
  switch (c->N) {
#define X(n)                                                                                                         \
  case n:                                                                                                            \
    return hii_run<n>(c, p, g, out);
    MXR_DIMS(X)
#undef X
    default: set_error("unsupported grid dimension %d", c->N); return MXR_ERR_INVALID_ARG;
  }
}

// the window of types 0 and 2 by (|n_x|, |n_y|, n_z): the same window_mul evaluation a mode
// with those magnitudes gets
__global__ void k_window_table(float* __restrict__ tab, FiltArgs fa)
{
  const int H = fa.N / 2 + 1;
  const long n = (long)H * H * H;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += (long)gridDim.x * blockDim.x) {
    const int c = (int)(idx % H), b = (int)((idx / H) % H), a = (int)(idx / ((long)H * H));
    tab[idx] = window_mul(a, b, c, fa);
  }
}

int window_table(Ctx* c, float dk, float R, int ftype, const float** tab)
{
  if (!(c->wtab && c->wtab_ftype == ftype && c->wtab_dk == dk && c->wtab_R == R)) {
    const long H = c->N / 2 + 1, n = H * H * H;
    c->wtab_ftype = -1;
    if (!c->wtab)
      MXR_HIP(hipMalloc(&c->wtab, sizeof(float) * n));
    FiltArgs fa{};
    fa.N = c->N;
    fa.dk = dk;
    fa.R = R;
    fa.ftype = ftype;
    fa.apply = 1;
    hipLaunchKernelGGL(k_window_table, dim3(ew_blocks(n)), dim3(256), 0, c->stream, c->wtab, fa);
    MXR_HIP(hipGetLastError());
    c->wtab_ftype = ftype;
    c->wtab_dk = dk;
    c->wtab_R = R;
  }
  *tab = c->wtab;
  return MXR_OK;
}

int run_filter(Ctx* c, float2* box, int ix0, int nx, int N, float dk, float R, int ftype)
{
  FiltArgs fa{};
  fa.N = N;
  fa.dk = dk;
  fa.R = R;
  fa.ftype = ftype;
  fa.apply = 1;
  const long n = (long)nx * N * (N / 2 + 1);
  hipLaunchKernelGGL(k_filter, dim3(ew_blocks(n)), dim3(256), 0, c->stream, box, ix0, nx, fa, N / 2 + 1);
  MXR_HIP(hipGetLastError());
  MXR_HIP(hipStreamSynchronize(c->stream));
  return MXR_OK;
}

int run_c2r_store(Ctx* c, float2* spec)
{
  CGridPtrs cs{};
  cs.p[0] = spec;
  FiltArgs none{};
  none.N = c->N;
  HiiEpi ea{};
  switch (c->N) {
#define X(n)                                                                                                         \
  case n: return Hii<n>::template c2r<C2R_STORE>(c, cs, 1, none, ea);
    MXR_DIMS(X)
#undef X
    default: set_error("unsupported grid dimension %d", c->N); return MXR_ERR_INVALID_ARG;
  }
}

int run_velocity_gradient(Ctx* c, float2* box, int slab_nx, int grid_dim, float dk)
{
  const int H = grid_dim / 2 + 1;
  const long n = (long)slab_nx * grid_dim * H;
  hipLaunchKernelGGL(k_vgrad, dim3(ew_blocks(n)), dim3(256), 0, c->stream, box, n, H, dk, 1.0, 0);
  MXR_HIP(hipGetLastError());
  MXR_HIP(hipStreamSynchronize(c->stream));
  return MXR_OK;
}

// hubble(float z) (XRayHeatingFunctions.c:443-453): 1 + z is a float sum
static double tb_hubble(const mxr_tb_params* p, float z)
{
  const float opz = 1 + z;
  return HUBBLE * p->Hubble_h *
         std::sqrt(p->OmegaM * std::pow((double)opz, 3) + p->OmegaR * std::pow((double)opz, 4) + p->OmegaLambda);
}

// dv/dz of the (converted) velocity grid into c->spec[0] as a real grid (rows of 2 spec_pitch(N) floats)
// (BrightnessTemperature.c:200-215): r2c, /N^3 and x Hubble_h, x i k_z, c2r
template <int N>
static int tb_velocity_gradient(Ctx* c, const float* vel, const mxr_tb_params* p)
{
  using K = Hii<N>;
  CRealPtrs rin{};
  rin.p[0] = vel;
  GridPtrs spec{};
  spec.p[0] = c->spec[0];
  CGridPtrs cspec{};
  cspec.p[0] = c->spec[0];
  FiltArgs none{};
  none.N = N;
  RC(K::r2c(c, rin, spec, 1));
  RC((K::template col<-1, COL_PLAIN>(c, cspec, spec, 1, K::ygeo(), none, 1.0)));
  RC((K::template col<-1, COL_SCALE>(c, cspec, spec, 1, K::xgeo(), none, std::pow((double)N, 3))));
  const float box_size = (float)p->BoxSize;  // reionization.c:3438-3439
  const float dk = (float)(2.0 * M_PI / box_size);
  hipLaunchKernelGGL(k_vgrad, dim3(ew_blocks(n_spec(N))), dim3(256), 0, c->stream, c->spec[0], n_spec(N),
                     spec_pitch(N), dk, p->Hubble_h, 1);
  MXR_HIP(hipGetLastError());
  RC((K::template col<+1, COL_PLAIN>(c, cspec, spec, 1, K::xgeo(), none, 1.0)));
  RC((K::template col<+1, COL_PLAIN>(c, cspec, spec, 1, K::ygeo(), none, 1.0)));
  HiiEpi ea{};
  return K::template c2r<C2R_STORE>(c, cspec, 1, none, ea);
}

int run_brightness_temperature(Ctx* c, const mxr_tb_params* p, const mxr_tb_grids* g, double* ave)
{
  const int N = c->N;
  const long nreal = n_real(N), npad = 2 * n_complex(N);
  const int spin = p->Flag_IncludeSpinTemp ? 1 : 0;
  const int pecvel = spin ? p->Flag_IncludePecVelsFor21cm : 0;  // velocity terms need T_S
  RC(ensure_buffers(c, pecvel ? 2 : 0));
  const float* xH = g->xH;
  const float* dx = g->deltax;
  const float* TS = g->TS_box;
  float* vel = g->vel;
  float* dT = g->delta_T;
  if (g->mem == MXR_MEM_HOST) {
    RC(stage_in(c, 0, xH, nreal, &xH));
    RC(stage_in(c, 1, dx, npad, &dx));
    if (spin)
      RC(stage_in(c, 2, TS, nreal, &TS));
    if (pecvel)
      RC(stage_out(c, 1, vel, npad, &vel));
    if (!c->out_stage[0])
      MXR_HIP(hipMalloc(&c->out_stage[0], sizeof(float) * npad));
    dT = c->out_stage[0];
  }
  // BrightnessTemperature.c:47-55, 82
  const float redshift = (float)p->redshift;
  const float Hubble_h = (float)p->Hubble_h;
  const float OmegaM = (float)p->OmegaM;
  const float OmegaB = (float)(OmegaM * p->BaryonFrac);
  TbArgs a{};
  a.N = N;
  a.const_factor = (float)(27.0 * (OmegaB * Hubble_h * Hubble_h / 0.023) *
                           std::sqrt((0.15 / OmegaM / Hubble_h / Hubble_h) * (1.0 + redshift) / 10.0));
  a.T_rad = TCMB * (1. + redshift);
  a.spin = spin;
  a.tau = spin && p->Flag_IncludePecVelsFor21cm > 1;
  a.z = redshift;
  a.xH = xH;
  a.deltax = dx;
  a.TS = TS;
  a.dT = dT;
  a.partials = c->partials;
  const int nblk = (int)ew_blocks(nreal);
  hipLaunchKernelGGL((k_tb<256>), dim3(nblk), dim3(256), 0, c->stream, a);
  MXR_HIP(hipGetLastError());
  if (pecvel) {
    const float H_z = (float)(float)tb_hubble(p, redshift);
    hipLaunchKernelGGL(k_tb_velconv, dim3(ew_blocks(nreal)), dim3(256), 0, c->stream, vel, N, p->TreesID, redshift);
    MXR_HIP(hipGetLastError());
    int rc = MXR_ERR_UNSUPPORTED;
    switch (N) {
#define X(n)                                                                                                         \
  case n:                                                                                                            \
    rc = tb_velocity_gradient<n>(c, vel, p);                                                                         \
    break;
      MXR_DIMS(X)
#undef X
    }
    RC(rc);
    TbVelArgs va{};
    va.N = N;
    va.pecvel = pecvel;
    va.H_z = H_z;
    va.z = redshift;
    va.T_rad = a.T_rad;
    va.max_v_deriv = std::fabs(0.2f * H_z);  // fabs(MAX_DVDR * H_z), a float product
    va.grad = (const float*)c->spec[0];
    va.TS = TS;
    va.dT = dT;
    hipLaunchKernelGGL(k_tb_pecvel, dim3(ew_blocks(nreal)), dim3(256), 0, c->stream, va);
    MXR_HIP(hipGetLastError());
    if (pecvel == 3) {
      TbRsdArgs ra{};
      ra.N = N;
      ra.H_z = H_z;
      const double BoxSize = p->BoxSize / p->Hubble_h;
      ra.sw = (BoxSize / ((double)N)) / (double)N_RSD_STEPS;
      ra.cw = BoxSize / (float)N;
      for (int s = 0; s < N_RSD_STEPS; s++) {
        ra.x_off[s] = (float)(ra.sw * (float)s + ra.sw / 2.);
        ra.x_pos[s] = (float)(ra.x_off[s] / (BoxSize / (float)N));
      }
      ra.dT = dT;
      ra.vel = vel;
      ra.xH = xH;
      ra.los = (float*)c->work[1];
      const long rows = (long)N * N;
      hipLaunchKernelGGL(k_tb_rsd, dim3((unsigned)((rows + 63) / 64)), dim3(64), 0, c->stream, ra);
      MXR_HIP(hipGetLastError());
    }
    RC(launch_sum_f(c, dT, nreal, nblk));
  }
  hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(256), 0, c->stream, c->partials, nblk, 1, c->sums);
  MXR_HIP(hipGetLastError());
  MXR_HIP(hipMemcpyAsync(c->sums_host, c->sums, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (g->mem == MXR_MEM_HOST) {
    MXR_HIP(hipMemcpyAsync(g->delta_T, dT, sizeof(float) * nreal, hipMemcpyDeviceToHost, c->stream));
    if (pecvel)  // the reference converts the velocity grid in place
      MXR_HIP(hipMemcpyAsync(g->vel, vel, sizeof(float) * npad, hipMemcpyDeviceToHost, c->stream));
  }
  MXR_HIP(hipStreamSynchronize(c->stream));
  *ave = c->sums_host[0] / std::pow((double)N, 3);
  return MXR_OK;
}

}  // namespace mxr
