// mxr_ts.hip -- the spin-temperature step (_ComputeTs, src/core/ComputeTs.c:85-1554) on one
// gfx950 GPU.
//
// Per call, inhomogeneous branch (z < ReionMaxHeatingRedshift):
//   for each of the F shells R_ct (radius R_ct in Mpc, ComputeTs.c:339-746):
//     source   : when the history plan changes the SFR / X-ray source grids before this shell,
//                one element-wise launch applies the plan's float operations to both grids
//                (ComputeTs.c:366-478, load_reion_sfr_grids reionization.c:2750-2797);
//     forward  : only then r2c + y + x (/ (float)N^3); an unchanged source keeps its spectrum,
//                which the reference recomputes bit-identically every shell (:480-507);
//     inverse  : x-pass with the window on load (TsHeatingFilterType, R_ct > 0), y-pass, and a
//                z-pass c2r whose epilogue clamps the SFR at 0 and stores both smoothed fields
//                as FP32 [shell][cell] in HBM (:528-743) -- the reference's FP64 SMOOTHED_*
//                arrays hold exactly (double)float of these values times unit constants, so
//                FP32 storage loses nothing and halves the traffic;
//     R_ct = 0 : FP64 sums of the capped collapse fraction, x_e and the X-ray emissivity
//                (:585-683), read back while the remaining shells run, so the host builds the
//                frequency tables (mxr_ts_host.hip) concurrently with the GPU's FFTs;
//   evolve     : one pass per cell: table interpolation at x_e, the shell sums of evolveInt,
//                the forward-Euler x_e / T_K update and the T_S fixed point, with FP64 sums of
//                the volume averages (:1171-1505; XRayHeatingFunctions.c:1308-1561, 1639-2008).
// Homogeneous branch (:252-331): x_e = xion_RECFAST, T_K = T_RECFAST (1 + cT delta), T_S.
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "mxr_passes.hpp"
#include "mxr_ts.hpp"

namespace mxr {

// ---- source grids ---------------------------------------------------------------------------
// G source grids: 0 sfr, 1 xray_luminosity, then the AGN bands the call includes.  The AGN
// grids take the plan's loads and accumulations only; clearing and rescaling touch the first
// two (load_reion_bh_grids, reionization.c:2799-2870, runs beside load_reion_sfr_grids, and
// ComputeTs.c:366-384, 418-422 clear and rescale sfr and xray_luminosity alone).
constexpr int TS_MAXG = 4;
struct TsOpsArgs {
  int nops;
  int init;  // start from history[0] (the current snapshot's grid) instead of the source buffer
  TsOp ops[TS_MAX_OPS];
};
struct TsSrcArgs {
  const float* h[TS_MAXG];   // histories of each grid (history slot s at h[g] + s * hstride)
  const float* h0[TS_MAXG];  // the current snapshot's grid (history 0, wherever a ring keeps it)
  float* s[TS_MAXG];         // the source grids the shells transform
};

template <int G>
__global__ void k_ts_source(TsOpsArgs oa, TsSrcArgs sa, long hstride, long n)
{
  // ops in batches of OB: the batch's history loads are issued together (the op list is
  // uniform, so the branches are scalar), then the ops are applied in order
  constexpr int OB = 8;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    float v[G];
#pragma unroll
    for (int g = 0; g < G; g++)
      v[g] = oa.init ? sa.h0[g][i] : sa.s[g][i];
    for (int k0 = 0; k0 < oa.nops; k0 += OB) {
      float hv[OB][G];
#pragma unroll
      for (int u = 0; u < OB; u++) {
#pragma unroll
        for (int g = 0; g < G; g++)
          hv[u][g] = 0.0f;
        if (k0 + u < oa.nops) {
          const TsOp o = oa.ops[k0 + u];
          if (o.kind == TS_OP_LOAD || o.kind == TS_OP_ADD)
#pragma unroll
            for (int g = 0; g < G; g++)
              hv[u][g] = sa.h[g][(long)o.hist * hstride + i];
        }
      }
#pragma unroll
      for (int u = 0; u < OB; u++) {
        if (k0 + u >= oa.nops)
          break;
        const TsOp o = oa.ops[k0 + u];
#pragma unroll
        for (int g = 0; g < G; g++) {
          if (o.kind == TS_OP_CLEAR) {
            if (g < 2)
              v[g] = 0.0f;
          } else if (o.kind == TS_OP_SCALE) {
            if (g < 2)
              v[g] = (float)((double)v[g] * o.wd);
          } else if (o.kind == TS_OP_LOAD) {
            v[g] = mul_rn(hv[u][g], o.wf);
          } else {
            v[g] = add_rn(v[g], mul_rn(hv[u][g], o.wf));
          }
        }
      }
    }
#pragma unroll
    for (int g = 0; g < G; g++)
      sa.s[g][i] = v[g];
  }
}

// ---- inverse z-pass storing one smoothed shell ------------------------------------------------
struct TsStore {
  float* out[TS_MAXG];  // sfr and the AGN bands clamped at 0, xray as is; real layout
};

template <int N, int NT, int TR>
__global__ void __launch_bounds__(NT) k_row_c2r_ts(CGridPtrs in, long nrows, const float2* __restrict__ twg,
                                                   FiltArgs fa, TsStore st)
{
  constexpr int M = N / 2, H = N / 2 + 1, T = NT / TR, P = M + M / 8 + 1;
  using SL = Slots<M, T>;
  constexpr int NS = SL::QL * SL::RL;
  __shared__ float2 lds[TR * P];
  __shared__ float2 tw[N];
  const int c = threadIdx.x / T, t = threadIdx.x % T;
  const long row = (long)blockIdx.x * TR + c;
  const bool valid = row < nrows;
  for (int m = threadIdx.x; m < N; m += NT)
    tw[m] = twg[m];
  int kc = H;
  if (fa.apply && fa.prune) {
    kc = 0;
    for (int k0 = 0; k0 < H; k0 += NT) {
      const int k = k0 + (int)threadIdx.x;
      kc += __syncthreads_count(k < H && sharpk_keep(0, 0, k, fa));
    }
  }
  __syncthreads();
  const int g = blockIdx.y;
  float2 res[NS];
  row_c2r_grid<N, T, TR>(pick(in, g), valid, row, lds, tw, kc, c, t, res);
  if (!valid)
    return;
  float* __restrict__ dst = (g == 0 ? st.out[0] : g == 1 ? st.out[1] : g == 2 ? st.out[2] : st.out[3]) + row * N;
#pragma unroll
  for (int q = 0; q < SL::QL; q++) {
    const int j = t + T * q;
    if (SL::BL % T == 0 || j < SL::BL)
#pragma unroll
      for (int r = 0; r < SL::RL; r++) {
        const int n = j + r * SL::BL;
        float2 y = res[q * SL::RL + r];
        if (g != 1)
          y = make_float2(fmaxf(y.x, 0.0f), fmaxf(y.y, 0.0f));  // ComputeTs.c:592, 612-628, 703-735
        *reinterpret_cast<float2*>(dst + 2 * n) = y;
      }
  }
}

// ---- R_ct = 0 averages the tables need (ComputeTs.c:629-677): capped collapse fraction of
// the unfiltered stars and <x_e>, both from the inputs alone, so they run before any FFT
struct TsRed0 {
  int N;
  const float* stars;   // padded
  const float* deltax;  // padded
  const float* xe;      // padded (x_e_box_prev before the update)
  double K1, R03, pv;
  double* partials;
};

template <int NT>
__global__ void __launch_bounds__(NT) k_ts_red0(TsRed0 a)
{
#pragma clang fp contract(off)
  const int N = a.N, H = N / 2 + 1;
  const long n = (long)N * N * N;
  double acc[2] = {0.0, 0.0};
  for (long ir = (long)blockIdx.x * NT + threadIdx.x; ir < n; ir += (long)gridDim.x * NT) {
    const long ip = (ir / N) * 2 * H + ir % N;
    const double dom = 1.0 + (double)a.deltax[ip];
    double cf = (double)a.stars[ip] / (a.K1 * dom);
    cf = cf * (4.0 / 3.0);
    cf = cf * M_PI;
    cf = cf * a.R03;
    cf = cf / a.pv;
    if (cf > 1.0)
      cf = 1.0;
    acc[0] += cf;
    acc[1] += (double)a.xe[ip];
  }
  block_sum<NT, 2>(acc, a.partials + (long)blockIdx.x * 2);
}

// ---- microphysics (XRayHeatingFunctions.c), FP64, in the reference's evaluation order ---------
// natural cubic spline s of the device block (GSL's cspline form), knots kx[s][0..n-1]
__device__ __forceinline__ double spline_eval(const TsDev* __restrict__ T, int s, double xq)
{
#pragma clang fp contract(off)
  const int n = T->kn[s];
  int ilo = 0, ihi = n - 1;
  while (ihi > ilo + 1) {
    const int i = (ihi + ilo) / 2;
    if (T->kx[s][i] > xq)
      ihi = i;
    else
      ilo = i;
  }
  const double dx = T->kx[s][ilo + 1] - T->kx[s][ilo];
  const double dy = T->ky[s][ilo + 1] - T->ky[s][ilo];
  const double delx = xq - T->kx[s][ilo];
  const double ci = T->kc[s][ilo], ci1 = T->kc[s][ilo + 1];
  const double b = (dy / dx) - dx * (ci1 + 2.0 * ci) / 3.0;
  const double d = (ci1 - ci) / (3.0 * dx);
  return T->ky[s][ilo] + delx * (b + delx * (ci + delx * d));
}

// kappa_10 (:1728-1822): log-log spline, 1 K floor, T^0.381 above 10^4 K (that branch as one
// exp: exp(log kap_26 + 0.381 (log T - log T_26)) = kap_26 (T / T_26)^0.381)
__device__ __noinline__ double kappa_10(const TsDev* __restrict__ T, double lT)
{
#pragma clang fp contract(off)
  double ans;
  if (lT < T->kx[0][0])
    ans = T->ky[0][0];
  else if (lT > T->kx[0][KAP_N - 1])
    ans = T->ky[0][KAP_N - 1] + 0.381 * (lT - T->kx[0][KAP_N - 1]);
  else
    ans = spline_eval(T, 0, lT);
  return exp(ans);
}

// kappa_10_elec / kappa_10_pH (:1827-1966): log-log spline, floor below, linear extrapolation above
__device__ __noinline__ double kappa_tab(const TsDev* __restrict__ T, int s, double lT)
{
#pragma clang fp contract(off)
  const int n = T->kn[s];
  double ans;
  if (lT < T->kx[s][0])
    ans = T->ky[s][0];
  else if (lT > T->kx[s][n - 1])
    ans = T->ky[s][n - 1] +
          ((T->ky[s][n - 1] - T->ky[s][n - 2]) / (T->kx[s][n - 1] - T->kx[s][n - 2]) * (lT - T->kx[s][n - 1]));
  else
    ans = spline_eval(T, s, lT);
  return exp(ans);
}

// get_Ts (:1639-1684) with xcoll (:1686-1726), xalpha_tilde, taugp, Salpha_tilde, Tc_eff
// (:1973-2008).  z is the (float) redshift of the call, so Trad and (1+z)^3 come from the block.
// The fixed point is iterated on 1/T_S: with u = 1/T_S the reference's update
//   T_S' = (1 + xa + xc) / (1/Trad + xa / Tc_eff + xc / T_K),  1/Tc_eff = 1/T_K + 0.405535/T_K (u - 1/T_K)
// and S_alpha = (A + B u) / D need no division but the one per step.  Its loop condition
// |T_S' - T_S| / T_S' > 1e-3 is sign(u') |u - u'| / u > 1e-3 (u > 0 throughout): it goes on while
// u' > 0 and |u - u'| > 1e-3 u, and stops at the first T_S' <= 0 (whose absolute value, or an
// infinity at u' = 0, the reference returns) -- in very cold cells the update can cross zero.
__device__ __noinline__ float get_Ts(const TsDev* __restrict__ T, float delta, float TKf, float xef, float Jalpha,
                                     float* curr_xalpha)
{
  const double TK = TKf, xe = xef, dl = delta;
  const double iTK = 1.0 / TK;
  const double Trad = T->Trad;
  const double opz3 = T->opz3;
  const double lT = log(TK);
  const double c21 = T21 / Trad / A10_HYPERFINE * opz3 * (1.0 + dl);
  const double x_HI = c21 * ((1.0 - xe) * T->No) * kappa_10(T, lT);
  const double x_e = c21 * (xe * T->N_b0) * kappa_tab(T, 1, lT);
  const double x_p = c21 * (xe * T->No) * kappa_tab(T, 2, lT);
  const double xc = (x_HI + x_e) + x_p;
  double u;
  if (Jalpha > 1.0e-20) {
    const double tgp = T->c_tgp * (1.0 + dl) * (1.0 - xe);
    const double xi = cbrt(1.0e-7 * tgp * iTK * iTK);
    const double iden = 1.0 / (((1.0 + 2.98394 * xi) + (1.53583 * xi) * xi) + ((3.85289 * xi) * xi) * xi);
    const double A = ((1.0 - 0.0631789 * iTK) + 0.115995 * iTK * iTK) * iden;
    const double B = (0.336463 * iTK * iTK - 0.401403 * iTK) * iden;
    const double ka = T->c_xa * (double)Jalpha;
    const double base = T->iTrad + xc * iTK;
    double xa = 0.0;
    u = T->iTrad;
    for (int it = 0; it < TS_MAX_ITER; it++) {
      xa = ka * (A + B * u);
      const double iTce = iTK + 0.405535 * iTK * (u - iTK);
      const double un = (base + xa * iTce) / ((1.0 + xa) + xc);
      const bool more = un > 0.0 && fabs(u - un) > 1.0e-3 * u;
      u = un;
      if (!more)
        break;
    }
    *curr_xalpha = (float)xa;
  } else {
    u = (T->iTrad + xc * iTK) / (1.0 + xc);
    *curr_xalpha = 0.0f;
  }
  double TS = 1.0 / u;
  if (TS < 0.)
    TS = fabs(TS);
  return (float)TS;
}

// alpha_A (:1576-1584): case-A recombination coefficient, Abel et al. 1997 (the powers of log T
// as products)
__device__ __forceinline__ double alpha_A(double Tk)
{
#pragma clang fp contract(off)
  const double L = log(Tk / (double)1.1604505e4);
  const double L2 = L * L, L3 = L2 * L, L4 = L3 * L, L5 = L4 * L, L6 = L5 * L, L7 = L6 * L, L8 = L7 * L,
               L9 = L8 * L;
  const double e = -28.6130338 - 0.72411256 * L - 2.02604473e-2 * L2 - 2.38086188e-3 * L3 - 3.21260521e-4 * L4 -
                   1.42150291e-5 * L5 + 4.98910892e-6 * L6 + 5.75561414e-7 * L7 - 1.85676704e-8 * L8 -
                   3.07113524e-9 * L9;
  return exp(e);
}

// ---- per-cell evolve (ComputeTs.c:1171-1407, evolveInt :1308-1561) ------------------------------
struct TsEvolve {
  int N;
  const float* deltax;  // padded
  float* xe;            // padded, in/out (x_e_box_prev)
  float* Tk;            // real, in/out
  float* TS;            // real, out
  const float* sm;      // [2 + NB][F][N^3]: smoothed SFR, X-ray, then the NB AGN bands' shells
  double* partials;
};

// The shell sums run in FP64 over the folded per-shell factors staged in LDS (6 x 14 x F for
// the heat / ionisation / Lyman-alpha rates, F for the stellar Lyman-alpha flux, and with NB
// AGN bands NB x 6 x 14 x F more): per cell and shell 2 + NB coalesced FP32 loads and 7 + 6 NB
// FP64 FMAs.  The per-cell tail (Euler step, floors, T_S) follows the reference's evaluation
// order.  Launched with NB = 1 or 2 AGN bands (k_ts_evolve_g below serves the galaxy-only step,
// cell for cell the same arithmetic as NB = 0 here): the bands' tables take up to 130 KB of LDS
// at 64 shells, so one 768-thread workgroup per CU keeps three waves per SIMD.  For the
// galaxy-only form three waves per SIMD with 16 loads in flight measured 1.86 ms at 256^3 x 40
// shells, against 2.18 ms at two waves and 2.32 ms with 4 loads in flight.
template <int NT, int NB>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(3)))
k_ts_evolve(const TsDev* __restrict__ T, TsEvolve a)
{
#pragma clang fp contract(off)
  extern __shared__ double fl[];  // [6][XNH][F], bs[F], then [NB][6][XNH][F]
  const int N = a.N, H = N / 2 + 1, F = T->F;
  const long n = (long)N * N * N;
  for (int i = threadIdx.x; i < 6 * XNH * F; i += NT) {
    const int k = i / (XNH * F), m = (i / F) % XNH, r = i % F;
    fl[i] = T->fold[k][m][r];
  }
  for (int r = threadIdx.x; r < F; r += NT)
    fl[6 * XNH * F + r] = T->bs[r];
  if constexpr (NB > 0) {
    for (int i = threadIdx.x; i < NB * 6 * XNH * F; i += NT) {
      const int b = i / (6 * XNH * F), k = (i / (XNH * F)) % 6, m = (i / F) % XNH, r = i % F;
      fl[(6 * XNH + 1) * F + i] = T->fold_agn[b][k][m][r];
    }
  }
  __syncthreads();
  const double* __restrict__ bsl = fl + 6 * XNH * F;
  const double* __restrict__ fag = fl + (6 * XNH + 1) * F;
  const float* __restrict__ sm_sfr = a.sm;
  const float* __restrict__ sm_xr = a.sm + (long)F * n;
  const float* __restrict__ sm_agn = a.sm + 2L * F * n;  // band b, shell r at (b F + r) n
  const int no_light = T->no_light;
  // 0 J_alpha, 1 x_alpha, 2 Xheat, 3 Xion, 4 T_S, 5 T_K, 6 x_e, 7 HMXB emissivity, then per
  // AGN band its heating term and its R_ct = 0 emissivity
  constexpr int NV = 8 + 2 * NB;
  double acc[NV];
#pragma unroll
  for (int q = 0; q < NV; q++)
    acc[q] = 0.0;
  for (long ir = (long)blockIdx.x * NT + threadIdx.x; ir < n; ir += (long)gridDim.x * NT) {
    const long ip = (ir / N) * 2 * H + ir % N;
    const float dx = a.deltax[ip];
    const float xe0 = a.xe[ip];
    float tk = a.Tk[ir];
    // table interpolation at the cell's x_e (:1190-1287)
    double xcall = xe0;
    const double xtop = (double)T->xint[XNH - 1] * 0.999;
    if (xcall > xtop)
      xcall = xtop;
    else if (xcall < (double)T->xint[0])
      xcall = 1.001 * (double)T->xint[0];
    const float xcf = (float)xcall;
    int m = XNH - 1;
    while (m > 0 && xcf < T->xint[m])
      m--;
    const double dxl = xcall - (double)T->xint[m];
    acc[7] += (double)sm_xr[ir] * T->cx;  // <L_X> of the unfiltered shell (:603, 677-682)
    if constexpr (NB > 0) {
#pragma unroll
      for (int b = 0; b < NB; b++)  // <AGN emissivity> of the unfiltered shell (:612-628, 669-681)
        acc[8 + NB + b] += (double)sm_agn[(long)b * F * n + ir] * T->cagn;
    }
    double hs = 0, hb = 0, is = 0, ib = 0, ls = 0, lb = 0, st = 0;
    // AGN: heat per band (slope, intercept), ionisation and Lyman-alpha summed over the bands
    double ha[NB > 0 ? NB : 1][2], ia0 = 0, ia1 = 0, la0 = 0, la1 = 0;
#pragma unroll
    for (int b = 0; b < (NB > 0 ? NB : 1); b++)
      ha[b][0] = ha[b][1] = 0.0;
    if (!no_light) {
      const double* __restrict__ f0 = fl + (0 * XNH + m) * F;
      const double* __restrict__ f1 = fl + (1 * XNH + m) * F;
      const double* __restrict__ f2 = fl + (2 * XNH + m) * F;
      const double* __restrict__ f3 = fl + (3 * XNH + m) * F;
      const double* __restrict__ f4 = fl + (4 * XNH + m) * F;
      const double* __restrict__ f5 = fl + (5 * XNH + m) * F;
      // shells in batches of SB: a batch's loads are issued together (one round trip), then
      // its FMAs in shell order, so the sums are those of the plain loop
      constexpr int SB = NB == 0 ? 8 : NB == 1 ? 4 : 2;
      auto add = [&](int r, double xr, double sf) {
        hs = fma(xr, f0[r], hs);
        hb = fma(xr, f1[r], hb);
        is = fma(xr, f2[r], is);
        ib = fma(xr, f3[r], ib);
        ls = fma(xr, f4[r], ls);
        lb = fma(xr, f5[r], lb);
        st = fma(sf, bsl[r], st);
      };
      auto add_agn = [&](int b, int r, double xb) {
        const double* __restrict__ fb = fag + ((long)(b * 6) * XNH + m) * F + r;
        ha[b][0] = fma(xb, fb[0 * XNH * F], ha[b][0]);
        ha[b][1] = fma(xb, fb[1 * XNH * F], ha[b][1]);
        ia0 = fma(xb, fb[2 * XNH * F], ia0);
        ia1 = fma(xb, fb[3 * XNH * F], ia1);
        la0 = fma(xb, fb[4 * XNH * F], la0);
        la1 = fma(xb, fb[5 * XNH * F], la1);
      };
      int r = 0;
      for (; r + SB <= F; r += SB) {
        float xv[SB], sv[SB], av[NB > 0 ? NB : 1][SB];
#pragma unroll
        for (int u = 0; u < SB; u++) {
          xv[u] = sm_xr[(long)(r + u) * n + ir];
          sv[u] = sm_sfr[(long)(r + u) * n + ir];
          if constexpr (NB > 0) {
#pragma unroll
            for (int b = 0; b < NB; b++)
              av[b][u] = sm_agn[((long)b * F + r + u) * n + ir];
          }
        }
#pragma unroll
        for (int u = 0; u < SB; u++) {
          add(r + u, (double)xv[u], (double)sv[u]);
          if constexpr (NB > 0) {
#pragma unroll
            for (int b = 0; b < NB; b++)
              add_agn(b, r + u, (double)av[b][u]);
          }
        }
      }
      for (; r < F; r++) {
        add(r, (double)sm_xr[(long)r * n + ir], (double)sm_sfr[(long)r * n + ir]);
        if constexpr (NB > 0) {
#pragma unroll
          for (int b = 0; b < NB; b++)
            add_agn(b, r, (double)sm_agn[((long)b * F + r) * n + ir]);
        }
      }
    }
    const double x_e = xe0, Td = tk;
    const double zpd = (double)T->zpf;
    const double n_b = (T->N_b0 * T->nb3) * (double)add_rn(1.0f, dx);
    const double dheat = (hs * dxl + hb) * T->prefac_gal;
    double dion = (is * dxl + ib) * T->prefac_gal;
    double dlya = (ls * dxl + lb) * (T->prefac_gal * n_b);
    if constexpr (NB > 0) {  // the bands' prefactors are in their folds (XRayHeatingFunctions.c:1466-1472)
      dion = dion + (ia0 * dxl + ia1);
      dlya = dlya + (la0 * dxl + la1) * n_b;
    }
    const double dstar = st * T->conv;
    const double sink = alpha_A(Td) * CLUMPING_FACTOR * x_e * x_e * T->f_H * n_b;
    const double dxe_dzp = T->dt_dzp * (dion - sink);
    const double i1x = 1.0 / (1.0 + x_e);
    double dadia = T->c_adia;
    if (fabs((double)dx) > FRACT_FLOAT_ERR)
      dadia += T->dgrowth * (double)dx / (1.0 + T->growth * (double)dx);
    dadia *= (2.0 / 3.0) * Td;
    const double dspec = -dxe_dzp * Td * i1x;
    const double dcomp = T->c_comp * (x_e / ((1.0 + x_e) + T->f_He)) * (T->Trad - Td);
    double dxheat_dzp = dheat * T->c_heat * i1x;
    if constexpr (NB > 0) {  // deriv[5], deriv[6] (:1517-1529)
      double dsum = 0.0;
#pragma unroll
      for (int b = 0; b < NB; b++) {
        const double dA = (ha[b][0] * dxl + ha[b][1]) * T->c_heat * i1x;
        acc[8 + b] += dA;
        dsum = b == 0 ? dA : dsum + dA;
      }
      dxheat_dzp = dxheat_dzp + dsum;
    }
    const double d1 = ((dxheat_dzp + dcomp) + dspec) + dadia;
    const double d2 = dlya + dstar;
    const double d4 = T->dt_dzp * dion;
    // Euler step and floors (:1345-1377)
    float xe1 = (float)((double)xe0 + dxe_dzp * T->dzp);
    if (xe1 > 1)
      xe1 = (float)(1 - FRACT_FLOAT_ERR);
    else if (xe1 < 0)
      xe1 = 0;
    if (tk < MAX_TK)
      tk = (float)((double)tk + d1 * T->dzp);
    if (tk < MIN_TK)
      tk = (float)(TCMB * (1 + T->zp));
    float xa;
    const float ts = get_Ts(T, dx, tk, xe1, (float)d2, &xa);
    a.xe[ip] = xe1;
    a.Tk[ir] = tk;
    a.TS[ir] = ts;
    acc[0] += d2;
    acc[1] += (double)xa;
    acc[2] += dxheat_dzp;
    acc[3] += d4;
    acc[4] += (double)ts;
    acc[5] += (double)tk;
    acc[6] += (double)xe1;
  }
  block_sum<NT, NV>(acc, a.partials + (long)blockIdx.x * NV);
}

// Galaxy sources only (no AGN bands), CPT adjacent cells per thread (launched with CPT = 1):
// k_ts_evolve<NT, 0>'s arithmetic in the same order (shell sums as FMAs in shell order, then the
// same tail), so every cell's result is the same bit for bit, with the cell's prologue, shell loop
// and tail as separate phases: 1.74 ms at 256^3 x 40 shells against 1.82 ms.  Measured and not
// kept: two adjacent cells per thread with float2 shell loads (1.89 ms with 4 shells per batch,
// 1.97 ms with 8: the accumulators spill), a bucket lookup of the kappa splines' interval instead
// of the binary search (no change), and the kappa helpers inlined (1.79 ms).  Alone, the FP64
// tail (rates, Euler step, T_S) takes 1.08 ms and the shell loop with a trivial tail 1.26 ms.
template <int NT, int CPT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(3)))
k_ts_evolve_g(const TsDev* __restrict__ T, TsEvolve a)
{
#pragma clang fp contract(off)
  using VecF = typename std::conditional<CPT == 2, float2, float>::type;
  extern __shared__ double fl[];  // [6][XNH][F], bs[F]
  const int N = a.N, H = N / 2 + 1, F = T->F;
  const long n = (long)N * N * N;
  for (int i = threadIdx.x; i < 6 * XNH * F; i += NT) {
    const int k = i / (XNH * F), m = (i / F) % XNH, r = i % F;
    fl[i] = T->fold[k][m][r];
  }
  for (int r = threadIdx.x; r < F; r += NT)
    fl[6 * XNH * F + r] = T->bs[r];
  __syncthreads();
  const double* __restrict__ bsl = fl + 6 * XNH * F;
  const float* __restrict__ sm_sfr = a.sm;
  const float* __restrict__ sm_xr = a.sm + (long)F * n;
  const int no_light = T->no_light;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  // CPT cells share a row (N is even), so their padded indices are adjacent too
  for (long ir0 = ((long)blockIdx.x * NT + threadIdx.x) * CPT; ir0 < n; ir0 += (long)gridDim.x * NT * CPT) {
    const long ip0 = (ir0 / N) * 2 * H + ir0 % N;
    float dx[CPT], xe0[CPT], tk[CPT];
    int m[CPT];
    double dxl[CPT];
    double hs[CPT], hb[CPT], is[CPT], ib[CPT], ls[CPT], lb[CPT], st[CPT];
#pragma unroll
    for (int c = 0; c < CPT; c++) {
      dx[c] = a.deltax[ip0 + c];
      xe0[c] = a.xe[ip0 + c];
      tk[c] = a.Tk[ir0 + c];
      // table interpolation at the cell's x_e (:1190-1287)
      double xcall = xe0[c];
      const double xtop = (double)T->xint[XNH - 1] * 0.999;
      if (xcall > xtop)
        xcall = xtop;
      else if (xcall < (double)T->xint[0])
        xcall = 1.001 * (double)T->xint[0];
      const float xcf = (float)xcall;
      int mm = XNH - 1;
      while (mm > 0 && xcf < T->xint[mm])
        mm--;
      m[c] = mm;
      dxl[c] = xcall - (double)T->xint[mm];
      acc[7] += (double)sm_xr[ir0 + c] * T->cx;  // <L_X> of the unfiltered shell (:603, 677-682)
      hs[c] = hb[c] = is[c] = ib[c] = ls[c] = lb[c] = st[c] = 0.0;
    }
    if (!no_light) {
      constexpr int SB = 8;
      auto add = [&](int c, int r, double xr, double sf) {
        const double* __restrict__ f = fl + m[c] * F + r;
        hs[c] = fma(xr, f[0 * XNH * F], hs[c]);
        hb[c] = fma(xr, f[1 * XNH * F], hb[c]);
        is[c] = fma(xr, f[2 * XNH * F], is[c]);
        ib[c] = fma(xr, f[3 * XNH * F], ib[c]);
        ls[c] = fma(xr, f[4 * XNH * F], ls[c]);
        lb[c] = fma(xr, f[5 * XNH * F], lb[c]);
        st[c] = fma(sf, bsl[r], st[c]);
      };
      auto add_v = [&](int r, VecF xv, VecF sv) {
        if constexpr (CPT == 2) {
          add(0, r, (double)xv.x, (double)sv.x);
          add(1, r, (double)xv.y, (double)sv.y);
        } else {
          add(0, r, (double)xv, (double)sv);
        }
      };
      int r = 0;
      for (; r + SB <= F; r += SB) {
        VecF xv[SB], sv[SB];
#pragma unroll
        for (int u = 0; u < SB; u++) {
          xv[u] = *reinterpret_cast<const VecF*>(sm_xr + (long)(r + u) * n + ir0);
          sv[u] = *reinterpret_cast<const VecF*>(sm_sfr + (long)(r + u) * n + ir0);
        }
#pragma unroll
        for (int u = 0; u < SB; u++)
          add_v(r + u, xv[u], sv[u]);
      }
      for (; r < F; r++)
        add_v(r, *reinterpret_cast<const VecF*>(sm_xr + (long)r * n + ir0),
              *reinterpret_cast<const VecF*>(sm_sfr + (long)r * n + ir0));
    }
#pragma unroll
    for (int c = 0; c < CPT; c++) {
      const long ir = ir0 + c, ip = ip0 + c;
      const double x_e = xe0[c], Td = tk[c];
      const double n_b = (T->N_b0 * T->nb3) * (double)add_rn(1.0f, dx[c]);
      const double dheat = (hs[c] * dxl[c] + hb[c]) * T->prefac_gal;
      const double dion = (is[c] * dxl[c] + ib[c]) * T->prefac_gal;
      const double dlya = (ls[c] * dxl[c] + lb[c]) * (T->prefac_gal * n_b);
      const double dstar = st[c] * T->conv;
      const double sink = alpha_A(Td) * CLUMPING_FACTOR * x_e * x_e * T->f_H * n_b;
      const double dxe_dzp = T->dt_dzp * (dion - sink);
      const double i1x = 1.0 / (1.0 + x_e);
      double dadia = T->c_adia;
      if (fabs((double)dx[c]) > FRACT_FLOAT_ERR)
        dadia += T->dgrowth * (double)dx[c] / (1.0 + T->growth * (double)dx[c]);
      dadia *= (2.0 / 3.0) * Td;
      const double dspec = -dxe_dzp * Td * i1x;
      const double dcomp = T->c_comp * (x_e / ((1.0 + x_e) + T->f_He)) * (T->Trad - Td);
      const double dxheat_dzp = dheat * T->c_heat * i1x;
      const double d1 = ((dxheat_dzp + dcomp) + dspec) + dadia;
      const double d2 = dlya + dstar;
      const double d4 = T->dt_dzp * dion;
      // Euler step and floors (:1345-1377)
      float xe1 = (float)((double)xe0[c] + dxe_dzp * T->dzp);
      if (xe1 > 1)
        xe1 = (float)(1 - FRACT_FLOAT_ERR);
      else if (xe1 < 0)
        xe1 = 0;
      float t1 = tk[c];
      if (t1 < MAX_TK)
        t1 = (float)((double)t1 + d1 * T->dzp);
      if (t1 < MIN_TK)
        t1 = (float)(TCMB * (1 + T->zp));
      float xa;
      const float ts = get_Ts(T, dx[c], t1, xe1, (float)d2, &xa);
      a.xe[ip] = xe1;
      a.Tk[ir] = t1;
      a.TS[ir] = ts;
      acc[0] += d2;
      acc[1] += (double)xa;
      acc[2] += dxheat_dzp;
      acc[3] += d4;
      acc[4] += (double)ts;
      acc[5] += (double)t1;
      acc[6] += (double)xe1;
    }
  }
  block_sum<NT, 8>(acc, a.partials + (long)blockIdx.x * 8);
}

// homogeneous branch (ComputeTs.c:252-287) with the final averages
template <int NT>
__global__ void __launch_bounds__(NT) k_ts_homog(const TsDev* __restrict__ T, TsEvolve a)
{
#pragma clang fp contract(off)
  const int N = a.N, H = N / 2 + 1;
  const long n = (long)N * N * N;
  double acc[3] = {0, 0, 0};
  for (long ir = (long)blockIdx.x * NT + threadIdx.x; ir < n; ir += (long)gridDim.x * NT) {
    const long ip = (ir / N) * 2 * H + ir % N;
    const float dx = a.deltax[ip];
    const float xe = T->xe_h;
    const float tk = (float)((double)T->TK_h * (1.0 + (double)mul_rn(T->cT_ad, dx)));
    float xa;
    const float ts = get_Ts(T, dx, tk, xe, 0.0f, &xa);
    a.xe[ip] = xe;
    a.Tk[ir] = tk;
    a.TS[ir] = ts;
    acc[0] += (double)ts;
    acc[1] += (double)tk;
    acc[2] += (double)xe;
  }
  block_sum<NT, 3>(acc, a.partials + (long)blockIdx.x * 3);
}

// ---- driver -----------------------------------------------------------------------------------
static int ts_ensure(Ctx* c, int F, int G)
{
  const long npad = 2 * n_complex(c->N), nreal = n_real(c->N);
  RC(ensure_buffers(c, F > 0 ? G : 0));
  for (int g = 0; g < G && F > 0; g++)
    if (!c->ts_src[g])
      MXR_HIP(hipMalloc(&c->ts_src[g], sizeof(float) * npad));
  const size_t need = (size_t)G * F * nreal;
  if (F > 0 && c->ts_store_cap < need) {
    (void)hipFree(c->ts_store);
    c->ts_store = nullptr;
    c->ts_store_cap = 0;
    MXR_HIP(hipMalloc(&c->ts_store, sizeof(float) * need));
    c->ts_store_cap = need;
  }
  if (!c->ts_dev)
    MXR_HIP(hipMalloc(&c->ts_dev, sizeof(TsDev)));
  if (!c->ts_dev_host)
    MXR_HIP(hipHostMalloc(&c->ts_dev_host, sizeof(TsDev), hipHostMallocDefault));
  if (!c->ev_red)
    MXR_HIP(hipEventCreateWithFlags(&c->ev_red, hipEventDisableTiming));
  return MXR_OK;
}

template <int N>
static int ts_run(Ctx* c, const mxr_ts_params* p, const mxr_ts_grids* g, mxr_ts_globals* out, mxr_ts_tables* tables)
{
  using K = Passes<N>;
  constexpr int NT = 256;
  const long Nc = n_complex(N), nreal = n_real(N), npad = 2 * Nc;
  const double total_n_cells = std::pow((double)N, 3);

  TsPlan pl;
  RC(ts_plan(p, &pl));
  const int F = p->TsNumFilterSteps;
  // source grids: 0 sfr, 1 xray_luminosity, then the AGN soft and hard bands the flag includes
  const int NB = pl.homogeneous ? 0 : pl.agn_soft + pl.agn_hard;
  const int G = 2 + NB;
  const float* hist_in[TS_MAXG] = {g->sfr_histories, g->xray_luminosity_histories, nullptr, nullptr};
  {
    int b = 2;
    if (pl.agn_soft && !pl.homogeneous)
      hist_in[b++] = g->bh_xray_histories_soft;
    if (pl.agn_hard && !pl.homogeneous)
      hist_in[b++] = g->bh_xray_histories_hard;
  }
  if (!pl.homogeneous) {
    for (int k = 0; k < G; k++)
      if (!hist_in[k]) {
        set_error("ComputeTs: the %s histories are NULL",
                  k == 0 ? "sfr" : k == 1 ? "xray_luminosity" : "AGN X-ray emissivity");
        return MXR_ERR_INVALID_ARG;
      }
    if (g->n_histories <= pl.hist_max || g->n_histories < 1) {
      set_error("ComputeTs: the history plan reads history %d but %d histories were given", pl.hist_max,
                g->n_histories);
      return MXR_ERR_INVALID_ARG;
    }
    if (g->history_head < 0 || g->history_head >= g->n_histories) {
      set_error("ComputeTs: history_head %d is not a slot of %d", g->history_head, g->n_histories);
      return MXR_ERR_INVALID_ARG;
    }
  }
  RC(ts_ensure(c, pl.homogeneous ? 0 : F, G));
  TsDev* hd = reinterpret_cast<TsDev*>(c->ts_dev_host);
  const TsDev* dd = reinterpret_cast<const TsDev*>(c->ts_dev);

  auto T = [&](int cls, double bytes, auto&& fn) -> int {
    const int ev = prof_begin(c);
    if (ev < -1)
      return ev;
    const int rc = fn();
    if (rc)
      return rc;
    return prof_end(c, ev, cls, bytes);
  };

  // ---- inputs on the device
  const float* dx = g->deltax;
  const float* stars = g->stars;
  float* xe = g->x_e_box_prev;
  float* Tk = g->Tk_box;
  float* TS = g->TS_box;
  float* xe_out = g->x_e_box;
  const float* hist[TS_MAXG] = {hist_in[0], hist_in[1], hist_in[2], hist_in[3]};
  int hist_head = pl.homogeneous ? 0 : g->history_head;  // ring slot of history 0 on the device
  const int nstage = pl.homogeneous ? 0 : pl.hist_max + 1 > 1 ? pl.hist_max + 1 : 1;
  if (g->mem == MXR_MEM_HOST) {
    auto stage = [&](int slot, const float* host, long nfl, bool in) -> int {
      if (!c->in_stage[slot])
        MXR_HIP(hipMalloc(&c->in_stage[slot], sizeof(float) * npad));
      if (in)
        MXR_HIP(hipMemcpyAsync(c->in_stage[slot], host, sizeof(float) * nfl, hipMemcpyHostToDevice, c->stream));
      return MXR_OK;
    };
    RC(stage(0, dx, npad, true));
    dx = c->in_stage[0];
    if (!pl.homogeneous) {
      RC(stage(1, stars, npad, true));
      stars = c->in_stage[1];
    }
    RC(stage(2, xe, npad, true));
    xe = c->in_stage[2];
    RC(stage(3, Tk, nreal, !pl.homogeneous));
    Tk = c->in_stage[3];
    RC(stage(4, nullptr, 0, false));
    TS = c->in_stage[4];
    if (!c->out_stage[0])
      MXR_HIP(hipMalloc(&c->out_stage[0], sizeof(float) * npad));
    xe_out = c->out_stage[0];
    if (!pl.homogeneous) {
      const size_t need = (size_t)G * nstage * npad;
      if (c->ts_hist_cap < need) {
        (void)hipFree(c->ts_hist);
        c->ts_hist = nullptr;
        c->ts_hist_cap = 0;
        MXR_HIP(hipMalloc(&c->ts_hist, sizeof(float) * need));
        c->ts_hist_cap = need;
      }
      // staged in logical order (history s at s), whatever ring slot holds it
      for (int k = 0; k < G; k++) {
        for (int s = 0; s < nstage; s++) {
          const size_t src = (size_t)((g->history_head + s) % g->n_histories) * npad;
          MXR_HIP(hipMemcpyAsync(c->ts_hist + ((size_t)k * nstage + s) * npad, hist[k] + src, sizeof(float) * npad,
                                 hipMemcpyHostToDevice, c->stream));
        }
        hist[k] = c->ts_hist + (size_t)k * nstage * npad;
      }
      hist_head = 0;
    }
  }

  // the histories' stride: a staged copy holds nstage grids, a caller's array n_histories
  TsSrcArgs sa{};
  for (int k = 0; k < G && !pl.homogeneous; k++) {
    sa.h[k] = hist[k];
    sa.h0[k] = hist[k] + (size_t)hist_head * npad;
    sa.s[k] = c->ts_src[k];
  }
  const int nblk = (int)ew_blocks(nreal);
  double* sums = c->sums_host;
  MXR_HIP(hipEventRecord(c->ev0, c->stream));
  if (pl.homogeneous) {
    RC(ts_homogeneous(p, hd));
    MXR_HIP(hipMemcpyAsync(c->ts_dev, hd, sizeof(TsDev), hipMemcpyHostToDevice, c->stream));
    TsEvolve ea{N, dx, xe, Tk, TS, nullptr, c->partials};
    RC(T(K_TS_EVOLVE, (double)nreal * 16, [&] {
      hipLaunchKernelGGL((k_ts_homog<NT>), dim3(nblk), dim3(NT), 0, c->stream, dd, ea);
      MXR_HIP(hipGetLastError());
      return launch_final_sum(c, c->partials, nblk, 3, c->sums);
    }));
    MXR_HIP(hipMemcpyAsync(sums, c->sums, sizeof(double) * 3, hipMemcpyDeviceToHost, c->stream));
    if (tables) {
      std::memset(tables, 0, sizeof(*tables));
      tables->homogeneous = 1;
      tables->zp = pl.zp;
      tables->dzp = pl.dzp;
      tables->dt_dzp = pl.dt_dzp;
      tables->hist_max = -1;
    }
  } else {
    // R_ct = 0 averages first: the host builds the frequency tables on its own thread while
    // the GPU works through every shell
    {
      TsRed0 ra{};
      ra.N = N;
      ra.stars = stars;
      ra.deltax = dx;
      ra.xe = xe;
      const double h = p->Hubble_h;
      const double box_size = p->BoxSize / h;
      ra.K1 = (ts_rtom(p, pl.R[0]) * h) * h;
      ra.R03 = std::pow(pl.R[0], 3.0);
      ra.pv = std::pow(box_size / (double)N, 3);
      ra.partials = c->partials;
      RC(T(K_TS_RED, (double)nreal * 12, [&] {
        hipLaunchKernelGGL((k_ts_red0<NT>), dim3(nblk), dim3(NT), 0, c->stream, ra);
        MXR_HIP(hipGetLastError());
        return launch_final_sum(c, c->partials, nblk, 2, c->sums);
      }));
      MXR_HIP(hipMemcpyAsync(sums, c->sums, sizeof(double) * 2, hipMemcpyDeviceToHost, c->stream));
      MXR_HIP(hipEventRecord(c->ev_red, c->stream));
    }
    int trc = MXR_OK;
    std::string terr;
    std::thread tables_thread([&] {
      if (hipSetDevice(c->device) != hipSuccess || hipEventSynchronize(c->ev_red) != hipSuccess) {
        trc = MXR_ERR_HIP;
        terr = "waiting for the R_ct = 0 averages failed";
        return;
      }
      const double collapse_fraction = sums[0] / total_n_cells;
      const double x_e_ave = sums[1] / total_n_cells;
      g->stored_fcoll[p->snapshot] = collapse_fraction;
      trc = ts_tables(p, pl, collapse_fraction, x_e_ave, g->stored_fcoll, hd, tables);
      if (trc)
        terr = mxr_last_error();
    });
    auto shells = [&]() -> int {
      FiltArgs fa{};
      fa.N = N;
      {
        const float bs = (float)p->BoxSize;  // filter() uses run_globals.params.BoxSize (Mpc/h)
        fa.dk = (float)(2.0 * M_PI / bs);
      }
      fa.ftype = p->TsHeatingFilterType;
      fa.prune = p->TsHeatingFilterType == 1 ? 1 : 0;
      CRealPtrs rin{};
      GridPtrs spec{}, work{};
      CGridPtrs cspec{}, cwork{};
      for (int i = 0; i < G; i++) {
        spec.p[i] = c->spec[i];
        work.p[i] = c->work[i];
        cspec.p[i] = c->spec[i];
        cwork.p[i] = c->work[i];
      }
      FiltArgs none{};
      none.N = N;
      const double cplx = 8.0 * (double)Nc;
      bool src_valid = false;
      for (int r = 0; r < F; r++) {
        if (r == 0 || pl.nops[r] > 0) {
          if (pl.nops[r] > 0) {
            TsOpsArgs oa{};
            oa.nops = pl.nops[r];
            oa.init = src_valid ? 0 : 1;
            int nread = 1;
            for (int k = 0; k < pl.nops[r]; k++) {
              oa.ops[k] = pl.ops[r][k];
              if (oa.ops[k].kind >= TS_OP_LOAD)
                oa.ops[k].hist = (hist_head + oa.ops[k].hist) % g->n_histories;
              nread += (pl.ops[r][k].kind >= TS_OP_LOAD) ? 1 : 0;
            }
            RC(T(K_TS_SRC, G * 4.0 * npad * (nread + 1), [&] {
              if (G == 2)
                hipLaunchKernelGGL(k_ts_source<2>, dim3(ew_blocks(npad)), dim3(256), 0, c->stream, oa, sa, npad, npad);
              else if (G == 3)
                hipLaunchKernelGGL(k_ts_source<3>, dim3(ew_blocks(npad)), dim3(256), 0, c->stream, oa, sa, npad, npad);
              else
                hipLaunchKernelGGL(k_ts_source<4>, dim3(ew_blocks(npad)), dim3(256), 0, c->stream, oa, sa, npad, npad);
              MXR_HIP(hipGetLastError());
              return MXR_OK;
            }));
            src_valid = true;
          }
          for (int k = 0; k < G; k++)
            rin.p[k] = src_valid ? c->ts_src[k] : sa.h0[k];
          RC(T(K_TS_FWD, G * 3 * 2 * cplx, [&] {
            RC(K::r2c(c, rin, spec, G));
            RC((K::template col<-1, COL_PLAIN>(c, cspec, spec, G, K::ygeo(), none, 1.0)));
            return K::template col<-1, COL_FSCALE>(c, cspec, spec, G, K::xgeo(), none, total_n_cells);
          }));
        }
        fa.R = (float)pl.R[r];
        fa.apply = r > 0 ? 1 : 0;
        const KeptModes km = (c->prof && fa.apply && fa.prune) ? kept_modes(N, fa.dk, fa.R) : KeptModes::all(N);
        const int kc = kept_kz(fa);
        fa.live = (kc < N / 2 + 1 && c->live_tiles) ? kc : 0;  // live-tile x / y grids
        TsStore st{};
        for (int k = 0; k < G; k++)
          st.out[k] = c->ts_store + ((size_t)k * F + r) * nreal;
        RC(T(K_TS_INV_X, G * 8.0 * ((double)km.k3 + (double)N * km.c2),
             [&] { return K::template col<+1, COL_FILTER>(c, cspec, work, G, K::xgeo(), fa, 1.0, c->col_tc_x); }));
        RC(T(K_TS_INV_Y, G * 8.0 * ((double)N * km.c2 + (double)N * N * km.k1),
             [&] { return K::template col<+1, COL_YPRUNE>(c, cwork, work, G, K::ygeo(), fa, 1.0, c->col_tc_y); }));
        RC(T(K_TS_INV_Z, G * (8.0 * (double)N * N * km.k1 + 4.0 * nreal), [&] {
          const long nrows = (long)N * N;
          dim3 grid((unsigned)((nrows + Cfg<N>::TR - 1) / Cfg<N>::TR), (unsigned)G);
          hipLaunchKernelGGL((k_row_c2r_ts<N, Cfg<N>::ROW_NT, Cfg<N>::TR>), grid, dim3(Cfg<N>::ROW_NT), 0,
                             c->stream, cwork, nrows, c->tw, fa, st);
          MXR_HIP(hipGetLastError());
          return MXR_OK;
        }));
      }
      return MXR_OK;
    };
    const int src = shells();
    tables_thread.join();
    RC(src);
    if (trc) {
      set_error("%s", terr.c_str());
      return trc;
    }
    if (hd->nagn != NB) {
      set_error("ComputeTs: %d AGN bands in the tables, %d in the plan", hd->nagn, NB);
      return MXR_ERR_INVALID_ARG;
    }
    MXR_HIP(hipMemcpyAsync(c->ts_dev, hd, sizeof(TsDev), hipMemcpyHostToDevice, c->stream));
    TsEvolve ea{N, dx, xe, Tk, TS, c->ts_store, c->partials};
    const size_t lds = sizeof(double) * ((6 * XNH + 1) + (size_t)NB * 6 * XNH) * (size_t)F;
    const int nv = 8 + 2 * NB;
    // with AGN bands: 768-thread workgroups, one per CU (their tables fill most of the LDS)
    constexpr int NTA = 768;
    const int nblk_ev = NB == 0 ? nblk : (int)std::min<long>((nreal + NTA - 1) / NTA, 1024);
    RC(T(K_TS_EVOLVE,
         (double)nreal * (4 + 8 + 8 + 4 + (hd->no_light ? 4.0 * (1 + NB) : 4.0 * G * F)), [&] {
           if (NB == 0) {
             hipLaunchKernelGGL((k_ts_evolve_g<NT, 1>), dim3(nblk_ev), dim3(NT), lds, c->stream, dd, ea);
           } else {
             int lmax = 0;
             MXR_HIP(hipDeviceGetAttribute(&lmax, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
             if ((size_t)lmax < lds + 1024) {
               set_error("ComputeTs: the AGN bands' tables need %zu bytes of LDS at %d shells, the device has %d",
                         lds, F, lmax);
               return MXR_ERR_UNSUPPORTED;
             }
             if (NB == 1) {
               if (hipFuncSetAttribute((const void*)k_ts_evolve<NTA, 1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds) != hipSuccess)
                 (void)hipGetLastError();  // the launch below reports a size the device refuses
               hipLaunchKernelGGL((k_ts_evolve<NTA, 1>), dim3(nblk_ev), dim3(NTA), lds, c->stream, dd, ea);
             } else {
               if (hipFuncSetAttribute((const void*)k_ts_evolve<NTA, 2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds) != hipSuccess)
                 (void)hipGetLastError();  // the launch below reports a size the device refuses
               hipLaunchKernelGGL((k_ts_evolve<NTA, 2>), dim3(nblk_ev), dim3(NTA), lds, c->stream, dd, ea);
             }
           }
           MXR_HIP(hipGetLastError());
           return launch_final_sum(c, c->partials, nblk_ev, nv, c->sums);
         }));
    MXR_HIP(hipMemcpyAsync(sums, c->sums, sizeof(double) * nv, hipMemcpyDeviceToHost, c->stream));
  }
  // x_e_box = x_e_box_prev (ComputeTs.c:1459), the whole padded array
  MXR_HIP(hipMemcpyAsync(xe_out, xe, sizeof(float) * npad, hipMemcpyDeviceToDevice, c->stream));
  MXR_HIP(hipEventRecord(c->ev1, c->stream));
  if (g->mem == MXR_MEM_HOST) {
    MXR_HIP(hipMemcpyAsync(g->x_e_box_prev, xe, sizeof(float) * npad, hipMemcpyDeviceToHost, c->stream));
    MXR_HIP(hipMemcpyAsync(g->x_e_box, xe_out, sizeof(float) * npad, hipMemcpyDeviceToHost, c->stream));
    MXR_HIP(hipMemcpyAsync(g->Tk_box, Tk, sizeof(float) * nreal, hipMemcpyDeviceToHost, c->stream));
    MXR_HIP(hipMemcpyAsync(g->TS_box, TS, sizeof(float) * nreal, hipMemcpyDeviceToHost, c->stream));
  }
  MXR_HIP(hipStreamSynchronize(c->stream));
  float ms = 0.f;
  MXR_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  RC(prof_resolve(c));
  if (pl.homogeneous) {
    out->volume_ave_TS = sums[0] / total_n_cells;
    out->volume_ave_TK = sums[1] / total_n_cells;
    out->volume_ave_xe = sums[2] / total_n_cells;
  } else {
    out->volume_ave_J_alpha = sums[0] / total_n_cells;
    out->volume_ave_xalpha = sums[1] / total_n_cells;
    out->volume_ave_Xheat = sums[2] / total_n_cells;
    out->volume_ave_Xion = sums[3] / total_n_cells;
    out->volume_ave_TS = sums[4] / total_n_cells;
    out->volume_ave_TK = sums[5] / total_n_cells;
    out->volume_ave_xe = sums[6] / total_n_cells;
    out->XrayEmissivity_HMXB = sums[7] / total_n_cells;
    out->XrayEmissivity_soft = out->XrayEmissivity_hard = 0.0;
    out->volume_ave_Xheat_AGN_soft = out->volume_ave_Xheat_AGN_hard = 0.0;
    for (int b = 0; b < NB; b++) {
      const bool soft = hd->agn_kind[b] == 0;
      (soft ? out->volume_ave_Xheat_AGN_soft : out->volume_ave_Xheat_AGN_hard) = sums[8 + b] / total_n_cells;
      (soft ? out->XrayEmissivity_soft : out->XrayEmissivity_hard) = sums[8 + NB + b] / total_n_cells;
    }
  }
  out->device_ms = ms;
  return MXR_OK;
}

int run_compute_ts(Ctx* c, const mxr_ts_params* p, const mxr_ts_grids* g, mxr_ts_globals* out,
                   mxr_ts_tables* tables)
{
  switch (c->N) {
#define X(n)                                                                                                         \
  case n:                                                                                                            \
    return ts_run<n>(c, p, g, out, tables);
    MXR_DIMS(X)
#undef X
    default: set_error("unsupported grid dimension %d", c->N); return MXR_ERR_INVALID_ARG;
  }
}

}  // namespace mxr
