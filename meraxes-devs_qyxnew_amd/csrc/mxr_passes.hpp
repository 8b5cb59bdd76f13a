// mxr_passes.hpp -- the FFT passes shared by the find_HII_bubbles and spin-temperature
// kernels: grid pointer packs, the reference's k-space window arithmetic, the column pass
// (x and y), the forward row pass (r2c), one grid's inverse row pass (c2r) into registers,
// FP64 block sums and the per-dimension launch configuration.
#pragma once

#include <cmath>

#include "mxr_fft.hpp"
#include "mxr_internal.hpp"

namespace mxr {

// up to 7 grids: deltax, stars, weighted_sfr, [x_e], [effective_bhm, effective_bhar], [N_rec]
constexpr int MAX_GRIDS = 7;
struct GridPtrs {
  float2* p[MAX_GRIDS];
};
struct CGridPtrs {
  const float2* p[MAX_GRIDS];
};
struct CRealPtrs {
  const float* p[MAX_GRIDS];
  unsigned bf16;  // bit g: p[g] holds bfloat16 bit patterns (uint16), widened on load
  unsigned real;  // bit g: p[g] is a real-layout grid (rows of N floats, no padding)
  // deldel loads (power spectrum, one grid): when dd_sum is set, p[0] is a real-layout grid
  // v and each value loads as dd = (float)((v / ave - 1) volume / total), then
  // (float)(dd ave), with ave = *dd_sum / total (ComputePowerSpectrum.c:97-104)
  const double* dd_sum;
  double dd_volume, dd_total, dd_rtotal;  // dd_rtotal = 1 / dd_total rounded to nearest
  int dd_total_pow2;                      // dd_total is a power of two: dd_rtotal is exact
};

// x / d rounded to nearest from r = RN(1/d): q = RN(x r) is within an ulp of x / d, the residual
// x - q d is exact through the FMA, and RN(q + residual r) is RN(x / d) (Markstein's theorem)
__device__ __forceinline__ double div_rcp(double x, double d, double r)
{
  const double q = x * r;
  return __fma_rn(__fma_rn(-q, d, x), r, q);
}

// pow2: total is a power of two, so multiplying by its reciprocal is the exact division
__device__ __forceinline__ float deldel(float v, double ave, double rave, double volume, double total,
                                        double rtotal, bool pow2)
{
#pragma clang fp contract(off)
  const double t = (div_rcp((double)v, ave, rave) - 1) * volume;
  const float x = (float)(pow2 ? t * rtotal : div_rcp(t, total, rtotal));
  return (float)((double)x * ave);
}

__device__ __forceinline__ float bf16_to_f32(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// The columns of a pass over a half spectrum of dimension L are (outer, k): outer = j (x pass) or
// i (y pass), k < L/2 + 1 the mode along z.  Column (outer, k) starts at outer * ostride + k and
// its L elements are es apart.  Columns are numbered with the L/2 leading modes of every outer
// index first (outer L/2 + k), then the Nyquist modes k = L/2 of all outer indices: every tile
// of TC consecutive columns but the last is full, and a tile of leading modes covers TC adjacent,
// aligned values of one spectrum row (the row padding is never touched).
struct ColGeom {
  long ostride, es;
};

// Window of one radius.  apply = 0 on the last (unfiltered) step.
struct FiltArgs {
  int N;
  float dk;   // (float)(2 pi / (float)BoxSize)
  float R;    // (float)R
  int ftype;
  int apply;
  int prune;  // skip the modes the sharp-k window zeroes
  int live;   // > 0: the grid covers only tiles that can hold kept columns, `live` = kept k_z count
  // types 0 and 2: when set, the window's values by (|n_x|, |n_y|, n_z), each in 0..N/2 (index
  // (a (N/2+1) + b) (N/2+1) + c), evaluated by window_mul beforehand; |k| depends on the three
  // magnitudes only, so a table entry is the value window_mul gives at any of its modes
  const float* wtab;
};

// COL_SCALE divides by `div` in double (find_HII_bubbles.c:235-253; as an exact reciprocal
// product when div is a power of two), COL_FSCALE by (float)div
// in float (ComputeTs.c:485-487)
enum { COL_PLAIN = 0, COL_SCALE = 1, COL_FILTER = 2, COL_YPRUNE = 3, COL_FSCALE = 4 };

// grid pointer selection without dynamic indexing of a by-value kernel argument
template <class P>
__device__ __forceinline__ auto pick(const P& ps, int g)
{
  return g == 0   ? ps.p[0]
         : g == 1 ? ps.p[1]
         : g == 2 ? ps.p[2]
         : g == 3 ? ps.p[3]
         : g == 4 ? ps.p[4]
         : g == 5 ? ps.p[5]
                  : ps.p[6];
}
// ---- the reference's window arithmetic (float, unfused) --------------------------------
// correctly rounded float square root of x = 0 or a normal x > 0 (sqrtf's result): the hardware
// root, within an ulp, then the residuals of its two neighbours decide, as in the usual correctly
// rounded expansion (HIP's __fsqrt_rn is the hardware root itself)
__device__ __forceinline__ float sqrt_rn(float x)
{
  const float s = __builtin_amdgcn_sqrtf(x);
  const float dn = __int_as_float(__float_as_int(s) - 1), up = __int_as_float(__float_as_int(s) + 1);
  float r = __builtin_fmaf(-dn, s, x) <= 0.f ? dn : s;
  r = __builtin_fmaf(-up, s, x) > 0.f ? up : r;
  return r;
}

__device__ __forceinline__ float kmag_ref(int nx, int ny, int nz, int N, float dk)
{
  const int mid = N / 2;
  const float kx = mul_rn((float)(nx > mid ? nx - N : nx), dk);
  const float ky = mul_rn((float)(ny > mid ? ny - N : ny), dk);
  const float kz = mul_rn((float)nz, dk);
  return sqrt_rn(add_rn(add_rn(mul_rn(kx, kx), mul_rn(ky, ky)), mul_rn(kz, kz)));
}

// sharp-k (type 1): kR *= 0.413566994 (double product, float result); zero if kR > 1
__device__ __forceinline__ bool sharpk_keep(int nx, int ny, int nz, const FiltArgs& fa)
{
  const float kR = mul_rn(kmag_ref(nx, ny, nz, fa.N, fa.dk), fa.R);
  const float s = (float)((double)kR * 0.413566994);
  return !(s > 1.0f);
}

// Along one axis of a column the sharp-k test depends on |n| only (the index n and N - n give
// the same |k| bit for bit) and every step of it is monotone in |n|, so it keeps exactly the
// |n| <= extent: the largest such |n| in 0..N/2 by bisection, or -1 when |n| = 0 is dropped.
template <class Keep>
__device__ __forceinline__ int kept_extent(Keep keep, int N)
{
  if (!keep(0))
    return -1;
  int lo = 0, hi = N / 2;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (keep(mid))
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// multiplicative window of types 0 (real-space top-hat) and 2 (Gaussian); one out-of-line copy
// serves every kernel, so the fused and plain paths evaluate identical code
static __device__ __noinline__ float window_mul(int nx, int ny, int nz, const FiltArgs& fa)
{
  float kR = mul_rn(kmag_ref(nx, ny, nz, fa.N, fa.dk), fa.R);
  if (fa.ftype == 0) {
    if ((double)kR > 1e-4) {
      const float a = __fdiv_rn(sinf(kR), powf(kR, 3.0f));
      const float b = __fdiv_rn(cosf(kR), powf(kR, 2.0f));
      return (float)(3.0 * (double)sub_rn(a, b));
    }
    return 1.0f;
  }
  kR = (float)((double)kR * 0.643);
  return powf(2.71828182845904523536f, (float)((double)mul_rn(-kR, kR) / 2.0));
}

// ---- column pass (x: columns = (j,k) flattened, elements i; y: columns = (i,k), elements j)
template <int L, int NT, int TC, int S, int MODE, int TPB>
__global__ void __launch_bounds__(NT) k_col(CGridPtrs in, GridPtrs out, ColGeom geo, const float2* __restrict__ twg,
                                            FiltArgs fa, double div, long ntile)
{
  constexpr int T = NT / TC;
  using SL = Slots<L, T>;
  __shared__ float2 lds[L * TC];
  __shared__ float2 tw[L];
  const int g = blockIdx.y;
  const int c = threadIdx.x % TC;
  const int t = threadIdx.x / TC;
  for (int m = threadIdx.x; m < L; m += NT)
    tw[m] = twg[m];
  const float2* __restrict__ src = pick(in, g);

  // TPB adjacent tiles per block: every tile's loads are issued first, so the later tiles'
  // memory latency hides behind the earlier tiles' transforms (the block barriers inside the
  // transform wait for LDS only)
  long base[TPB];
  bool colnz[TPB];
  int fj[TPB], fk[TPB];
  float2 v[TPB][SL::NREG];
#pragma unroll
  for (int i = 0; i < TPB; i++) {
    const long tile = (long)blockIdx.x * TPB + i;
    constexpr int HM = L / 2;
    const long m = tile * TC + c;
    bool valid = tile < ntile && m < (long)L * (HM + 1);
    int outer = 0, kk = 0;
    if (valid) {
      if (m < (long)L * HM) {
        outer = (int)(m / HM);
        kk = (int)(m % HM);
      } else {
        outer = (int)(m - (long)L * HM);
        kk = HM;
      }
    }
    if constexpr (MODE == COL_FILTER || MODE == COL_YPRUNE) {
      if (fa.live > 0) {
        // live-tile grid: tiles enumerate (outer index, k_z tile) pairs whose k_z tile starts
        // below the kept count; for the x pass the outer index is j with |n_y| < live
        const int nkt = (fa.live + TC - 1) / TC;
        const int jj = (int)(tile / nkt), kt = (int)(tile % nkt);
        outer = jj;
        if constexpr (MODE == COL_FILTER)
          outer = jj < fa.live ? jj : fa.N - (2 * fa.live - 1 - jj);
        kk = kt * TC + c;
        valid = tile < ntile && kk <= HM;
      }
    }
    base[i] = valid ? (long)outer * geo.ostride + kk : 0;
    colnz[i] = valid;
    fj[i] = 0;
    fk[i] = 0;
    if constexpr (MODE == COL_FILTER) {
      fj[i] = outer;
      fk[i] = kk;
      if (valid && fa.apply && fa.prune && fa.ftype == 1)
        colnz[i] = sharpk_keep(0, fj[i], fk[i], fa);
    } else if constexpr (MODE == COL_YPRUNE) {
      fk[i] = kk;
      if (valid && fa.apply && fa.prune)
        colnz[i] = sharpk_keep(0, 0, fk[i], fa);
    }
    // the sharp-k window's kept extent of |n| along the column, once per column; elements past
    // it load the column's first element (a cache hit) and are zeroed, so every load issues
    // together
    int ext = L;
    if constexpr (MODE == COL_FILTER) {
      if (colnz[i] && fa.apply && fa.ftype == 1) {
        const int j0 = fj[i], k0 = fk[i];
        ext = kept_extent([&](int m) { return sharpk_keep(m, j0, k0, fa); }, L);
      }
    } else if constexpr (MODE == COL_YPRUNE) {
      if (colnz[i] && fa.apply && fa.prune) {
        const int k0 = fk[i];
        ext = kept_extent([&](int m) { return sharpk_keep(0, m, k0, fa); }, L);
      }
    }
#pragma unroll
    for (int q = 0; q < SL::Q0; q++) {
      const int j = t + T * q;
#pragma unroll
      for (int r = 0; r < SL::R0; r++) {
        const int e = j + r * SL::B0;
        const bool in_col = (SL::B0 % T == 0 || j < SL::B0) && colnz[i];
        const bool keep = in_col && (e < L - e ? e : L - e) <= ext;
        float2 x = src[keep ? base[i] + (long)e * geo.es : base[i]];
        if constexpr (MODE == COL_FILTER) {
          if (fa.apply && fa.ftype != 1 && !fa.wtab && in_col) {
            const float w = window_mul(e, fj[i], fk[i], fa);
            // rounded on its own (no fusion into the first butterfly), as filter() does
            x = make_float2(mul_rn(x.x, w), mul_rn(x.y, w));
          }
        }
        v[i][q * SL::R0 + r] = keep ? x : make_float2(0.f, 0.f);
      }
    }
    if constexpr (MODE == COL_FILTER) {
      // tabulated window (types 0 and 2): every entry's load issued, then the products
      if (fa.apply && fa.ftype != 1 && fa.wtab) {
        const int Hn = fa.N / 2 + 1;
        const long wrow = (long)min(fj[i], fa.N - fj[i]) * Hn + fk[i];
        float w[SL::NREG];
#pragma unroll
        for (int q = 0; q < SL::Q0; q++) {
          const int j = t + T * q;
#pragma unroll
          for (int r = 0; r < SL::R0; r++) {
            const int e = j + r * SL::B0;
            const bool in_col = (SL::B0 % T == 0 || j < SL::B0) && colnz[i];
            w[q * SL::R0 + r] = fa.wtab[in_col ? (long)min(e, fa.N - e) * Hn * Hn + wrow : 0];
          }
        }
#pragma unroll
        for (int q = 0; q < SL::Q0; q++) {
          const int j = t + T * q;
#pragma unroll
          for (int r = 0; r < SL::R0; r++) {
            const bool in_col = (SL::B0 % T == 0 || j < SL::B0) && colnz[i];
            float2& x = v[i][q * SL::R0 + r];
            const float wv = w[q * SL::R0 + r];
            if (in_col)
              x = make_float2(mul_rn(x.x, wv), mul_rn(x.y, wv));
          }
        }
      }
    }
  }

  float2* __restrict__ dst = pick(out, g);
  // COL_SCALE with a power-of-two divisor (N^3 for N = 2^a): the product with its reciprocal,
  // which is exact, equals the double division bit for bit (no product of a float's value and
  // 2^-k leaves the normal double range), at the cost of one multiply
  double rdiv = 0.0;
  if constexpr (MODE == COL_SCALE) {
    if ((__double_as_longlong(div) & 0x000FFFFFFFFFFFFFLL) == 0)
      rdiv = 1.0 / div;
  }
#pragma unroll
  for (int i = 0; i < TPB; i++) {
    // block-uniform: a tile whose columns are all zero for this radius is skipped; the barrier
    // also orders the LDS between consecutive tiles' transforms and publishes the twiddles
    if (!__syncthreads_or(colnz[i]))
      continue;
    tile_fft<L, T, S, L, LdsCol<TC>>(v[i], lds, tw, c, t);
    if (!colnz[i])
      continue;
#pragma unroll
    for (int q = 0; q < SL::QL; q++) {
      const int j = t + T * q;
      if (SL::BL % T == 0 || j < SL::BL) {
#pragma unroll
        for (int r = 0; r < SL::RL; r++) {
          const int e = j + r * SL::BL;
          float2 y = v[i][q * SL::RL + r];
          if constexpr (MODE == COL_SCALE) {
            if (rdiv != 0.0)
              y = make_float2((float)((double)y.x * rdiv), (float)((double)y.y * rdiv));
            else
              y = make_float2((float)((double)y.x / div), (float)((double)y.y / div));
          }
          if constexpr (MODE == COL_FSCALE)
            y = make_float2(__fdiv_rn(y.x, (float)div), __fdiv_rn(y.y, (float)div));
          dst[base[i] + (long)e * geo.es] = y;
        }
      }
    }
  }
}

// r2c post-twiddle of bin k from Z[k] and conj Z[M-k] (zm), t = exp(-2 pi i k / N):
// X[k] = e + t o, e = (Z[k] + zm) / 2, o = (Z[k] - zm) / (2i); two FMAs per part, contraction off
__device__ __forceinline__ float2 post_twiddle(float2 zk, float2 zm, float2 t)
{
#pragma clang fp contract(off)
  const float2 e = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y + zm.y));
  const float2 d = make_float2(0.5f * (zk.x - zm.x), 0.5f * (zk.y - zm.y));
  const float2 o = make_float2(d.y, -d.x);  // d / i
  return make_float2(__builtin_fmaf(t.x, o.x, __builtin_fmaf(-t.y, o.y, e.x)),
                     __builtin_fmaf(t.x, o.y, __builtin_fmaf(t.y, o.x, e.y)));
}

// ---- z-pass forward: padded real rows -> half-spectrum rows ------------------------------
// x[2n] + i x[2n+1] packed into an N/2-point complex FFT Z; then
// X[k] = (Z[k] + conj Z[M-k]) / 2 + exp(-2 pi i k / N) (Z[k] - conj Z[M-k]) / (2i), k = 0..M.
template <int N, int NT, int TR>
__global__ void __launch_bounds__(NT) k_row_r2c(CRealPtrs in, GridPtrs out, long nrows, const float2* __restrict__ twg)
{
  constexpr int M = N / 2, H = N / 2 + 1, T = NT / TR, P = M + M / 8 + 1;
  using SL = Slots<M, T>;
  using LI = LdsRowPad<P>;
  __shared__ float2 lds[TR * P];
  __shared__ float2 tw[N];
  const int g = blockIdx.y;
  const int c = threadIdx.x / T, t = threadIdx.x % T;
  const long row = (long)blockIdx.x * TR + c;
  const bool valid = row < nrows;
  for (int m = threadIdx.x; m < N; m += NT)
    tw[m] = twg[m];
  // one uniform choice of the load form, then every load issued (from a clamped in-row index)
  // before any is used; slots past the row's data are zeroed afterwards
  const long rb = valid ? row : 0;
  auto nc = [&](int q, int r) { return min(t + T * q + r * SL::B0, M - 1); };
  float2 v[SL::NREG];
  if (in.dd_sum) {
    const double ave = *in.dd_sum / in.dd_total, rave = 1.0 / ave;
    const float* __restrict__ src = pick(in, g) + rb * N;
#pragma unroll
    for (int q = 0; q < SL::Q0; q++)
#pragma unroll
      for (int r = 0; r < SL::R0; r++)
        v[q * SL::R0 + r] = *reinterpret_cast<const float2*>(src + 2 * nc(q, r));
#pragma unroll
    for (int i = 0; i < SL::Q0 * SL::R0; i++)
      v[i] = make_float2(deldel(v[i].x, ave, rave, in.dd_volume, in.dd_total, in.dd_rtotal, in.dd_total_pow2),
                         deldel(v[i].y, ave, rave, in.dd_volume, in.dd_total, in.dd_rtotal, in.dd_total_pow2));
  } else if ((in.bf16 >> g) & 1u) {
    // a row of 2H bfloat16 values is H 32-bit words
    const unsigned* __restrict__ src = reinterpret_cast<const unsigned*>(pick(in, g)) + rb * H;
    unsigned u[SL::Q0 * SL::R0];
#pragma unroll
    for (int q = 0; q < SL::Q0; q++)
#pragma unroll
      for (int r = 0; r < SL::R0; r++)
        u[q * SL::R0 + r] = src[nc(q, r)];
#pragma unroll
    for (int i = 0; i < SL::Q0 * SL::R0; i++)
      v[i] = make_float2(bf16_to_f32((unsigned short)(u[i] & 0xffffu)), bf16_to_f32((unsigned short)(u[i] >> 16)));
  } else {
    const float* __restrict__ src = pick(in, g) + rb * (((in.real >> g) & 1u) ? N : 2 * H);
#pragma unroll
    for (int q = 0; q < SL::Q0; q++)
#pragma unroll
      for (int r = 0; r < SL::R0; r++)
        v[q * SL::R0 + r] = *reinterpret_cast<const float2*>(src + 2 * nc(q, r));
  }
#pragma unroll
  for (int q = 0; q < SL::Q0; q++) {
    const bool ok = valid && (SL::B0 % T == 0 || t + T * q < SL::B0);
#pragma unroll
    for (int r = 0; r < SL::R0; r++)
      if (!ok)
        v[q * SL::R0 + r] = make_float2(0.f, 0.f);
  }
  __syncthreads();  // twiddles visible (single-stage transforms use them below)
  tile_fft<M, T, -1, N, LI, SyncWave>(v, lds, tw, c, t);
#pragma unroll
  for (int q = 0; q < SL::QL; q++) {
    const int j = t + T * q;
    if (SL::BL % T == 0 || j < SL::BL)
#pragma unroll
      for (int r = 0; r < SL::RL; r++)
        lds[LI::at(c, j + r * SL::BL)] = v[q * SL::RL + r];
  }
  __syncthreads();
  if (!valid)
    return;
  float2* __restrict__ dst = pick(out, g) + row * spec_pitch(N);
  for (int k = t; k <= M; k += T) {
    const float2 zk = lds[LI::at(c, k % M)];
    const float2 zm0 = lds[LI::at(c, (M - k) % M)];
    const float2 zm = make_float2(zm0.x, -zm0.y);
    dst[k] = post_twiddle(zk, zm, tw[k]);
  }
}

// wave-wide then block-wide FP64 sum of NV values; thread 0 writes them to out
template <int NT, int NV>
__device__ __forceinline__ void block_sum(double* acc, double* out)
{
  __shared__ double red[NT / 64][NV];
#pragma unroll
  for (int q = 0; q < NV; q++) {
    double x = acc[q];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
      x += __shfl_down(x, o, 64);
    acc[q] = x;
  }
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < NV; q++)
      red[w][q] = acc[q];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; q++) {
      double s = 0.0;
      for (int ww = 0; ww < NT / 64; ww++)
        s += red[ww][q];
      out[q] = s;
    }
  }
}

// c2r pre-twiddle: z = s + i w d with w = conj(t) = exp(+2 pi i n / N), i.e.
// z = (s.x - (t.x d.y - t.y d.x), s.y + (t.x d.x + t.y d.y)); two FMAs per part, contraction off
__device__ __forceinline__ float2 pre_twiddle(float2 s, float2 d, float2 t)
{
#pragma clang fp contract(off)
  return make_float2(__builtin_fmaf(t.y, d.x, __builtin_fmaf(-t.x, d.y, s.x)),
                     __builtin_fmaf(t.y, d.y, __builtin_fmaf(t.x, d.x, s.y)));
}

// One grid's rows through the inverse z-pass: load the half-spectrum rows into LDS (only the
// kc leading k the window keeps -- a prefix, since |k| grows with k_z; the real parts of the
// k_z = 0 and N/2 bins, as a c2r of the Hermitian extension uses), pre-twiddle, M-point
// inverse FFT.  out receives the
// last-stage slots: slot (q, r) = y[2n] + i y[2n+1], n = t + T q + r BL.  A row's T threads
// are consecutive lanes of one wavefront, so the stages synchronise per wave, not per block.
template <int N, int T, int TR>
__device__ __forceinline__ void row_c2r_grid(const float2* __restrict__ base, bool valid, long row, float2* lds,
                                             const float2* tw, int kc, int c, int t, float2* out)
{
  constexpr int M = N / 2, H = N / 2 + 1, P = M + M / 8 + 1;
  static_assert(64 % T == 0, "a row must not straddle two wavefronts");
  using SL = Slots<M, T>;
  using LI = LdsRowPad<P>;
  const float2* __restrict__ src = base + (valid ? row : 0) * spec_pitch(N);
  // the loads of a batch first, then their LDS stores: the loads are independent, so a batch is
  // in flight together instead of one round trip per element (two batches bound the registers)
  constexpr int KI = (H + T - 1) / T;
  constexpr int KB = (KI + 1) / 2;
#pragma unroll
  for (int b0 = 0; b0 < KI; b0 += KB) {
    float2 xin[KB];
#pragma unroll
    for (int i = 0; i < KB; i++) {
      const int k = t + (b0 + i) * T;
      float2 x = make_float2(0.f, 0.f);
      if (b0 + i < KI && valid && k < kc)
        x = src[k];
      xin[i] = x;
    }
#pragma unroll
    for (int i = 0; i < KB; i++) {
      const int k = t + (b0 + i) * T;
      float2 x = xin[i];
      if (k == 0 || k == M)
        x.y = 0.f;  // c2r semantics: the imaginary parts of the k_z = 0 and N/2 bins do not contribute
      if (b0 + i < KI && k < H)
        lds[LI::at(c, k)] = x;
    }
  }
  SyncWave::sync();
  float2 v[SL::NREG];
#pragma unroll
  for (int q = 0; q < SL::Q0; q++) {
    const int j = t + T * q;
#pragma unroll
    for (int r = 0; r < SL::R0; r++) {
      const int n = j + r * SL::B0;
      float2 z = make_float2(0.f, 0.f);
      if (SL::B0 % T == 0 || j < SL::B0) {
        const float2 xn = lds[LI::at(c, n)];
        const float2 xm0 = lds[LI::at(c, M - n)];
        const float2 xm = make_float2(xm0.x, -xm0.y);
        const float2 s = cadd(xn, xm);
        const float2 d = csub(xn, xm);
        z = pre_twiddle(s, d, tw[n]);
      }
      v[q * SL::R0 + r] = z;
    }
  }
  SyncWave::sync();
  tile_fft<M, T, +1, N, LI, SyncWave>(v, lds, tw, c, t);
#pragma unroll
  for (int s = 0; s < SL::QL * SL::RL; s++)
    out[s] = v[s];
  SyncWave::sync();
}

// ---- launch configuration per grid dimension ------------------------------------------
constexpr int pow2_floor(int x)
{
  int p = 1;
  while (p * 2 <= x)
    p *= 2;
  return p;
}

template <int N>
struct Cfg {
  static constexpr int M = N / 2;
  static constexpr int H = N / 2 + 1;
  static constexpr int TC = 16;
  // threads per column: 16 values each at N = 512 and 1024 (a 1024-point column over 16
  // threads would hold 64 per thread)
  static constexpr int COL_T = N >= 1024 ? 64 : N == 512 ? 32 : (N / 8 >= 16) ? 16 : (N / 8 < 1 ? 1 : N / 8);
  static constexpr int COL_NT = TC * COL_T;
  static constexpr int COL_TPB = 1;  // column tiles per block (2 overlaps loads with transforms but needs ~280 VGPRs)
  // threads per row (one wavefront at most): 32 at N = 1024 keeps 16 values per thread and
  // the row tile's LDS at 45 KB
  static constexpr int ROW_T = N >= 1024 ? 32 : pow2_floor((M / 8 >= 16) ? 16 : (M / 8 < 1 ? 1 : M / 8));
  static constexpr int ROW_NT = 256;
  static constexpr int TR = ROW_NT / ROW_T;
};

static inline unsigned ew_blocks(long n)
{
  long b = (n + 255) / 256;
  if (b > 8192)
    b = 8192;
  return (unsigned)(b < 1 ? 1 : b);
}

// Modes the sharp-k window keeps at radius R (host restatement of sharpk_keep, used only to
// count algorithmic bytes): k1 = kept k_z on the k_x = k_y = 0 line, c2 = kept (k_y, k_z)
// columns at k_x = 0, k3 = kept modes overall.
struct KeptModes {
  long k1, c2, k3;
  static KeptModes all(int N)
  {
    const long H = N / 2 + 1;
    return KeptModes{H, (long)N * H, (long)N * N * H};
  }
};

static inline bool keep_host(int nx, int ny, int nz, int N, float dk, float R)
{
  const int mid = N / 2;
  volatile float kx = (float)(nx > mid ? nx - N : nx) * dk;
  volatile float ky = (float)(ny > mid ? ny - N : ny) * dk;
  volatile float kz = (float)nz * dk;
  volatile float x2 = kx * kx, y2 = ky * ky, z2 = kz * kz;
  volatile float s = x2 + y2;
  s = s + z2;
  volatile float kR = std::sqrt((float)s) * R;
  return !((float)((double)kR * 0.413566994) > 1.0f);
}

static inline KeptModes kept_modes(int N, float dk, float R)
{
  const int H = N / 2 + 1;
  KeptModes km{0, 0, 0};
  for (int k = 0; k < H; k++) {
    if (!keep_host(0, 0, k, N, dk, R))
      break;
    km.k1++;
    for (int j = 0; j < N; j++) {
      if (!keep_host(0, j, k, N, dk, R))
        continue;
      km.c2++;
      // kept |n_x| form a prefix of 0..N/2 (|k| grows with |n_x|)
      int lo = 0, hi = N / 2;
      while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (keep_host(mid, j, k, N, dk, R))
          lo = mid;
        else
          hi = mid - 1;
      }
      km.k3 += 1 + 2L * lo - ((N % 2 == 0 && lo == N / 2) ? 1 : 0);
    }
  }
  return km;
}

// kept k_z on the k_x = k_y = 0 line at this radius (a prefix: |k| grows with k_z); H when the
// window is not applied or does not zero modes
static inline int kept_kz(const FiltArgs& fa)
{
  const int H = fa.N / 2 + 1;
  if (!(fa.apply && fa.prune && fa.ftype == 1))
    return H;
  int kc = 0;
  while (kc < H && keep_host(0, 0, kc, fa.N, fa.dk, fa.R))
    kc++;
  return kc;
}


// launchers of the shared passes for one grid dimension
template <int N>
struct Passes {
  using C = Cfg<N>;

  static constexpr int HP = spec_pitch(N);
  static ColGeom xgeo() { return ColGeom{(long)HP, (long)N * HP}; }
  static ColGeom ygeo() { return ColGeom{(long)N * HP, (long)HP}; }

  // one column pass with TC columns per workgroup (the per-column arithmetic does not depend on TC)
  template <int S, int MODE, int TC>
  static int col_tc(Ctx* c, CGridPtrs in, GridPtrs out, int ngrid, ColGeom geo, FiltArgs fa, double div)
  {
    constexpr int NT = TC * C::COL_T;
    constexpr int TPB = C::COL_TPB;
    long ntile = ((long)N * C::H + TC - 1) / TC;
    if ((MODE == COL_FILTER || MODE == COL_YPRUNE) && fa.live > 0) {
      const long nkt = (fa.live + TC - 1) / TC;
      ntile = (MODE == COL_FILTER ? 2L * fa.live - 1 : (long)N) * nkt;
    }
    dim3 grid((unsigned)((ntile + TPB - 1) / TPB), (unsigned)ngrid);
    hipLaunchKernelGGL((k_col<N, NT, TC, S, MODE, TPB>), grid, dim3(NT), 0, c->stream, in, out, geo, c->tw, fa, div,
                       ntile);
    MXR_HIP(hipGetLastError());
    return MXR_OK;
  }

  // the per-radius inverse passes take their tile width from `tc` (4, 8, 16 or 32); the others use 16
  template <int S, int MODE>
  static int col(Ctx* c, CGridPtrs in, GridPtrs out, int ngrid, ColGeom geo, FiltArgs fa, double div, int tc = C::TC)
  {
    if constexpr (S > 0 && (MODE == COL_FILTER || MODE == COL_YPRUNE)) {
      if (tc == 4)
        return col_tc<S, MODE, 4>(c, in, out, ngrid, geo, fa, div);
      if (tc == 8)
        return col_tc<S, MODE, 8>(c, in, out, ngrid, geo, fa, div);
      if constexpr (32 * C::COL_T <= 1024)  // a workgroup holds at most 1024 threads
        if (tc == 32)
          return col_tc<S, MODE, 32>(c, in, out, ngrid, geo, fa, div);
    }
    return col_tc<S, MODE, C::TC>(c, in, out, ngrid, geo, fa, div);
  }

  static int r2c(Ctx* c, CRealPtrs in, GridPtrs out, int ngrid)
  {
    const long nrows = (long)N * N;
    dim3 grid((unsigned)((nrows + C::TR - 1) / C::TR), (unsigned)ngrid);
    hipLaunchKernelGGL((k_row_r2c<N, C::ROW_NT, C::TR>), grid, dim3(C::ROW_NT), 0, c->stream, in, out, nrows, c->tw);
    MXR_HIP(hipGetLastError());
    return MXR_OK;
  }
};

}  // namespace mxr
