// mxr_fft.hpp -- LDS-staged Stockham FFT building blocks for gfx950.
//
// A "tile" is TC independent transforms of length L held by one workgroup of NT threads.
// Stage s (radix R, Ns = product of the earlier radices) maps butterfly j (0 <= j < L/R) to
//   inputs  x[j + r L/R] * w^(k r),  w = exp(S 2 pi i / (Ns R)),  k = j mod Ns
//   outputs y[(j / Ns) Ns R + k + q Ns]   (Stockham autosort: natural order after the last stage)
// The first stage takes its inputs straight from registers filled by the caller's loads (Ns = 1,
// no twiddles) and the last stage hands its outputs (natural positions j + q L/R) straight to
// the caller, so a length-R0*R1*R2 transform makes two LDS round trips.
// Twiddles come from a table tw[m] = exp(-2 pi i m / LT) (evaluated in FP64 on the host,
// stored FP32) kept in LDS; the inverse (S = +1) uses its conjugate.
#pragma once

#include <hip/hip_runtime.h>

namespace mxr {

// The transform's rounding is fixed by the source: every building block below turns FP
// contraction off and writes its fused multiply-adds explicitly, so the compiler cannot fuse a
// product with a neighbouring sum differently in different kernels (with contraction left to
// the compiler, the same butterfly compiled into two kernels with different epilogues took
// different FMA choices and rounded differently).
// Single-rounding float multiply, add and subtract that FP contraction never merges with a
// neighbouring operation.  HIP's __fmul_rn and __fadd_rn are the plain operators, which the
// compiler fuses into an FMA when it can (x*y + z rounded once), unlike the unfused float
// arithmetic they restate; an operation written under contract(off) keeps its own rounding
// wherever it is inlined.
__device__ __forceinline__ float mul_rn(float a, float b)
{
#pragma clang fp contract(off)
  return a * b;
}
__device__ __forceinline__ float add_rn(float a, float b)
{
#pragma clang fp contract(off)
  return a + b;
}
__device__ __forceinline__ float sub_rn(float a, float b)
{
#pragma clang fp contract(off)
  return a - b;
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b)
{
#pragma clang fp contract(off)
  return make_float2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ float2 csub(float2 a, float2 b)
{
#pragma clang fp contract(off)
  return make_float2(a.x - b.x, a.y - b.y);
}
// a b = (a.x b.x - a.y b.y, a.x b.y + a.y b.x), each part one product and one FMA
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
#pragma clang fp contract(off)
  return make_float2(__builtin_fmaf(a.x, b.x, -(a.y * b.y)), __builtin_fmaf(a.x, b.y, a.y * b.x));
}
// a * (S == -1 ? w : conj(w))
template <int S>
__device__ __forceinline__ float2 ctw(float2 a, float2 w)
{
#pragma clang fp contract(off)
  if constexpr (S < 0)
    return make_float2(__builtin_fmaf(a.x, w.x, -(a.y * w.y)), __builtin_fmaf(a.x, w.y, a.y * w.x));
  else
    return make_float2(__builtin_fmaf(a.x, w.x, a.y * w.y), __builtin_fmaf(a.y, w.x, -(a.x * w.y)));
}
// a * (S i)
template <int S>
__device__ __forceinline__ float2 cmul_si(float2 a)
{
  if constexpr (S < 0)
    return make_float2(a.y, -a.x);
  else
    return make_float2(-a.y, a.x);
}

// ---- radix-R DFTs in registers: V[q] = sum_r v[r] exp(S 2 pi i r q / R) -------------------
template <int S>
__device__ __forceinline__ void dft2(float2* v)
{
#pragma clang fp contract(off)
  float2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <int S>
__device__ __forceinline__ void dft4(float2* v)
{
#pragma clang fp contract(off)
  float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
  float2 t2 = cadd(v[1], v[3]), t3 = cmul_si<S>(csub(v[1], v[3]));
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3);
  v[3] = csub(t1, t3);
}

template <int S>
__device__ __forceinline__ void dft8(float2* v)
{
  // even/odd split: V[2p] = DFT4(v_r + v_{r+4}), V[2p+1] = DFT4((v_r - v_{r+4}) w8^r)
#pragma clang fp contract(off)
  constexpr float h = 0.70710678118654752440f;
  float2 b[4], c[4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    b[r] = cadd(v[r], v[r + 4]);
    c[r] = csub(v[r], v[r + 4]);
  }
  // w8 = (1 + S i)/sqrt2, w8^2 = S i, w8^3 = (-1 + S i)/sqrt2
  c[1] = make_float2(h * (c[1].x - (float)S * c[1].y), h * (c[1].y + (float)S * c[1].x));
  c[2] = cmul_si<S>(c[2]);
  c[3] = make_float2(h * (-c[3].x - (float)S * c[3].y), h * (-c[3].y + (float)S * c[3].x));
  dft4<S>(b);
  dft4<S>(c);
#pragma unroll
  for (int p = 0; p < 4; p++) {
    v[2 * p] = b[p];
    v[2 * p + 1] = c[p];
  }
}

template <int S>
__device__ __forceinline__ void dft3(float2* v)
{
#pragma clang fp contract(off)
  constexpr float c3 = -0.5f, s3 = 0.86602540378443864676f;
  float2 t = cadd(v[1], v[2]);
  float2 d = csub(v[1], v[2]);
  float2 m = make_float2(__builtin_fmaf(c3, t.x, v[0].x), __builtin_fmaf(c3, t.y, v[0].y));
  float2 r = cmul_si<S>(make_float2(s3 * d.x, s3 * d.y));
  v[0] = cadd(v[0], t);
  v[1] = cadd(m, r);
  v[2] = csub(m, r);
}

template <int S>
__device__ __forceinline__ void dft5(float2* v)
{
#pragma clang fp contract(off)
  constexpr float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
  constexpr float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
  float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
  float2 d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
  float2 a1 = make_float2(__builtin_fmaf(c2, t2.x, __builtin_fmaf(c1, t1.x, v[0].x)),
                          __builtin_fmaf(c2, t2.y, __builtin_fmaf(c1, t1.y, v[0].y)));
  float2 a2 = make_float2(__builtin_fmaf(c1, t2.x, __builtin_fmaf(c2, t1.x, v[0].x)),
                          __builtin_fmaf(c1, t2.y, __builtin_fmaf(c2, t1.y, v[0].y)));
  float2 b1 = cmul_si<S>(make_float2(__builtin_fmaf(s2, d2.x, s1 * d1.x), __builtin_fmaf(s2, d2.y, s1 * d1.y)));
  float2 b2 = cmul_si<S>(make_float2(__builtin_fmaf(-s1, d2.x, s2 * d1.x), __builtin_fmaf(-s1, d2.y, s2 * d1.y)));
  v[0] = cadd(v[0], cadd(t1, t2));
  v[1] = cadd(a1, b1);
  v[4] = csub(a1, b1);
  v[2] = cadd(a2, b2);
  v[3] = csub(a2, b2);
}

template <int R, int S>
__device__ __forceinline__ void dft(float2* v)
{
  if constexpr (R == 2)
    dft2<S>(v);
  else if constexpr (R == 3)
    dft3<S>(v);
  else if constexpr (R == 4)
    dft4<S>(v);
  else if constexpr (R == 5)
    dft5<S>(v);
  else
    dft8<S>(v);
}

// ---- compile-time factorisation: 8s first, then 4, 2, 3, 5 -------------------------------
template <int L>
struct Factor {
  static constexpr int nst()
  {
    int n = L, c = 0;
    while (n % 8 == 0) { n /= 8; c++; }
    while (n % 4 == 0) { n /= 4; c++; }
    while (n % 2 == 0) { n /= 2; c++; }
    while (n % 3 == 0) { n /= 3; c++; }
    while (n % 5 == 0) { n /= 5; c++; }
    return n == 1 ? c : -1;
  }
  static constexpr int radix(int s)
  {
    int n = L, c = 0;
    const int rs[5] = {8, 4, 2, 3, 5};
    for (int t = 0; t < 5; t++)
      while (n % rs[t] == 0) {
        if (c == s)
          return rs[t];
        n /= rs[t];
        c++;
      }
    return 0;
  }
  static constexpr int ns(int s)
  {
    int p = 1;
    for (int t = 0; t < s; t++)
      p *= radix(t);
    return p;
  }
  static_assert(nst() >= 1, "transform length must be 2^a 3^b 5^c");
};

constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }

// Register footprint of one thread for a tile: the largest stage's butterflies x radix.
template <int L, int T>
struct TileRegs {
  static constexpr int value()
  {
    int m = 0;
    for (int s = 0; s < Factor<L>::nst(); s++) {
      int R = Factor<L>::radix(s);
      int q = cdiv(L / R, T) * R;
      if (q > m)
        m = q;
    }
    return m;
  }
};

// Index policies for the LDS copy of a tile.
//   ColMajor: element e of transform c at lds[e * TC + c]   (transforms adjacent in memory)
//   RowMajor: element e of transform c at lds[c * P + e]     (elements adjacent in memory)
template <int TC>
struct LdsCol {
  __device__ static int at(int c, int e) { return e * TC + c; }
};
template <int P>
struct LdsRow {
  __device__ static int at(int c, int e) { return c * P + e; }
};
// Row-major with one pad slot after every 8 elements: the first radix-8 stage writes its
// outputs 8 apart (j*8 + r), which on the unpadded layout lands 16 lanes on 4 bank groups;
// with the pad consecutive lanes are 9 slots (18 banks) apart and the 16 lanes never collide.
// P must be at least L + L/8 + 1.
template <int P>
struct LdsRowPad {
  __device__ static int at(int c, int e) { return c * P + e + (e >> 3); }
};

// Synchronisation between stages.  SyncBlock: the transforms of a tile span several waves.
// SyncWave: every transform lives inside one wavefront (its threads are consecutive lanes), so
// LDS traffic between stages only has to be ordered within the wave: LDS instructions of a wave
// execute in order, and the wave barrier keeps the compiler from moving LDS accesses across it.
struct SyncBlock {
  __device__ static void sync() { __syncthreads(); }
};
struct SyncWave {
  __device__ static void sync()
  {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
};

// Middle and last stages.  `v` holds the outputs of stage s-1 in its natural slots; on return
// it holds the outputs of the last stage: slot (q, r) <-> element j + r L/Rlast,
// j = t + T q.  lds is reused between stages; tw is the LDS twiddle table of length LT.
template <int L, int T, int S, int LT, class Idx, int s, class Sync = SyncBlock>
__device__ __forceinline__ void stages_from(float2* v, float2* lds, const float2* tw, int c, int t)
{
  using F = Factor<L>;
  if constexpr (s < F::nst()) {
    constexpr int Rp = F::radix(s - 1);
    constexpr int Bp = L / Rp;
    constexpr int Qp = cdiv(Bp, T);
    constexpr int Nsp = F::ns(s - 1);
    // write the outputs of stage s-1 to their Stockham positions
#pragma unroll
    for (int q = 0; q < Qp; q++) {
      const int j = t + T * q;
      if (Bp % T == 0 || j < Bp) {
        const int k = j % Nsp;
        const int base = (j / Nsp) * Nsp * Rp + k;
#pragma unroll
        for (int r = 0; r < Rp; r++)
          lds[Idx::at(c, base + r * Nsp)] = v[q * Rp + r];
      }
    }
    Sync::sync();
    constexpr int R = F::radix(s);
    constexpr int B = L / R;
    constexpr int Q = cdiv(B, T);
    constexpr int Ns = F::ns(s);
    constexpr int step = (LT / L) * (L / (Ns * R));
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const int j = t + T * q;
      if (B % T == 0 || j < B) {
        const int k = j % Ns;
        v[q * R] = lds[Idx::at(c, j)];
#pragma unroll
        for (int r = 1; r < R; r++)
          v[q * R + r] = ctw<S>(lds[Idx::at(c, j + r * B)], tw[k * r * step]);
      }
    }
    Sync::sync();
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const int j = t + T * q;
      if (B % T == 0 || j < B)
        dft<R, S>(v + q * R);
    }
    stages_from<L, T, S, LT, Idx, s + 1, Sync>(v, lds, tw, c, t);
  }
}

// Full transform given stage-0 inputs in v: slot (q, r) = x[j + r L/R0], j = t + T q.
template <int L, int T, int S, int LT, class Idx, class Sync = SyncBlock>
__device__ __forceinline__ void tile_fft(float2* v, float2* lds, const float2* tw, int c, int t)
{
  using F = Factor<L>;
  constexpr int R0 = F::radix(0);
  constexpr int B0 = L / R0;
  constexpr int Q0 = cdiv(B0, T);
#pragma unroll
  for (int q = 0; q < Q0; q++) {
    const int j = t + T * q;
    if (B0 % T == 0 || j < B0)
      dft<R0, S>(v + q * R0);
  }
  stages_from<L, T, S, LT, Idx, 1, Sync>(v, lds, tw, c, t);
}

// Slot geometry of the first and last stage, for callers doing their own loads / stores.
template <int L, int T>
struct Slots {
  using F = Factor<L>;
  static constexpr int R0 = F::radix(0);
  static constexpr int B0 = L / R0;
  static constexpr int Q0 = cdiv(B0, T);
  static constexpr int RL = F::radix(F::nst() - 1);
  static constexpr int BL = L / RL;
  static constexpr int QL = cdiv(BL, T);
  static constexpr int NREG = TileRegs<L, T>::value();
};

}  // namespace mxr
