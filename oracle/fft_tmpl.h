/*
 * fft_tmpl.h -- CPU FFT of the oracle, instantiated twice by fft.c (REAL = float, double).
 *
 * TEST INFRASTRUCTURE ONLY: the oracle is the checker of the HIP path, never part of it.
 *
 * Restates the transforms the reference takes from FFTW3 (third-party, not vendored; the
 * reference's plans are fftwf_mpi_plan_dft_r2c_3d / _c2r_3d at src/core/reionization.c:
 * 907-997, executed at src/core/find_HII_bubbles.c:166-229 and 351-370):
 *   r2c: X[k] = sum_n x[n] exp(-2 pi i k.n / N)   (unnormalised, out of place, padded input)
 *   c2r: x[n] = sum_k X[k] exp(+2 pi i k.n / N)   (unnormalised, in place, padded output)
 * over the half spectrum k_z in [0, N/2].  The 1-D kernel is a Stockham autosort FFT
 * (Van Loan, "Computational Frameworks for the FFT", 1992, section 1.7), radices 4, 2, 3, 5,
 * twiddles evaluated in double and rounded to REAL.
 * FFTW's own rounding order cannot be reproduced: parity at this boundary is to tolerance.
 */
#ifndef REAL
#error "define REAL before including fft_tmpl.h"
#endif

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define CPLX CAT(cplx_, SUFFIX)
#define PLAN CAT(fftplan_, SUFFIX)

typedef struct {
  REAL re, im;
} CPLX;

typedef struct {
  int n;
  int nstage;
  int radix[32];
  long twoff[32];
  CPLX* tw; /* per-stage twiddle blocks, see plan_init */
} PLAN;

static int CAT(plan_init_, SUFFIX)(PLAN* p, int n)
{
  p->n = n;
  p->nstage = 0;
  int m = n;
  static const int order[4] = {4, 2, 3, 5};
  for (int t = 0; t < 4; t++) {
    int r = order[t];
    while (m % r == 0) {
      p->radix[p->nstage++] = r;
      m /= r;
    }
  }
  if (m != 1)
    return -1;
  /* per-stage twiddles exp(-2 pi i k r / (Ns R)), k < Ns, r in [1, R), evaluated in double */
  long total = 0;
  int Ns = 1;
  for (int s = 0; s < p->nstage; s++) {
    total += (long)Ns * (p->radix[s] - 1);
    Ns *= p->radix[s];
  }
  p->tw = (CPLX*)malloc(sizeof(CPLX) * (size_t)(total > 0 ? total : 1));
  long off = 0;
  Ns = 1;
  for (int s = 0; s < p->nstage; s++) {
    const int R = p->radix[s];
    p->twoff[s] = off;
    for (int k = 0; k < Ns; k++)
      for (int r = 1; r < R; r++) {
        double a = -2.0 * M_PI * (double)k * (double)r / (double)(Ns * R);
        p->tw[off].re = (REAL)cos(a);
        p->tw[off].im = (REAL)sin(a);
        off++;
      }
    Ns *= R;
  }
  return 0;
}

static void CAT(plan_free_, SUFFIX)(PLAN* p)
{
  free(p->tw);
  p->tw = NULL;
}

/* (a + ib) * (c + i s d): s = -1 forward table, s = +1 conjugated for the inverse */
#define CMUL(o, x, w, sgn)                                                                     \
  do {                                                                                         \
    const REAL wi_ = (sgn) > 0 ? -(w).im : (w).im;                                             \
    const REAL xr_ = (x).re, xi_ = (x).im;                                                     \
    (o).re = xr_ * (w).re - xi_ * wi_;                                                         \
    (o).im = xr_ * wi_ + xi_ * (w).re;                                                         \
  } while (0)

/* In-place 1-D transform of x[0..n) (contiguous); work holds n values.
 * Butterflies: radix-2/4 (exact +-1, +-i rotations), radix-3 and radix-5 from the
 * closed-form cos/sin of 2 pi/3 and 2 pi/5, 4 pi/5. */
static void CAT(fft1d_, SUFFIX)(const PLAN* p, CPLX* x, CPLX* work, int sign)
{
  const int n = p->n;
  CPLX* a = x;
  CPLX* b = work;
  int Ns = 1;
  const REAL sg = (REAL)sign;
  const REAL c3 = (REAL)(-0.5), s3 = (REAL)(0.86602540378443864676);
  const REAL c51 = (REAL)(0.30901699437494742410), c52 = (REAL)(-0.80901699437494742410);
  const REAL s51 = (REAL)(0.95105651629515357212), s52 = (REAL)(0.58778525229247312917);
  for (int s = 0; s < p->nstage; s++) {
    const int R = p->radix[s];
    const int nR = n / R;
    const CPLX* tw = p->tw + p->twoff[s];
    for (int j = 0; j < nR; j++) {
      const int k = j % Ns;
      CPLX v[5];
      v[0] = a[j];
      for (int r = 1; r < R; r++)
        CMUL(v[r], a[j + r * nR], tw[k * (R - 1) + r - 1], sign);
      CPLX V[5];
      if (R == 2) {
        V[0].re = v[0].re + v[1].re; V[0].im = v[0].im + v[1].im;
        V[1].re = v[0].re - v[1].re; V[1].im = v[0].im - v[1].im;
      } else if (R == 4) {
        const REAL t0r = v[0].re + v[2].re, t0i = v[0].im + v[2].im;
        const REAL t1r = v[0].re - v[2].re, t1i = v[0].im - v[2].im;
        const REAL t2r = v[1].re + v[3].re, t2i = v[1].im + v[3].im;
        /* (v1 - v3) * (sign i) */
        const REAL dr = v[1].re - v[3].re, di = v[1].im - v[3].im;
        const REAL t3r = -sg * di, t3i = sg * dr;
        V[0].re = t0r + t2r; V[0].im = t0i + t2i;
        V[2].re = t0r - t2r; V[2].im = t0i - t2i;
        V[1].re = t1r + t3r; V[1].im = t1i + t3i;
        V[3].re = t1r - t3r; V[3].im = t1i - t3i;
      } else if (R == 3) {
        const REAL tr = v[1].re + v[2].re, ti = v[1].im + v[2].im;
        const REAL dr = v[1].re - v[2].re, di = v[1].im - v[2].im;
        const REAL mr = v[0].re + c3 * tr, mi = v[0].im + c3 * ti;
        /* sign i s3 (v1 - v2) */
        const REAL rr = -sg * s3 * di, ri = sg * s3 * dr;
        V[0].re = v[0].re + tr; V[0].im = v[0].im + ti;
        V[1].re = mr + rr; V[1].im = mi + ri;
        V[2].re = mr - rr; V[2].im = mi - ri;
      } else { /* R == 5 */
        const REAL t1r = v[1].re + v[4].re, t1i = v[1].im + v[4].im;
        const REAL t2r = v[2].re + v[3].re, t2i = v[2].im + v[3].im;
        const REAL d1r = v[1].re - v[4].re, d1i = v[1].im - v[4].im;
        const REAL d2r = v[2].re - v[3].re, d2i = v[2].im - v[3].im;
        const REAL a1r = v[0].re + c51 * t1r + c52 * t2r, a1i = v[0].im + c51 * t1i + c52 * t2i;
        const REAL a2r = v[0].re + c52 * t1r + c51 * t2r, a2i = v[0].im + c52 * t1i + c51 * t2i;
        const REAL b1r = s51 * d1r + s52 * d2r, b1i = s51 * d1i + s52 * d2i;
        const REAL b2r = s52 * d1r - s51 * d2r, b2i = s52 * d1i - s51 * d2i;
        V[0].re = v[0].re + t1r + t2r; V[0].im = v[0].im + t1i + t2i;
        V[1].re = a1r - sg * b1i; V[1].im = a1i + sg * b1r;
        V[4].re = a1r + sg * b1i; V[4].im = a1i - sg * b1r;
        V[2].re = a2r - sg * b2i; V[2].im = a2i + sg * b2r;
        V[3].re = a2r + sg * b2i; V[3].im = a2i - sg * b2r;
      }
      const int idxD = (j / Ns) * Ns * R + k;
      for (int q = 0; q < R; q++)
        b[idxD + q * Ns] = V[q];
    }
    CPLX* t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  if (a != x)
    memcpy(x, a, sizeof(CPLX) * (size_t)n);
}

#define COLBLK 16

/* Transform along a strided axis: ncol columns, column c starts at base(c), elements at
 * stride es.  Columns are grouped in blocks of COLBLK adjacent ones for cache locality. */
static void CAT(fft_axis_, SUFFIX)(const PLAN* p, CPLX* data, long n_outer, long outer_stride, long ncol,
                                   long es, int sign)
{
  const int n = p->n;
  const long nblk = (ncol + COLBLK - 1) / COLBLK;
#pragma omp parallel
  {
    CPLX* buf = (CPLX*)malloc(sizeof(CPLX) * (size_t)n * COLBLK);
    CPLX* work = (CPLX*)malloc(sizeof(CPLX) * (size_t)n);
#pragma omp for schedule(static)
    for (long t = 0; t < n_outer * nblk; t++) {
      const long o = t / nblk;
      const long c0 = (t % nblk) * COLBLK;
      const long nc = (ncol - c0 < COLBLK) ? (ncol - c0) : COLBLK;
      CPLX* base = data + o * outer_stride + c0;
      for (int e = 0; e < n; e++)
        for (long c = 0; c < nc; c++)
          buf[c * n + e] = base[(long)e * es + c];
      for (long c = 0; c < nc; c++)
        CAT(fft1d_, SUFFIX)(p, buf + c * n, work, sign);
      for (int e = 0; e < n; e++)
        for (long c = 0; c < nc; c++)
          base[(long)e * es + c] = buf[c * n + e];
    }
    free(work);
    free(buf);
  }
}

/* 3-D r2c, FFTW semantics: in = padded real (N*N*2H), out = half spectrum (N*N*H). */
static int CAT(r2c3d_, SUFFIX)(int N, const REAL* in, CPLX* out)
{
  const long H = N / 2 + 1;
  PLAN p;
  if (CAT(plan_init_, SUFFIX)(&p, N))
    return -1;
#pragma omp parallel
  {
    CPLX* row = (CPLX*)malloc(sizeof(CPLX) * (size_t)N);
    CPLX* work = (CPLX*)malloc(sizeof(CPLX) * (size_t)N);
#pragma omp for schedule(static)
    for (long r = 0; r < (long)N * N; r++) {
      const REAL* src = in + r * 2 * H;
      for (int k = 0; k < N; k++) {
        row[k].re = src[k];
        row[k].im = 0;
      }
      CAT(fft1d_, SUFFIX)(&p, row, work, -1);
      for (long k = 0; k < H; k++)
        out[r * H + k] = row[k];
    }
    free(work);
    free(row);
  }
  /* y: for each x-plane, H columns of stride H */
  CAT(fft_axis_, SUFFIX)(&p, out, N, (long)N * H, H, H, -1);
  /* x: one group of N*H columns of stride N*H */
  CAT(fft_axis_, SUFFIX)(&p, out, 1, 0, (long)N * H, (long)N * H, -1);
  CAT(plan_free_, SUFFIX)(&p);
  return 0;
}

/* 3-D c2r in place, FFTW semantics: data holds the half spectrum on entry and the padded
 * real field on exit.  The imaginary parts of the k_z = 0 and k_z = N/2 planes after the
 * x/y inverse transforms do not contribute (the real part of the Hermitian extension). */
static int CAT(c2r3d_, SUFFIX)(int N, CPLX* data)
{
  const long H = N / 2 + 1;
  PLAN p;
  if (CAT(plan_init_, SUFFIX)(&p, N))
    return -1;
  CAT(fft_axis_, SUFFIX)(&p, data, 1, 0, (long)N * H, (long)N * H, +1);
  CAT(fft_axis_, SUFFIX)(&p, data, N, (long)N * H, H, H, +1);
#pragma omp parallel
  {
    CPLX* row = (CPLX*)malloc(sizeof(CPLX) * (size_t)N);
    CPLX* work = (CPLX*)malloc(sizeof(CPLX) * (size_t)N);
#pragma omp for schedule(static)
    for (long r = 0; r < (long)N * N; r++) {
      CPLX* src = data + r * H;
      for (long k = 0; k < H; k++)
        row[k] = src[k];
      for (long k = 1; k < N - H + 1; k++) {
        row[N - k].re = src[k].re;
        row[N - k].im = -src[k].im;
      }
      CAT(fft1d_, SUFFIX)(&p, row, work, +1);
      REAL* dst = (REAL*)src;
      for (int k = 0; k < N; k++)
        dst[k] = row[k].re;
      dst[N] = 0;
      dst[N + 1] = 0;
    }
    free(work);
    free(row);
  }
  CAT(plan_free_, SUFFIX)(&p);
  return 0;
}

/* 1-D complex transform of a contiguous array (used by tests to pin the kernel). */
static int CAT(c2c1d_, SUFFIX)(int n, CPLX* x, int sign)
{
  PLAN p;
  if (CAT(plan_init_, SUFFIX)(&p, n))
    return -1;
  CPLX* work = (CPLX*)malloc(sizeof(CPLX) * (size_t)n);
  CAT(fft1d_, SUFFIX)(&p, x, work, sign);
  free(work);
  CAT(plan_free_, SUFFIX)(&p);
  return 0;
}

#undef CMUL
#undef CPLX
#undef PLAN
#undef CAT
#undef CAT_
