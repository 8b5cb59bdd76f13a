/* fft.c -- FP32 and FP64 instantiations of the oracle FFT (test infrastructure only). */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"

#define REAL float
#define SUFFIX f32
#include "fft_tmpl.h"
#undef REAL
#undef SUFFIX

#define REAL double
#define SUFFIX f64
#include "fft_tmpl.h"
#undef REAL
#undef SUFFIX

int oracle_c2c1d_f32(int n, float* data, int sign) { return c2c1d_f32(n, (cplx_f32*)data, sign); }
int oracle_c2c1d_f64(int n, double* data, int sign) { return c2c1d_f64(n, (cplx_f64*)data, sign); }
int oracle_r2c3d_f32(int N, const float* in, float* out) { return r2c3d_f32(N, in, (cplx_f32*)out); }
int oracle_r2c3d_f64(int N, const double* in, double* out) { return r2c3d_f64(N, in, (cplx_f64*)out); }
int oracle_c2r3d_f32(int N, float* inout) { return c2r3d_f32(N, (cplx_f32*)inout); }
int oracle_c2r3d_f64(int N, double* inout) { return c2r3d_f64(N, (cplx_f64*)inout); }

int oracle_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
