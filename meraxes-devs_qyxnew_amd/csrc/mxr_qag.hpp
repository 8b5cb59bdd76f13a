// mxr_qag.hpp -- adaptive Gauss-Kronrod (61) quadrature with the bisection, error and ordering
// rules of GSL's gsl_integration_qag (GSL_INTEG_GAUSS61), which the reference calls, restated
// from QUADPACK's published QAG algorithm for host and device code.  The interval workspace
// holds CAP intervals: a run that would need more stops and reports status 4 (GSL's own limit,
// `limit`, still decides its ordering arithmetic, as in GSL).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "mxr_gk61.hpp"

namespace mxr {

constexpr double QAG_DBL_EPS = 2.2204460492503131e-16;
constexpr double QAG_DBL_MIN = 2.2250738585072014e-308;

__host__ __device__ inline double qag_rescale_error(double err, double result_abs, double result_asc)
{
#pragma clang fp contract(off)
  err = fabs(err);
  if (result_asc != 0 && err != 0) {
    const double scale = pow((200 * err / result_asc), 1.5);
    err = scale < 1 ? result_asc * scale : result_asc;
  }
  if (result_abs > 2 * QAG_DBL_MIN / (50 * QAG_DBL_EPS)) {
    const double min_err = 50 * QAG_DBL_EPS * result_abs;
    if (min_err > err)
      err = min_err;
  }
  return err;
}

// one 61-point Gauss-Kronrod panel on [a, b]
template <class F>
__host__ __device__ inline void qk61(const F& f, double a, double b, double* result, double* abserr, double* resabs,
                                     double* resasc)
{
#pragma clang fp contract(off)
  constexpr int n = 31;
  double fv1[31], fv2[31];
  const double center = 0.5 * (a + b);
  const double half_length = 0.5 * (b - a);
  const double abs_half_length = fabs(half_length);
  const double f_center = f(center);
  double result_gauss = 0;
  double result_kronrod = f_center * GK61_WGK[n - 1];
  double result_abs = fabs(result_kronrod);
  for (int j = 0; j < (n - 1) / 2; j++) {
    const int jtw = j * 2 + 1;
    const double abscissa = half_length * GK61_XGK[jtw];
    const double fval1 = f(center - abscissa), fval2 = f(center + abscissa);
    const double fsum = fval1 + fval2;
    fv1[jtw] = fval1;
    fv2[jtw] = fval2;
    result_gauss += GK61_WG[j] * fsum;
    result_kronrod += GK61_WGK[jtw] * fsum;
    result_abs += GK61_WGK[jtw] * (fabs(fval1) + fabs(fval2));
  }
  for (int j = 0; j < n / 2; j++) {
    const int jtwm1 = j * 2;
    const double abscissa = half_length * GK61_XGK[jtwm1];
    const double fval1 = f(center - abscissa), fval2 = f(center + abscissa);
    fv1[jtwm1] = fval1;
    fv2[jtwm1] = fval2;
    result_kronrod += GK61_WGK[jtwm1] * (fval1 + fval2);
    result_abs += GK61_WGK[jtwm1] * (fabs(fval1) + fabs(fval2));
  }
  const double mean = result_kronrod * 0.5;
  double result_asc = GK61_WGK[n - 1] * fabs(f_center - mean);
  for (int j = 0; j < n - 1; j++)
    result_asc += GK61_WGK[j] * (fabs(fv1[j] - mean) + fabs(fv2[j] - mean));
  const double err = (result_kronrod - result_gauss) * half_length;
  result_kronrod *= half_length;
  result_abs *= abs_half_length;
  result_asc *= abs_half_length;
  *result = result_kronrod;
  *resabs = result_abs;
  *resasc = result_asc;
  *abserr = qag_rescale_error(err, result_abs, result_asc);
}

template <int CAP>
struct QagWork {
  double alist[CAP], blist[CAP], rlist[CAP], elist[CAP];
  int order[CAP], level[CAP];
  int size, nrmax, i, maximum_level, limit;
};

// keeps order[] sorted by decreasing error over the part of the list still to be bisected
template <int CAP>
__host__ __device__ inline void qag_qpsrt(QagWork<CAP>& w)
{
#pragma clang fp contract(off)
  const int last = w.size - 1, limit = w.limit;
  int i_nrmax = w.nrmax, i_maxerr = w.order[i_nrmax];
  if (last < 2) {
    w.order[0] = 0;
    w.order[1] = 1;
    w.i = i_maxerr;
    return;
  }
  const double errmax = w.elist[i_maxerr];
  while (i_nrmax > 0 && errmax > w.elist[w.order[i_nrmax - 1]]) {
    w.order[i_nrmax] = w.order[i_nrmax - 1];
    i_nrmax--;
  }
  const int top = (last < (limit / 2 + 2)) ? last : (limit - last + 1);
  int i = i_nrmax + 1;
  while (i < top && errmax < w.elist[w.order[i]]) {
    w.order[i - 1] = w.order[i];
    i++;
  }
  w.order[i - 1] = i_maxerr;
  const double errmin = w.elist[last];
  int k = top - 1;
  while (k > i - 2 && errmin >= w.elist[w.order[k]]) {
    w.order[k + 1] = w.order[k];
    k--;
  }
  w.order[k + 1] = last;
  i_maxerr = w.order[i_nrmax];
  w.i = i_maxerr;
  w.nrmax = i_nrmax;
}

template <int CAP>
__host__ __device__ inline void qag_update(QagWork<CAP>& w, double a1, double b1, double area1, double error1,
                                           double a2, double b2, double area2, double error2)
{
#pragma clang fp contract(off)
  const int i_max = w.i, i_new = w.size;
  const int new_level = w.level[i_max] + 1;
  if (error2 > error1) {
    w.alist[i_max] = a2;
    w.rlist[i_max] = area2;
    w.elist[i_max] = error2;
    w.level[i_max] = new_level;
    w.alist[i_new] = a1;
    w.blist[i_new] = b1;
    w.rlist[i_new] = area1;
    w.elist[i_new] = error1;
    w.level[i_new] = new_level;
  } else {
    w.blist[i_max] = b1;
    w.rlist[i_max] = area1;
    w.elist[i_max] = error1;
    w.level[i_max] = new_level;
    w.alist[i_new] = a2;
    w.blist[i_new] = b2;
    w.rlist[i_new] = area2;
    w.elist[i_new] = error2;
    w.level[i_new] = new_level;
  }
  w.size++;
  if (new_level > w.maximum_level)
    w.maximum_level = new_level;
  qag_qpsrt(w);
}

// Returns 0 when the tolerance is reached and GSL's nonzero status otherwise (1 roundoff on the
// first panel, 2 limit 1, 3 tolerance not reached), or 4 when the CAP-interval workspace is full.
template <int CAP, class F>
__host__ __device__ inline int qag(QagWork<CAP>& w, const F& f, double a, double b, double epsabs, double epsrel,
                                   int limit, double* result, double* abserr)
{
#pragma clang fp contract(off)
  w.limit = limit;
  w.size = 0;
  w.nrmax = 0;
  w.i = 0;
  w.alist[0] = a;
  w.blist[0] = b;
  w.rlist[0] = 0.0;
  w.elist[0] = 0.0;
  w.order[0] = 0;
  w.level[0] = 0;
  w.maximum_level = 0;
  *result = 0;
  *abserr = 0;
  double result0, abserr0, resabs0, resasc0;
  qk61(f, a, b, &result0, &abserr0, &resabs0, &resasc0);
  w.size = 1;
  w.rlist[0] = result0;
  w.elist[0] = abserr0;
  double tolerance = fmax(epsabs, epsrel * fabs(result0));
  const volatile double round_off = 50 * QAG_DBL_EPS * resabs0;
  if (abserr0 <= round_off && abserr0 > tolerance) {
    *result = result0;
    *abserr = abserr0;
    return 1;
  } else if ((abserr0 <= tolerance && abserr0 != resasc0) || abserr0 == 0.0) {
    *result = result0;
    *abserr = abserr0;
    return 0;
  } else if (limit == 1) {
    *result = result0;
    *abserr = abserr0;
    return 2;
  }
  double area = result0, errsum = abserr0;
  int iteration = 1;
  int roundoff_type1 = 0, roundoff_type2 = 0, error_type = 0;
  do {
    if (w.size >= CAP) {
      error_type = 4;
      break;
    }
    const int ii = w.i;
    const double a_i = w.alist[ii], b_i = w.blist[ii], r_i = w.rlist[ii], e_i = w.elist[ii];
    const double a1 = a_i, b1 = 0.5 * (a_i + b_i), a2 = b1, b2 = b_i;
    double area1, error1, resabs1, resasc1, area2, error2, resabs2, resasc2;
    qk61(f, a1, b1, &area1, &error1, &resabs1, &resasc1);
    qk61(f, a2, b2, &area2, &error2, &resabs2, &resasc2);
    const double area12 = area1 + area2, error12 = error1 + error2;
    errsum += (error12 - e_i);
    area += area12 - r_i;
    if (resasc1 != error1 && resasc2 != error2) {
      const double delta = r_i - area12;
      if (fabs(delta) <= 1.0e-5 * fabs(area12) && error12 >= 0.99 * e_i)
        roundoff_type1++;
      if (iteration >= 10 && error12 > e_i)
        roundoff_type2++;
    }
    tolerance = fmax(epsabs, epsrel * fabs(area));
    if (errsum > tolerance) {
      if (roundoff_type1 >= 6 || roundoff_type2 >= 20)
        error_type = 2;
      const double tmp = (1 + 100 * QAG_DBL_EPS) * (fabs(a2) + 1000 * QAG_DBL_MIN);
      if (fabs(a1) <= tmp && fabs(b2) <= tmp)
        error_type = 3;
    }
    qag_update(w, a1, b1, area1, error1, a2, b2, area2, error2);
    iteration++;
  } while (iteration < limit && !error_type && errsum > tolerance);
  double s = 0;
  for (int k = 0; k < w.size; k++)
    s += w.rlist[k];
  *result = s;
  *abserr = errsum;
  if (error_type == 4)
    return 4;
  return errsum <= tolerance ? 0 : 3;
}

}  // namespace mxr
