/* A host for mxr_meraxes.c: the globals and helpers the reference defines elsewhere, and a main
 * that calls the binding's host-side entry point RtoM and the reference-named wrappers' symbols.
 * Test only. */
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "meraxes.h"
#include "mxr.h"
#include "reionization.h"

run_globals_t run_globals;
float timer_gpu = 0.f;
double* stored_fcoll;
double* stored_XrayEmissivity_HMXB;
double* stored_XrayEmissivity_hard;
double* stored_XrayEmissivity_soft;

void myexit(int signum) { exit(signum); }
void mlog_error(const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vfprintf(stderr, fmt, ap);
  va_end(ap);
  fputc('\n', stderr);
}
void timer_start(timer_info* t) { gettimeofday(&t->start, NULL); }
void timer_stop(timer_info* t) { gettimeofday(&t->stop, NULL); }
float timer_delta(timer_info t)
{
  return (float)(t.stop.tv_sec - t.start.tv_sec) + 1e-6f * (float)(t.stop.tv_usec - t.start.tv_usec);
}

int main(void)
{
  static double zz[2] = {9.0, 8.0};
  run_globals.ZZ = zz;
  run_globals.mpi_size = 1;
  run_globals.params.OmegaM = 0.3121;
  run_globals.params.OmegaLambda = 0.6879;
  run_globals.params.Hubble_h = 0.6751;
  run_globals.params.BaryonFrac = 0.15714;
  run_globals.params.BoxSize = 67.8;
  run_globals.params.ReionGridDim = 64;
  run_globals.params.ReionRtoMFilterType = 0;
  run_globals.units.UnitLength_in_cm = 3.08568e+24;
  run_globals.units.UnitMass_in_g = 1.989e+43;
  run_globals.units.UnitVelocity_in_cm_per_s = 100000;
  /* RtoM through the binding equals the library's own */
  mxr_hii_params p;
  memset(&p, 0, sizeof(p));
  p.OmegaM = run_globals.params.OmegaM;
  p.OmegaLambda = run_globals.params.OmegaLambda;
  p.Hubble_h = run_globals.params.Hubble_h;
  p.ReionRtoMFilterType = 0;
  p.UnitLength_in_cm = run_globals.units.UnitLength_in_cm;
  p.UnitMass_in_g = run_globals.units.UnitMass_in_g;
  p.UnitVelocity_in_cm_per_s = run_globals.units.UnitVelocity_in_cm_per_s;
  const double a = RtoM(5.0), b = mxr_RtoM(&p, 5.0);
  printf("RtoM %.17g %.17g\n", a, b);
  /* the replaced entry points exist with the reference's signatures */
  void (*f1)(int, timer_info*) = find_HII_bubbles;
  void (*f2)(int, timer_info*) = ComputeTs;
  void (*f3)(int) = ComputeBrightnessTemperatureBox;
  void (*f4)(fftwf_complex*, int, int, int, float, int) = filter;
  void (*f5)(fftwf_complex*, int, int) = velocity_gradient;
  printf("entry points %d\n", f1 && f2 && f3 && f4 && f5);
  return (a == b && a > 0) ? 0 : 1;
}
