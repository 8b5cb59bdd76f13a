/* Stand-in for src/core/reionization.h: the declarations mxr_meraxes.c replaces.  Test only. */
#ifndef MXR_SHIM_ENV_REIONIZATION_H
#define MXR_SHIM_ENV_REIONIZATION_H
#include "meraxes.h"
#include "utils.h"
void filter(fftwf_complex* box, int local_ix_start, int slab_nx, int grid_dim, float R, int filter_type);
void velocity_gradient(fftwf_complex* box, int slab_nx, int grid_dim);
double RtoM(double R);
void find_HII_bubbles(int snapshot, timer_info* timer_total);
void ComputeTs(int snapshot, timer_info* timer_total);
void ComputeBrightnessTemperatureBox(int snapshot);
#endif
