/*
 * grid.c -- CPU restatement of the input-grid preparation (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ as the checker.
 *
 * One rank of read_grid__gbptrees (src/core/read_grids-gbptrees.c:88-231) with smooth_grid and
 * subsample_grid (src/core/read_grids.c:128-287), and the density step of the VELOCIraptor
 * reader (read_grids-velociraptor.c:213-230).  FFTs are the oracle's FFTW-semantics FP32
 * transforms; the filter is oracle_filter_f32.  Padding entries of the output are zero.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define O_REL_TOL (float)1e-5 /* src/meraxes.h:66 */

int oracle_prepare_grid(int N, const float* hi, int n_hi, int property, double file_box, const mxr_grid_params* p,
                        float* out)
{
  if (n_hi < N || n_hi % N != 0)
    return MXR_ERR_INVALID_ARG;
  const long Hh = n_hi / 2 + 1, H = N / 2 + 1;
  /* the file's grid in the padded layout (:165-169 reorder) */
  float* slab_file = (float*)calloc((size_t)2 * n_hi * n_hi * Hh, sizeof(float));
  if (!slab_file)
    return MXR_ERR_OOM;
  for (long i = 0; i < (long)n_hi * n_hi; i++)
    memcpy(slab_file + i * 2 * Hh, hi + i * n_hi, sizeof(float) * n_hi);
  if (n_hi > N) {
    /* smooth_grid (read_grids.c:128-165) */
    float* spec = (float*)malloc(sizeof(float) * 2 * n_hi * n_hi * Hh);
    if (!spec) {
      free(slab_file);
      return MXR_ERR_OOM;
    }
    oracle_r2c3d_f32(n_hi, slab_file, spec);
    const double total_n_cells = (double)n_hi * n_hi * n_hi;
    for (long i = 0; i < 2L * n_hi * n_hi * Hh; i++)
      spec[i] = (float)((double)spec[i] / total_n_cells);
    oracle_filter_f32(spec, 0, n_hi, n_hi, p->BoxSize, (float)(p->BoxSize / (double)N / 2.0), 0);
    oracle_c2r3d_f32(n_hi, spec);
    free(slab_file);
    slab_file = spec;
  }
  /* subsample_grid (read_grids.c:167-287), one rank */
  const int n_every = n_hi / N;
  memset(out, 0, sizeof(float) * 2 * N * N * H);
  for (int ix = 0; ix < N; ix++)
    for (int iy = 0; iy < N; iy++)
      for (int iz = 0; iz < N; iz++)
        out[iz + 2 * H * (iy + (long)N * ix)] =
          slab_file[iz * n_every + 2 * Hh * (iy * n_every + (long)n_hi * ix * n_every)];
  free(slab_file);
  if (property == MXR_GRID_DENSITY) {
    float mean_inv, floor_v;
    if (p->TreesID == MXR_TREES_GBPTREES) {
      mean_inv = pow(file_box, 3) * p->Hubble_h / p->NPart / p->PartMass;
      floor_v = -1.0;
    } else {
      mean_inv = pow(file_box, 3) * p->Hubble_h / ((double)p->NPart * p->PartMass);
      floor_v = -1.0 + O_REL_TOL;
    }
    for (int ii = 0; ii < N; ii++)
      for (int jj = 0; jj < N; jj++)
        for (int kk = 0; kk < N; kk++) {
          float* val = &out[kk + 2 * H * (jj + (long)N * ii)];
          *val = fmaxf(*val * mean_inv - 1.0, floor_v);
        }
  }
  return 0;
}

int oracle_read_grid_gbptrees(const char* fname, int property, int N, const mxr_grid_params* p, float* out)
{
  FILE* fd = fopen(fname, "rb");
  if (!fd)
    return MXR_ERR_INVALID_ARG;
  int n_cell[3], n_grids, ma_scheme;
  double box_size[3];
  if (fread(n_cell, sizeof(int), 3, fd) != 3 || fread(box_size, sizeof(double), 3, fd) != 3 ||
      fread(&n_grids, sizeof(int), 1, fd) != 1 || fread(&ma_scheme, sizeof(int), 1, fd) != 1 || n_grids != 4 ||
      n_cell[0] != n_cell[1] || n_cell[1] != n_cell[2]) {
    fclose(fd);
    return MXR_ERR_INVALID_ARG;
  }
  const long n3 = (long)n_cell[0] * n_cell[1] * n_cell[2];
  char identifier[32];
  for (int ii = 0; ii < property; ii++) {
    if (fread(identifier, sizeof(identifier), 1, fd) != 1 || fseek(fd, sizeof(float) * n3, SEEK_CUR) != 0) {
      fclose(fd);
      return MXR_ERR_INVALID_ARG;
    }
  }
  float* hi = (float*)malloc(sizeof(float) * n3);
  if (!hi || fread(identifier, sizeof(identifier), 1, fd) != 1 || (long)fread(hi, sizeof(float), n3, fd) != n3) {
    free(hi);
    fclose(fd);
    return MXR_ERR_INVALID_ARG;
  }
  fclose(fd);
  mxr_grid_params q = *p;
  q.TreesID = MXR_TREES_GBPTREES;
  const int rc = oracle_prepare_grid(N, hi, n_cell[0], property, box_size[0], &q, out);
  free(hi);
  return rc;
}
