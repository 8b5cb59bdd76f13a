// mxr_grid.hip -- the input density and velocity grids (read_grids*.c) for one rank: a
// gbpTrees grid file read on the host, then on the GPU the smoothing and subsampling of a
// higher-resolution grid (smooth_grid, subsample_grid: src/core/read_grids.c:128-287) and the
// density normalisation of the two readers.
//
// Smoothing runs on a second context of the file's dimension (created on first use and kept
// with the caller's context): r2c, /n^3 in double, the real-space top-hat filter, c2r.  One
// kernel then takes every (n_hi/N)-th point into the caller's padded grid and normalises it.
#include <cmath>
#include <cstdio>
#include <vector>

#include "mxr_passes.hpp"

namespace mxr {

namespace {

struct Finish {
  int N, every, n_hi;
  int density;
  float mean_inv, floor_v;
  const float* hi;  // padded (smoothed) or real (n_hi == N, unsmoothed) source
  int hi_padded;
  float* out;       // padded N^3
};

// subsample_grid (read_grids.c:167-287) then the density step (read_grids-gbptrees.c:201-218,
// read_grids-velociraptor.c:213-230): *val = fmaxf(*val * mean_inv - 1.0, floor)
__global__ void k_grid_finish(Finish f)
{
#pragma clang fp contract(off)
  const int N = f.N, H2 = 2 * (N / 2 + 1);
  const long npad = (long)N * N * H2;
  // the smoothed grid is the c2r output in a spectrum buffer: rows of 2 spec_pitch(n_hi) floats
  const long hrow = f.hi_padded ? 2 * spec_pitch(f.n_hi) : f.n_hi;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < npad; p += (long)gridDim.x * blockDim.x) {
    const int iz = (int)(p % H2);
    float v = 0.f;
    if (iz < N) {
      const long r = p / H2;
      const int iy = (int)(r % N), ix = (int)(r / N);
      v = f.hi[((long)ix * f.every * f.n_hi + (long)iy * f.every) * hrow + (long)iz * f.every];
      if (f.density)
        v = fmaxf((float)((double)(v * f.mean_inv) - 1.0), f.floor_v);
    }
    f.out[p] = v;
  }
}

// smooth_grid (read_grids.c:128-165): r2c, /n^3, filter(R, 0), c2r; the result is the padded
// real grid in hc->spec[0].  r2c reads the file's real-layout rows directly; the filter is the
// window table's multiply on the inverse x pass's loads (the same product filter() forms, each
// mode rounded once), the table built once per (n, R) and kept.
template <int N>
int smooth(Ctx* hc, const float* hi_real, double BoxSize, int N_lo)
{
  using K = Passes<N>;
  RC(ensure_buffers(hc, 1));
  CRealPtrs rin{};
  rin.p[0] = hi_real;
  rin.real = 1u;
  GridPtrs spec{};
  spec.p[0] = hc->spec[0];
  CGridPtrs cspec{};
  cspec.p[0] = hc->spec[0];
  FiltArgs none{};
  none.N = N;
  RC(K::r2c(hc, rin, spec, 1));
  RC((K::template col<-1, COL_PLAIN>(hc, cspec, spec, 1, K::ygeo(), none, 1.0)));
  RC((K::template col<-1, COL_SCALE>(hc, cspec, spec, 1, K::xgeo(), none, std::pow((double)N, 3))));
  // filter(slab, ..., n_cell[0], (float)(BoxSize / ReionGridDim / 2.0), 0): delta_k from the
  // float box size (reionization.c:3376-3377)
  FiltArgs fa{};
  fa.N = N;
  fa.dk = (float)(2.0 * M_PI / (float)BoxSize);
  fa.R = (float)(BoxSize / (double)N_lo / 2.0);
  fa.ftype = 0;
  fa.apply = 1;
  RC(window_table(hc, fa.dk, fa.R, fa.ftype, &fa.wtab));
  RC((K::template col<+1, COL_FILTER>(hc, cspec, spec, 1, K::xgeo(), fa, 1.0)));
  RC((K::template col<+1, COL_PLAIN>(hc, cspec, spec, 1, K::ygeo(), none, 1.0)));
  return run_c2r_store(hc, hc->spec[0]);
}

}  // namespace

// out: the caller's padded N^3 grid on the device
int run_prepare_grid(Ctx* c, const float* hi, int n_hi, int mem_in, int property, double file_box,
                     const mxr_grid_params* p, float* out)
{
  const int N = c->N;
  const long nreal_hi = (long)n_hi * n_hi * n_hi;
  Finish f{};
  f.N = N;
  f.n_hi = n_hi;
  f.every = n_hi / N;
  f.density = property == MXR_GRID_DENSITY;
  if (p->TreesID == MXR_TREES_GBPTREES) {
    // read_grids-gbptrees.c:203-204 (the Hubble factor corrects the grids' units)
    f.mean_inv = (float)(std::pow(file_box, 3) * p->Hubble_h / (double)p->NPart / p->PartMass);
    f.floor_v = -1.0f;
  } else {
    // read_grids-velociraptor.c:215-216, floor -1 + REL_TOL with REL_TOL = (float)1e-5
    f.mean_inv = (float)(std::pow(file_box, 3) * p->Hubble_h / ((double)p->NPart * p->PartMass));
    f.floor_v = (float)(-1.0 + (double)(float)1e-5);
  }
  // the hi-res grid on the device, 8-byte aligned (r2c reads it in pairs)
  const float* hd = hi;
  float* staged = nullptr;
  if (mem_in == MXR_MEM_HOST || (reinterpret_cast<uintptr_t>(hi) & 7u) != 0) {
    MXR_HIP(hipMalloc(&staged, sizeof(float) * nreal_hi));
    hipError_t e = hipMemcpyAsync(staged, hi, sizeof(float) * nreal_hi,
                                  mem_in == MXR_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess)
      e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
      (void)hipFree(staged);
      return hip_fail(e, "staging the input grid", __FILE__, __LINE__);
    }
    hd = staged;
  }
  int rc = MXR_OK;
  if (n_hi == N) {
    f.hi = hd;
    f.hi_padded = 0;
  } else {
    if (!c->hires || c->hires->c.N != n_hi) {
      if (c->hires)
        mxr_destroy(c->hires);
      c->hires = nullptr;
      rc = mxr_create(&c->hires, c->device, n_hi);
    }
    if (rc == MXR_OK) {
      Ctx* hc = &c->hires->c;
      switch (n_hi) {
#define X(n)                                                                                                         \
  case n: rc = smooth<n>(hc, hd, p->BoxSize, N); break;
        MXR_DIMS(X)
#undef X
        default: set_error("grid dimension %d is not supported for smoothing", n_hi); rc = MXR_ERR_UNSUPPORTED;
      }
      if (rc == MXR_OK) {
        hipError_t e = hipStreamSynchronize(hc->stream);
        if (e != hipSuccess)
          rc = hip_fail(e, "smooth_grid", __FILE__, __LINE__);
      }
      f.hi = reinterpret_cast<const float*>(hc->spec[0]);
      f.hi_padded = 1;
    }
  }
  if (rc == MXR_OK) {
    f.out = out;
    hipLaunchKernelGGL(k_grid_finish, dim3(ew_blocks(2 * n_complex(N))), dim3(256), 0, c->stream, f);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
      e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess)
      rc = hip_fail(e, "subsample_grid", __FILE__, __LINE__);
  }
  (void)hipFree(staged);
  return rc;
}

int read_grid_file_info(const char* fname, mxr_grid_file_info* info, long* data_offset_of_grid0)
{
  FILE* fd = std::fopen(fname, "rb");
  if (!fd) {
    set_error("failed to open grid file %s", fname);
    return MXR_ERR_INVALID_ARG;
  }
  bool ok = std::fread(info->n_cell, sizeof(int), 3, fd) == 3 && std::fread(info->box_size, sizeof(double), 3, fd) == 3 &&
            std::fread(&info->n_grids, sizeof(int), 1, fd) == 1 && std::fread(&info->ma_scheme, sizeof(int), 1, fd) == 1;
  if (ok && data_offset_of_grid0)
    *data_offset_of_grid0 = std::ftell(fd);
  std::fclose(fd);
  if (!ok) {
    set_error("grid file %s: short header", fname);
    return MXR_ERR_INVALID_ARG;
  }
  return MXR_OK;
}

int run_read_grid_gbptrees(Ctx* c, const char* fname, int property, const mxr_grid_params* p, float* out)
{
  mxr_grid_file_info info{};
  long off0 = 0;
  RC(read_grid_file_info(fname, &info, &off0));
  const int n = info.n_cell[0];
  // read_header (read_grids-gbptrees.c:66-75) and calc_resample_factor (read_grids.c:33-57)
  if (info.n_grids != 4) {
    set_error("grid file %s: n_grids %d != 4", fname, info.n_grids);
    return MXR_ERR_INVALID_ARG;
  }
  if (info.n_cell[1] != n || info.n_cell[2] != n) {
    set_error("grid file %s: input grids are not cubic", fname);
    return MXR_ERR_INVALID_ARG;
  }
  if (n < c->N || n % c->N != 0) {
    set_error("grid file %s: n_cell %d is not a multiple of ReionGridDim %d (at least as large)", fname, n, c->N);
    return MXR_ERR_INVALID_ARG;
  }
  // grids follow one another: a 32-byte identifier, then n^3 floats
  const long nreal = (long)n * n * n;
  const long off = off0 + (long)property * (32 + (long)sizeof(float) * nreal) + 32;
  float* host = nullptr;
  MXR_HIP(hipHostMalloc(&host, sizeof(float) * nreal, hipHostMallocDefault));
  int rc = MXR_OK;
  FILE* fd = std::fopen(fname, "rb");
  if (!fd || std::fseek(fd, off, SEEK_SET) != 0 || (long)std::fread(host, sizeof(float), nreal, fd) != nreal) {
    set_error("grid file %s: failed to read grid %d (%ld floats at offset %ld)", fname, property, nreal, off);
    rc = MXR_ERR_INVALID_ARG;
  }
  if (fd)
    std::fclose(fd);
  if (rc == MXR_OK) {
    mxr_grid_params q = *p;
    q.TreesID = MXR_TREES_GBPTREES;
    rc = run_prepare_grid(c, host, n, MXR_MEM_HOST, property, info.box_size[0], &q, out);
  }
  (void)hipHostFree(host);
  return rc;
}

}  // namespace mxr
