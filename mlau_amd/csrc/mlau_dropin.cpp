// mlau_dropin.cpp -- reference-named shims over the libmlau_hip C ABI.
#include "../../include/mlau_dropin.h"

#include <cstdio>
#include <cstdlib>

// Boundary flags normally defined by the reference driver (global.hpp); weak
// so that a driver's own definitions take precedence.
__attribute__((weak)) int dnxs[8] = {0, 0, 0, 0, 0, 0, 0, 0};
__attribute__((weak)) int dnys[8] = {0, 0, 0, 0, 0, 0, 0, 0};
__attribute__((weak)) int dnzs[8] = {0, 0, 0, 0, 0, 0, 0, 0};

namespace {
mlau_scheme make(int odr, int rk, int flux, int ctw, int bc1d)
{
  mlau_scheme s{};
  s.odr = odr;
  s.rk = rk;
  s.flux = flux;
  s.ctw = ctw;
  s.bc1d = bc1d;
  return s;
}
// the reference files' own #define values (mhd_fd4c_1d.c:25-26,87; mhd_fd_ct_2d.c:25-27,128; mhd_fd_ct_3d.c:25-27,138)
mlau_scheme g_s1 = make(1, 2, MLAU_FLUX_LHLLD, 0, MLAU_BC1D_OPEN);
mlau_scheme g_s2 = make(2, 3, MLAU_FLUX_LHLLD, 1, 0);
mlau_scheme g_s3 = make(2, 3, MLAU_FLUX_LHLLD, 1, 0);

void die(const char *who)
{
  std::fprintf(stderr, "%s (GPU drop-in): %s\n", who, mlau_last_error());
  std::abort();
}
void with_flags(mlau_scheme &s)
{
  for (int n = 0; n < 8; n++) {
    s.dnxs[n] = dnxs[n];
    s.dnys[n] = dnys[n];
    s.dnzs[n] = dnzs[n];
  }
}
}  // namespace

extern "C" int mlau_dropin_set_scheme(int dim, const mlau_scheme *scheme)
{
  if (!scheme || dim < 1 || dim > 3) return MLAU_ERR_ARG;
  (dim == 1 ? g_s1 : dim == 2 ? g_s2 : g_s3) = *scheme;
  return MLAU_OK;
}

void mhd_fd4c_1d(double *ro, double *mx, double *my, double *mz, double *en, double *by, double *bz, double bx,
                 double dt, double dx, int nx, int xoff, double gamma)
{
  if (mlau_fd4c_1d(ro, mx, my, mz, en, by, bz, bx, dt, dx, nx, xoff, gamma, &g_s1)) die("mhd_fd4c_1d");
}

void mhd_fd_ct_2d(double *ro, double *mx, double *my, double *mz, double *en, double *bx, double *by, double *bz,
                  double dt, double dx, double dy, int nx, int ny, int xoff, int yoff, double gamma)
{
  mlau_scheme s = g_s2;
  with_flags(s);
  if (mlau_fd_ct_2d(ro, mx, my, mz, en, bx, by, bz, dt, dx, dy, nx, ny, xoff, yoff, gamma, &s)) die("mhd_fd_ct_2d");
}

void mhd_fd_ct_3d(double *ro, double *mx, double *my, double *mz, double *en, double *bx, double *by, double *bz,
                  double dt, double dx, double dy, double dz, int nx, int ny, int nz, int xoff, int yoff, int zoff,
                  double gamma, int mpi_rank, int mpi_numx, int mpi_numy, int mpi_numz)
{
  (void)mpi_rank;
  if (mpi_numx != 1 || mpi_numy != 1 || mpi_numz != 1) {
    std::fprintf(stderr, "mhd_fd_ct_3d (GPU drop-in): one GPU, one rank only (got %dx%dx%d ranks)\n", mpi_numx,
                 mpi_numy, mpi_numz);
    std::abort();
  }
  mlau_scheme s = g_s3;
  with_flags(s);
  if (mlau_fd_ct_3d(ro, mx, my, mz, en, bx, by, bz, dt, dx, dy, dz, nx, ny, nz, xoff, yoff, zoff, gamma, &s))
    die("mhd_fd_ct_3d");
}
