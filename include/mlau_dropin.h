/*
 * mlau_dropin.h -- the reference's per-step entry points, same names, same
 * C++ linkage and argument lists, implemented on the GPU by libmlau_dropin
 * (which links libmlau_hip).  A reference driver (main.cpp) links against
 * libmlau_dropin instead of the solver object and runs unchanged.
 *
 * As in the reference, the boundary flags come from the driver's globals
 * dnxs/dnys/dnzs (2D/SERIAL/OTvortex/global.hpp:15-20); the library carries
 * weak default definitions so it also loads without a driver.  The scheme the
 * reference fixes with #defines (ODR, R_K, CTW, func_flux) defaults to the
 * reference files' own values and can be changed with
 * mlau_dropin_set_scheme().  Errors abort with a message, since the reference
 * entry points return void.
 */
#ifndef MLAU_DROPIN_H
#define MLAU_DROPIN_H
#include "mlau.h"

/* replaces 1D/SERIAL/shock/mhd_fd4c_1d.c:60-63 (default ODR=1, R_K=2, LHLLD, open BC) */
void mhd_fd4c_1d(double *ro, double *mx, double *my, double *mz, double *en, double *by, double *bz, double bx,
                 double dt, double dx, int nx, int xoff, double gamma);
/* replaces 2D/SERIAL/OTvortex/mhd_fd_ct_2d.c:91-94 (default ODR=2, R_K=3, CTW=1, LHLLD) */
void mhd_fd_ct_2d(double *ro, double *mx, double *my, double *mz, double *en, double *bx, double *by, double *bz,
                  double dt, double dx, double dy, int nx, int ny, int xoff, int yoff, double gamma);
/* replaces 3D/MPI/OTvortex/mhd_fd_ct_3d.c:99-103 for mpi_numx=mpi_numy=mpi_numz=1 (default ODR=2, R_K=3, CTW=1, LHLLD) */
void mhd_fd_ct_3d(double *ro, double *mx, double *my, double *mz, double *en, double *bx, double *by, double *bz,
                  double dt, double dx, double dy, double dz, int nx, int ny, int nz, int xoff, int yoff, int zoff,
                  double gamma, int mpi_rank, int mpi_numx, int mpi_numy, int mpi_numz);

#ifdef __cplusplus
extern "C" {
#endif
/* Select the scheme used by the shims of dimension dim (1, 2 or 3); the
 * boundary-flag fields of the scheme are ignored (the driver globals rule). */
int mlau_dropin_set_scheme(int dim, const mlau_scheme *scheme);
#ifdef __cplusplus
}
#endif
#endif
