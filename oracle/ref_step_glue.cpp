// ref_step_glue.cpp -- host-program side of a reference per-step solver build.
// The reference solvers import the boundary flag arrays from the driver
// (2D/SERIAL/OTvortex/global.hpp:15-20, 3D/MPI/OTvortex/global.hpp:20-27) and
// export C++-linkage entry points; this file plays the driver's part and adds
// C-ABI wrappers so the test suite can call one reference step through ctypes.
// Built only by oracle/ref_build.sh into oracle/_ref/.  Select the solver with
// -DREF_DIM=1|2|3.

int dnxs[8] = {0, 0, 0, 0, 0, 0, 0, 0};
int dnys[8] = {0, 0, 0, 0, 0, 0, 0, 0};
int dnzs[8] = {0, 0, 0, 0, 0, 0, 0, 0};
int stxs[8] = {0, 0, 0, 0, 0, 1, 0, 0};
int stys[8] = {0, 0, 0, 0, 0, 0, 1, 0};
int stzs[8] = {0, 0, 0, 0, 0, 0, 0, 1};

#if REF_DIM == 1
void mhd_fd4c_1d(double *ro, double *mx, double *my, double *mz, double *en, double *by, double *bz, double bx,
                 double dt, double dx, int nx, int xoff, double gamma);
#elif REF_DIM == 2
void mhd_fd_ct_2d(double *ro, double *mx, double *my, double *mz, double *en, double *bx, double *by, double *bz,
                  double dt, double dx, double dy, int nx, int ny, int xoff, int yoff, double gamma);
void bc2d(double *f, int nx, int ny, int xoff, int yoff, int stx, int dnx, int sty, int dny);
#else
void mhd_fd_ct_3d(double *ro, double *mx, double *my, double *mz, double *en, double *bx, double *by, double *bz,
                  double dt, double dx, double dy, double dz, int nx, int ny, int nz, int xoff, int yoff, int zoff,
                  double gamma, int mpi_rank, int mpi_numx, int mpi_numy, int mpi_numz);
#endif

extern "C" {

void ref_set_bc(const int *dx, const int *dy, const int *dz)
{
  for (int n = 0; n < 8; n++) {
    dnxs[n] = dx[n];
    dnys[n] = dy[n];
    dnzs[n] = dz[n];
  }
}

#if REF_DIM == 1
void ref_step_1d(double *ro, double *mx, double *my, double *mz, double *en, double *by, double *bz, double bx,
                 double dt, double dx, int nx, int xoff, double gamma)
{
  mhd_fd4c_1d(ro, mx, my, mz, en, by, bz, bx, dt, dx, nx, xoff, gamma);
}
#elif REF_DIM == 2
void ref_step_2d(double *ro, double *mx, double *my, double *mz, double *en, double *bx, double *by, double *bz,
                 double dt, double dx, double dy, int nx, int ny, int xoff, int yoff, double gamma)
{
  mhd_fd_ct_2d(ro, mx, my, mz, en, bx, by, bz, dt, dx, dy, nx, ny, xoff, yoff, gamma);
}
void ref_bc2d(double *f, int nx, int ny, int xoff, int yoff, int stx, int dnx, int sty, int dny)
{
  bc2d(f, nx, ny, xoff, yoff, stx, dnx, sty, dny);
}
#else
void ref_step_3d(double *ro, double *mx, double *my, double *mz, double *en, double *bx, double *by, double *bz,
                 double dt, double dx, double dy, double dz, int nx, int ny, int nz, int xoff, int yoff, int zoff,
                 double gamma)
{
  mhd_fd_ct_3d(ro, mx, my, mz, en, bx, by, bz, dt, dx, dy, dz, nx, ny, nz, xoff, yoff, zoff, gamma, 0, 1, 1, 1);
}
#endif

}  // extern "C"
