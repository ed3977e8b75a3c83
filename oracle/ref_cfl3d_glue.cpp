// ref_cfl3d_glue.cpp -- C-ABI wrapper around the 3D driver's CFL check
// (3D/MPI/OTvortex/cflcheck.hpp:1-39) compiled from the reference headers
// where they lie.  global.hpp fixes the grid (XMESH/MNP_X = 50 cells per axis
// plus xoff=4 ghosts, so 58^3), gam=5/3, cfl=0.4 and dr=min(dx,dy,dz).  The
// check ends with an MPI_Allreduce over MPI_COMM_WORLD, so the wrapper starts
// MPI (a one-rank singleton) on first use.  Built only by oracle/ref_build.sh
// into oracle/_ref/.
#include "global.hpp"
using namespace global;
#include "cflcheck.hpp"

extern "C" {
int ref_cfl3d_n(void) { return nx * 1000000 + ny * 1000 + nz; }
double ref_cfl3d_dr(void) { return dr; }
double ref_cfl3d(double *ro_, double *mx_, double *my_, double *mz_, double *en_, double *bx_, double *by_,
                 double *bz_, double vfloor, double dt_in)
{
  int started = 0;
  MPI_Initialized(&started);
  if (!started) MPI_Init(NULL, NULL);
  global::ro = ro_; global::mx = mx_; global::my = my_; global::mz = mz_;
  global::en = en_; global::bx = bx_; global::by = by_; global::bz = bz_;
  global::vfast = vfloor;
  double d = dt_in;
  cflcheck(&d, 0);
  return d;
}
}
