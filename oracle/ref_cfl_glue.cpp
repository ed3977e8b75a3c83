// ref_cfl_glue.cpp -- C-ABI wrapper around the reference driver's CFL check
// (2D/SERIAL/OTvortex/cflcheck.hpp:1-32) compiled from the reference headers
// where they lie (global.hpp fixes the 200x200 grid, xoff=4, gam=5/3, cfl=0.4).
// Built only by oracle/ref_build.sh into oracle/_ref/.
#include "global.hpp"
using namespace global;
#include "cflcheck.hpp"

extern "C" {
int ref_cfl2d_nx(void) { return nx; }
double ref_cfl2d(double *ro_, double *mx_, double *my_, double *mz_, double *en_, double *bx_, double *by_,
                 double *bz_, double vfloor, double dt_in)
{
  global::ro = ro_; global::mx = mx_; global::my = my_; global::mz = mz_;
  global::en = en_; global::bx = bx_; global::by = by_; global::bz = bz_;
  global::vfast = vfloor;
  double d = dt_in;
  cflcheck(dr, &d, 0);
  return d;
}
double ref_cfl2d_dr(void) { return dr; }
}
