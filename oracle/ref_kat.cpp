// ref_kat.cpp -- C-ABI wrappers around the reference's header-only point
// kernels, compiled against the reference headers where they lie
// (include/funcs.h, interp.h, mhd_eigen.h, mhd_flux.h).  Built only by
// oracle/ref_build.sh into oracle/_ref/; used to pin the oracle.
#include "funcs.h"
#include "interp.h"
#include "mhd_eigen.h"
#include "mhd_flux.h"

// face interpolation used by the solvers' calc_flr for each spatial order
// (2D/SERIAL/OTvortex/mhd_fd_ct_2d.c:55-68)
static void lr1(const double *f, double *fl, double *fr) { *fl = f[2]; *fr = f[2]; }
static void lr2(const double *f, double *fl, double *fr) { muscl_mm_cal_flr(f[1], f[2], f[3], fl, fr); }
static void lr3(const double *f, double *fl, double *fr) { wcns3_cal_flr(f[1], f[2], f[3], fl, fr); }
static void lr4(const double *f, double *fl, double *fr) { wcns4_cal_flr(f[0], f[1], f[2], f[3], f[4], fl, fr); }

extern "C" {

void ref_muscl(int which, double f1, double f2, double f3, double *fl, double *fr)
{
  if (which == 0) muscl_mm_cal_flr(f1, f2, f3, fl, fr);
  else if (which == 1) muscl_mc_cal_flr(f1, f2, f3, fl, fr);
  else muscl_kr_cal_flr(f1, f2, f3, fl, fr);
}
void ref_wcns3(double f0, double f1, double f2, double *fl, double *fr) { wcns3_cal_flr(f0, f1, f2, fl, fr); }
void ref_wcns4(double f0, double f1, double f2, double f3, double f4, double *fl, double *fr)
{
  wcns4_cal_flr(f0, f1, f2, f3, f4, fl, fr);
}
double ref_minmod(double a, double b) { return minmod(a, b); }
double ref_minmod3(double a, double b, double c) { return minmod3(a, b, c); }

void ref_eigen(int kind, const double *p, double gamma, double *al, double *L, double *R)
{
  double a[7] = {0, 0, 0, 0, 0, 0, 0};
  if (kind == 0)
    mhd_eigen_st(p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], gamma, &a[0], &a[1], &a[2], &a[3], &a[4], &a[5],
                 &a[6], L, L + 7, L + 14, L + 21, L + 28, L + 35, L + 42, R, R + 7, R + 14, R + 21, R + 28, R + 35,
                 R + 42);
  else
    mhd_eigen_ae(p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], gamma, &a[0], &a[1], &a[2], &a[3], &a[4], &a[5],
                 &a[6], L, L + 7, L + 14, L + 21, L + 28, L + 35, L + 42, R, R + 7, R + 14, R + 21, R + 28, R + 35,
                 R + 42);
  for (int n = 0; n < 7; n++) al[n] = a[n];
}

void ref_reconst(int kind, int odr, const double *st, double bn, double gamma, double *ul, double *ur)
{
  void (*f)(const double *, double *, double *) = (odr == 1) ? lr1 : (odr == 2) ? lr2 : (odr == 3) ? lr3 : lr4;
  if (kind == 0)
    mhd_c_reconst(st, st + 5, st + 10, st + 15, st + 20, st + 25, st + 30, bn, gamma, 5, ul, ur, f);
  else
    mhd_a_reconst(st, st + 5, st + 10, st + 15, st + 20, st + 25, st + 30, bn, gamma, 5, ul, ur, f);
}

void ref_flux(int kind, const double *q, double bn, double gamma, const double *dv, double *out)
{
  double dvv[2] = {dv[0], dv[1]};
  void (*fn)(double, double, double, double, double, double, double, double, double, double, double, double, double,
             double, double, double, double *, double *, double *, double *, double *, double *, double *, double *);
  fn = (kind == 0) ? calc_flux_mlau : (kind == 1) ? calc_flux_roe : (kind == 2) ? calc_flux_hlld : calc_flux_lhlld;
  fn(q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], q[8], q[9], q[10], q[11], q[12], q[13], bn, gamma, dvv,
     &out[0], &out[1], &out[2], &out[3], &out[4], &out[5], &out[6]);
}

}  // extern "C"
