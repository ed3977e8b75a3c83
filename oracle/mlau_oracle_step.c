/*
 * mlau_oracle_step.c -- per-step updates of the CPU restatement (oracle).
 *
 * TEST INFRASTRUCTURE ONLY (see mlau_oracle.h).  Restates, loop by loop and in
 * the same evaluation order (OpenMP only splits loops whose iterations write
 * disjoint elements, so the result does not depend on the thread count):
 *   1D/SERIAL/shock/mhd_fd4c_1d.c:60-271 + openbc1d_mhd.c:21-175,
 *   1D/SERIAL/wave/prdc_1d.c:18-25 (periodic variant),
 *   2D/SERIAL/OTvortex/mhd_fd_ct_2d.c:32-660 + boundary.c:18-74,
 *   3D/MPI/OTvortex/mhd_fd_ct_3d.c:38-1198 + boundary.c single-rank paths,
 *   2D/SERIAL/OTvortex/cflcheck.hpp:1-32, 3D/MPI/OTvortex/cflcheck.hpp:1-39,
 *   1D/SERIAL/shock/main.cpp:113-128.
 * The reference reuses its ul/ur face buffers as 2-wide corner scratch; this
 * restatement keeps separate buffers, which is equivalent because every
 * element read was written earlier in the same stage.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "mlau_oracle.h"

static inline double mx_(double a, double b) { return (a >= b) ? a : b; }
static inline double mn_(double a, double b) { return (a >= b) ? b : a; }

int oc_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
void oc_set_num_threads(int n)
{
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* 2D/SERIAL/OTvortex/mhd_fd_ct_2d.c:32-40: face B -> cell centre */
static inline double s_calc_bc(int odr, const double *f)
{
  if (odr <= 2) return (0.5 * (f[1] + f[2]));
  return (0.0625 * (9.0 * (f[1] + f[2]) - (f[0] + f[3])));
}
/* mhd_fd_ct_2d.c:42-57: linear interpolation to cell faces */
static inline void s_lcal_flr(int odr, const double *f, double *fl, double *fr)
{
  if (odr == 1) {
    *fl = f[2]; *fr = f[2];
  } else if (odr == 2) {
    *fl = 0.25 * (-f[1] + 4.0 * f[2] + f[3]);
    *fr = 0.25 * (-f[3] + 4.0 * f[2] + f[1]);
  } else if (odr == 3) {
    *fl = 0.125 * (-f[1] + 6.0 * f[2] + 3.0 * f[3]);
    *fr = 0.125 * (-f[3] + 6.0 * f[2] + 3.0 * f[1]);
  } else {
    *fl = 0.003125 * (9.0 * f[0] - 56.0 * f[1] + 234.0 * f[2] + 144.0 * f[3] - 11.0 * f[4]);
    *fr = 0.003125 * (9.0 * f[4] - 56.0 * f[3] + 234.0 * f[2] + 144.0 * f[1] - 11.0 * f[0]);
  }
}
/* mhd_fd_ct_2d.c:74-82: first difference */
static inline double s_calc_df(int odr, const double *f)
{
  if (odr <= 2) return (f[2] - f[1]);
  return ((27.0 * (f[2] - f[1]) - (f[3] - f[0])) / 24.0);
}

/* Reconstruct one cell from 5-point stencils given as pointers with stride. */
static void recon_cell(int odr, int kind, const double *ro, const double *vn, const double *vt,
                       const double *vu, const double *bt, const double *bu, const double *pr, long st,
                       double bn, double gamma, double *vl, double *vr)
{
  double s[35];
  int m;
  for (m = 0; m < 5; m++) {
    long o = (m - 2) * st;
    s[m] = ro[o]; s[5 + m] = vn[o]; s[10 + m] = vt[o]; s[15 + m] = vu[o];
    s[20 + m] = bt[o]; s[25 + m] = bu[o]; s[30 + m] = pr[o];
  }
  oc_reconst(kind, odr, s, bn, gamma, vl, vr);
}

static inline void flux7(int kind, double rol, double vnl, double vtl, double vul, double btl, double bul,
                         double prl, double ror, double vnr, double vtr, double vur, double btr, double bur,
                         double prr, double bn, double gamma, const double *dv, double *f)
{
  double q[14] = {rol, vnl, vtl, vul, btl, bul, prl, ror, vnr, vtr, vur, btr, bur, prr};
  oc_flux(kind, q, bn, gamma, dv, f);
}

/* ------------------------------------------------------------------ 1D -- */

/* 1D/SERIAL/shock/openbc1d_mhd.c:21-175 with istt=0, iend=1 */
static void openbc1d(double *ro, double *mx, double *my, double *mz, double *en, double *by, double *bz,
                     double bx, double dx, double dt, int nx, int xoff, double gamma)
{
  int side;
  const int bpos[2] = {xoff, nx - xoff - 1};
  const int dirc[2] = {1, -1};
  const double bx2 = bx * bx;
  const double dtdx = dt / dx;
  for (side = 0; side <= 1; side++) {
    int sgn = dirc[side], si = bpos[side], so = bpos[side] - sgn, j;
    double rob[2], vxb[2], vyb[2], vzb[2], bxb[2], byb[2], bzb[2], prb[2];
    double w[7][2], al[7], L[49], R[49], p[8];
    int idx[2] = {si, so};
    for (j = 0; j < 2; j++) {
      int ss = idx[j];
      double iro, v2, b2;
      rob[j] = ro[ss];
      iro = 1.0 / rob[j];
      vxb[j] = mx[ss] * iro;
      vyb[j] = my[ss] * iro;
      vzb[j] = mz[ss] * iro;
      bxb[j] = bx;
      byb[j] = by[ss];
      bzb[j] = bz[ss];
      v2 = vxb[j] * vxb[j] + vyb[j] * vyb[j] + vzb[j] * vzb[j];
      b2 = bxb[j] * bxb[j] + byb[j] * byb[j] + bzb[j] * bzb[j];
      prb[j] = (gamma - 1) * (en[ss] - 0.5 * (rob[j] * v2 + b2));
    }
    p[0] = 0.5 * (rob[0] + rob[1]); p[1] = 0.5 * (vxb[0] + vxb[1]); p[2] = 0.5 * (vyb[0] + vyb[1]);
    p[3] = 0.5 * (vzb[0] + vzb[1]); p[4] = 0.5 * (bxb[0] + bxb[1]); p[5] = 0.5 * (byb[0] + byb[1]);
    p[6] = 0.5 * (bzb[0] + bzb[1]); p[7] = 0.5 * (prb[0] + prb[1]);
    oc_eigen_st(p, gamma, al, L, R);
    for (j = 0; j < 2; j++) {
      int n;
      for (n = 0; n < 7; n++) {
        const double *l = L + 7 * n;
        w[n][j] = (+l[0] * rob[j] + l[1] * vxb[j] + l[2] * vyb[j] + l[3] * vzb[j] + l[4] * byb[j] +
                   l[5] * bzb[j] + l[6] * prb[j]);
      }
    }
    {
      int n, ss = so;
      double v2, b2, q[7];
      for (n = 0; n < 7; n++)
        w[n][1] = (sgn * al[n] > 0) ? w[n][1]
                                    : (1 + sgn * al[n] * dtdx) * w[n][1] - sgn * al[n] * dtdx * w[n][0];
      for (n = 0; n < 7; n++)
        q[n] = R[0 + n] * w[0][1] + R[7 + n] * w[1][1] + R[14 + n] * w[2][1] + R[21 + n] * w[3][1] +
               R[28 + n] * w[4][1] + R[35 + n] * w[5][1] + R[42 + n] * w[6][1];
      ro[ss] = q[0]; mx[ss] = q[1]; my[ss] = q[2]; mz[ss] = q[3];
      by[ss] = q[4]; bz[ss] = q[5]; en[ss] = q[6];
      v2 = mx[ss] * mx[ss] + my[ss] * my[ss] + mz[ss] * mz[ss];
      b2 = bx2 + by[ss] * by[ss] + bz[ss] * bz[ss];
      en[ss] = en[ss] / (gamma - 1.0) + 0.5 * (ro[ss] * v2 + b2);
      mx[ss] *= ro[ss];
      my[ss] *= ro[ss];
      mz[ss] *= ro[ss];
    }
    for (j = 0; j < xoff - 1; j++) {
      int ss = side * (nx - 1) + sgn * j;
      ro[ss] = ro[so]; mx[ss] = mx[so]; my[ss] = my[so]; mz[ss] = mz[so];
      by[ss] = by[so]; bz[ss] = bz[so]; en[ss] = en[so];
    }
  }
}

/* 1D/SERIAL/wave/prdc_1d.c:18-25 */
static void prdc1d(double *f, int nx, int xoff)
{
  int i;
  for (i = 0; i < xoff; i++) {
    f[nx - 1 - i] = f[2 * xoff - 1 - i];
    f[i] = f[nx - 2 * xoff + i];
  }
}

/* 1D/SERIAL/shock/mhd_fd4c_1d.c:60-271 (bc1d open) / 1D/SERIAL/wave (periodic) */
void oc_step_1d(double *ro, double *mx, double *my, double *mz, double *en, double *by, double *bz,
                double bx, double dt, double dx, int nx, int xoff, double gamma,
                const oc_scheme *sc, int bc1d)
{
  const int odr = sc->odr, R_K = sc->rk;
  const double rkf[3][2] = {{0.0, 1.0}, {0.5 + (R_K - 2) * 0.25, 0.5 - (R_K - 2) * 0.25}, {1. / 3., 2. / 3.}};
  const double dtdx = dt / dx;
  const double bx2 = bx * bx;
  double *ut = (double *)malloc(sizeof(double) * 7 * nx);
  double *ul = (double *)malloc(sizeof(double) * 7 * nx);
  double *ur = (double *)malloc(sizeof(double) * 7 * nx);
  double *fx = (double *)malloc(sizeof(double) * 7 * nx);
  double *vx = (double *)malloc(sizeof(double) * nx);
  double *vy = (double *)malloc(sizeof(double) * nx);
  double *vz = (double *)malloc(sizeof(double) * nx);
  double *pr = (double *)malloc(sizeof(double) * nx);
  int i, rk;
  for (i = 0; i < nx; i++) {
    ut[7 * i + 0] = ro[i]; ut[7 * i + 1] = mx[i]; ut[7 * i + 2] = my[i]; ut[7 * i + 3] = mz[i];
    ut[7 * i + 4] = by[i]; ut[7 * i + 5] = bz[i]; ut[7 * i + 6] = en[i];
  }
  for (rk = 0; rk < R_K; rk++) {
    for (i = 0; i < nx; i++) {
      double iro = 1.0 / ro[i], v2, b2;
      vx[i] = mx[i] * iro;
      vy[i] = my[i] * iro;
      vz[i] = mz[i] * iro;
      v2 = vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i];
      b2 = bx2 + by[i] * by[i] + bz[i] * bz[i];
      pr[i] = (gamma - 1) * (en[i] - 0.5 * (ro[i] * v2 + b2));
    }
    for (i = 2; i < nx - 2; i++) {
      double vl[7], vr[7], cs2[3], vv2[3];
      int k, n;
      for (k = 0; k < 3; k++) {
        int s = i - 1 + k;
        cs2[k] = gamma * pr[s] / ro[s];
        vv2[k] = vx[s] * vx[s] + vy[s] * vy[s] + vz[s] * vz[s];
      }
      int kind = (mx_(vv2[0] / cs2[0], mx_(vv2[1] / cs2[1], vv2[2] / cs2[2])) >= 1.0) ? 0 : 1;
      recon_cell(odr, kind, ro + i, vx + i, vy + i, vz + i, by + i, bz + i, pr + i, 1, bx, gamma, vl, vr);
      for (n = 0; n < 7; n++) {
        ul[7 * (i + 1) + n] = vl[n];
        ur[7 * i + n] = vr[n];
      }
    }
    for (i = 3; i < nx - 2; i++) {
      double dv[2] = {(vx[i] - vx[i - 1]), 0.0};
      const double *a = ul + 7 * i, *b = ur + 7 * i;
      flux7(sc->flux, a[0], a[1], a[2], a[3], a[4], a[5], a[6], b[0], b[1], b[2], b[3], b[4], b[5], b[6],
            bx, gamma, dv, fx + 7 * i);
    }
    if (rk == 0 && bc1d == OC_BC1D_OPEN) openbc1d(ro, mx, my, mz, en, by, bz, bx, dx, dt, nx, xoff, gamma);
    for (i = 4; i < nx - 4; i++) {
      double du[7];
      double *f[7] = {ro, mx, my, mz, by, bz, en};
      int n;
      for (n = 0; n < 7; n++) {
        double v[4] = {fx[7 * (i - 1) + n], fx[7 * i + n], fx[7 * (i + 1) + n], fx[7 * (i + 2) + n]};
        du[n] = s_calc_df(odr, v);
      }
      for (n = 0; n < 7; n++) f[n][i] = rkf[rk][0] * ut[7 * i + n] + rkf[rk][1] * (f[n][i] - dtdx * du[n]);
    }
    if (bc1d == OC_BC1D_PERIODIC) {
      prdc1d(ro, nx, xoff); prdc1d(mx, nx, xoff); prdc1d(my, nx, xoff); prdc1d(mz, nx, xoff);
      prdc1d(by, nx, xoff); prdc1d(bz, nx, xoff); prdc1d(en, nx, xoff);
    }
  }
  free(ut); free(ul); free(ur); free(fx); free(vx); free(vy); free(vz); free(pr);
}

/* 1D/SERIAL/shock/main.cpp:114-127 (vfloor = 1.0 there) */
double oc_vmax_1d(const double *ro, const double *mx, const double *my, const double *mz, const double *en,
                  const double *by, const double *bz, double bx, int nx, double gamma, double vfloor)
{
  double vmax = vfloor;
  int i;
  for (i = 0; i < nx; i++) {
    double iro = 1.0 / ro[i];
    double vx = mx[i] * iro, vy = my[i] * iro, vz = mz[i] * iro;
    double v2 = vx * vx + vy * vy + vz * vz;
    double b2 = bx * bx + by[i] * by[i] + bz[i] * bz[i];
    double pr = (gamma - 1) * (en[i] - 0.5 * (ro[i] * v2 + b2));
    double cf = sqrt((gamma * pr + b2) * iro);
    double vtmp = fabs(vx) + cf;
    if (vtmp > vmax) vmax = vtmp;
  }
  return vmax;
}

/* ------------------------------------------------------------------ 2D -- */

/* 2D/SERIAL/OTvortex/boundary.c:18-74 */
void oc_bc2d(double *f, int nx, int ny, int xoff, int yoff, int stx, int dnx, int sty, int dny)
{
  int i, j;
  if (dnx == 0) {
    for (j = 0; j < ny; j++)
      for (i = 0; i < xoff; i++) {
        f[nx * j + (nx - 1 - i)] = f[nx * j + (2 * xoff - 1 - i)];
        f[nx * j + i] = f[nx * j + (nx - 2 * xoff + i)];
      }
  } else {
    for (j = 0; j < ny; j++) {
      for (i = 0; i < xoff; i++) f[nx * j + i] = dnx * f[nx * j + (2 * xoff - 1 + stx) - i];
      for (i = 0; i < xoff - stx; i++) f[nx * j + (nx - 1 - i)] = dnx * f[nx * j + (nx - 2 * xoff + stx) + i];
    }
  }
  if (dny == 0) {
    for (j = 0; j < yoff; j++)
      for (i = 0; i < nx; i++) {
        f[nx * (ny - 1 - j) + i] = f[nx * (2 * yoff - 1 - j) + i];
        f[nx * j + i] = f[nx * (ny - 2 * yoff + j) + i];
      }
  } else {
    for (j = 0; j < yoff; j++)
      for (i = 0; i < nx; i++) f[nx * j + i] = dny * f[nx * (2 * yoff - 1 + sty - j) + i];
    for (j = 0; j < yoff - sty; j++)
      for (i = 0; i < nx; i++) f[nx * (ny - 1 - j) + i] = dny * f[nx * (ny - 2 * yoff + sty + j) + i];
  }
}

/* 2D/SERIAL/OTvortex/mhd_fd_ct_2d.c:91-660 */
void oc_step_2d(double *ro, double *mx, double *my, double *mz, double *en, double *bx, double *by, double *bz,
                double dt, double dx, double dy, int nx, int ny, int xoff, int yoff, double gamma,
                const oc_scheme *sc, const int *dnxs, const int *dnys, const int *stxs, const int *stys)
{
  const int odr = sc->odr, R_K = sc->rk, fk = sc->flux, ctw = sc->ctw;
  const double rkf[3][2] = {{0.0, 1.0}, {0.5 + (R_K - 2) * 0.25, 0.5 - (R_K - 2) * 0.25}, {1. / 3., 2. / 3.}};
  const long nxy = (long)nx * ny;
  const double dtdx = dt / dx, dtdy = dt / dy;
  const double eps = 1e-6;
  double *ut = (double *)malloc(sizeof(double) * 8 * nxy);
  double *ul = (double *)malloc(sizeof(double) * 8 * nxy);
  double *ur = (double *)malloc(sizeof(double) * 8 * nxy);
  double *cl = (double *)malloc(sizeof(double) * 2 * nxy);
  double *cr = (double *)malloc(sizeof(double) * 2 * nxy);
  double *ql = (double *)malloc(sizeof(double) * nxy);
  double *qr = (double *)malloc(sizeof(double) * nxy);
  double *fx = (double *)malloc(sizeof(double) * 8 * nxy);
  double *fy = (double *)malloc(sizeof(double) * 8 * nxy);
  double *vx = (double *)malloc(sizeof(double) * nxy);
  double *vy = (double *)malloc(sizeof(double) * nxy);
  double *vz = (double *)malloc(sizeof(double) * nxy);
  double *pr = (double *)malloc(sizeof(double) * nxy);
  double *cx = (double *)malloc(sizeof(double) * nxy);
  double *cy = (double *)malloc(sizeof(double) * nxy);
  double *ez = (double *)malloc(sizeof(double) * nxy);
  double *ctx = (double *)malloc(sizeof(double) * nxy);
  double *fc = (double *)malloc(sizeof(double) * nxy);
  double *m2 = (double *)malloc(sizeof(double) * nxy);
  double *dvx = (double *)malloc(sizeof(double) * nxy);
  double *dvy = (double *)malloc(sizeof(double) * nxy);
  long ss;
  int rk;
  for (ss = 0; ss < nxy; ss++) {
    ut[8 * ss + 0] = ro[ss]; ut[8 * ss + 1] = mx[ss]; ut[8 * ss + 2] = my[ss]; ut[8 * ss + 3] = mz[ss];
    ut[8 * ss + 4] = bx[ss]; ut[8 * ss + 5] = by[ss]; ut[8 * ss + 6] = bz[ss]; ut[8 * ss + 7] = en[ss];
  }
  for (rk = 0; rk < R_K; rk++) {
    const double ra = rkf[rk][0], rb = rkf[rk][1];
#pragma omp parallel
    {
      int i, j;
      /* B at cell centre (:172-194) */
#pragma omp for
      for (j = 0; j < ny; j++)
        for (i = 1; i < nx - 2; i++) cx[nx * j + i] = s_calc_bc(odr, &bx[nx * j + i - 1]);
#pragma omp for
      for (j = 1; j < ny - 2; j++)
        for (i = 0; i < nx; i++) {
          double v[4] = {by[nx * (j - 1) + i], by[nx * j + i], by[nx * (j + 1) + i], by[nx * (j + 2) + i]};
          cy[nx * j + i] = s_calc_bc(odr, v);
        }
#pragma omp single
      {
        oc_bc2d(cx, nx, ny, xoff, yoff, 0, dnxs[5], 0, dnys[5]);
        oc_bc2d(cy, nx, ny, xoff, yoff, 0, dnxs[6], 0, dnys[6]);
      }
      /* primitives (:203-223) */
#pragma omp for
      for (j = 0; j < ny; j++)
        for (i = 0; i < nx; i++) {
          long s = (long)nx * j + i;
          double iro = 1.0 / ro[s], v2, b2;
          vx[s] = mx[s] * iro;
          vy[s] = my[s] * iro;
          vz[s] = mz[s] * iro;
          v2 = (vx[s] * vx[s] + vy[s] * vy[s] + vz[s] * vz[s]);
          b2 = (cx[s] * cx[s] + cy[s] * cy[s] + bz[s] * bz[s]);
          pr[s] = (gamma - 1) * (en[s] - 0.5 * (ro[s] * v2 + b2));
          m2[s] = v2 / (gamma * pr[s] * iro);
          ez[s] = 0.0;
          ctx[s] = 0.5;
        }
      /* CT upwind weight (:224-249) */
      if (ctw) {
#pragma omp for
        for (j = 1; j < ny; j++)
          for (i = 1; i < nx; i++) {
            long a = (long)nx * (j - 1), b = (long)nx * j;
            double vxc = 0.25 * (vx[a + i - 1] + vx[a + i] + vx[b + i - 1] + vx[b + i]);
            double vyc = 0.25 * (vy[a + i - 1] + vy[a + i] + vy[b + i - 1] + vy[b + i]);
            double roc = 0.25 * (ro[a + i - 1] + ro[a + i] + ro[b + i - 1] + ro[b + i]);
            roc = 1.0 / sqrt(roc);
            double bxc = 0.5 * (bx[a + i] + bx[b + i]);
            double byc = 0.5 * (by[b + i - 1] + by[b + i]);
            vxc = fabs(vxc) + fabs(bxc * roc);
            vyc = fabs(vyc) + fabs(byc * roc);
            ctx[b + i] = (vxc + 0.5 * eps) / (vxc + vyc + eps);
          }
      }
      /* shock sensors (:251-271) */
#pragma omp for
      for (j = 1; j < ny - 1; j++)
        for (i = 1; i < nx - 1; i++) {
          long s = (long)nx * j + i;
          dvx[s] = mn_((vx[s + 1] - vx[s]), (vx[s] - vx[s - 1]));
          dvy[s] = mn_((vy[s + nx] - vy[s]), (vy[s] - vy[s - nx]));
        }
#pragma omp single
      {
        oc_bc2d(dvx, nx, ny, xoff, yoff, 0, -dnxs[1], 0, +dnys[1]);
        oc_bc2d(dvy, nx, ny, xoff, yoff, 0, +dnxs[2], 0, -dnys[2]);
      }
      /* X reconstruction (:273-332) */
#pragma omp for
      for (j = 0; j < ny; j++)
        for (i = 2; i < nx - 2; i++) {
          long s = (long)nx * j + i, sl = s + 1, sr = s;
          double vl[7], vr[7];
          int kind = (mx_(m2[s - 1], mx_(m2[s], m2[s + 1])) >= 1.0) ? 0 : 1;
          recon_cell(odr, kind, ro + s, vx + s, vy + s, vz + s, cy + s, bz + s, pr + s, 1, cx[s], gamma, vl, vr);
          ul[8 * sl + 0] = vl[0]; ul[8 * sl + 1] = vl[1]; ul[8 * sl + 2] = vl[2]; ul[8 * sl + 3] = vl[3];
          ul[8 * sl + 4] = bx[sl]; ul[8 * sl + 5] = vl[4]; ul[8 * sl + 6] = vl[5]; ul[8 * sl + 7] = vl[6];
          ur[8 * sr + 0] = vr[0]; ur[8 * sr + 1] = vr[1]; ur[8 * sr + 2] = vr[2]; ur[8 * sr + 3] = vr[3];
          ur[8 * sr + 4] = bx[sr]; ur[8 * sr + 5] = vr[4]; ur[8 * sr + 6] = vr[5]; ur[8 * sr + 7] = vr[6];
          double val[5] = {cy[s - 2] * vx[s - 2], cy[s - 1] * vx[s - 1], cy[s] * vx[s], cy[s + 1] * vx[s + 1],
                           cy[s + 2] * vx[s + 2]};
          double a0, a1, b0, b1;
          s_lcal_flr(odr, val, &a0, &b0);
          s_lcal_flr(odr, vy + s - 2, &a1, &b1);
          ql[sl] = a0 - bx[sl] * a1;
          qr[sr] = b0 - bx[sr] * b1;
        }
      /* X Riemann (:333-362) */
#pragma omp for
      for (j = 0; j < ny; j++)
        for (i = 3; i < nx - 2; i++) {
          long s = (long)nx * j + i;
          double f[7];
          double bn = 0.5 * (ul[8 * s + 4] + ur[8 * s + 4]);
          double dv[2] = {(vx[s] - vx[s - 1]), mn_(dvy[s - 1], dvy[s])};
          const double *a = ul + 8 * s, *b = ur + 8 * s;
          flux7(fk, a[0], a[1], a[2], a[3], a[5], a[6], a[7], b[0], b[1], b[2], b[3], b[5], b[6], b[7], bn,
                gamma, dv, f);
          fx[8 * s + 0] = f[0]; fx[8 * s + 1] = f[1]; fx[8 * s + 2] = f[2]; fx[8 * s + 3] = f[3];
          fx[8 * s + 4] = 0; fx[8 * s + 5] = f[4]; fx[8 * s + 6] = f[5]; fx[8 * s + 7] = f[6];
          fc[s] = 0.5 * (ql[s] + qr[s]);
          fx[8 * s + 5] -= fc[s];
        }
      /* By flux to corners along Y (:364-391) */
#pragma omp for
      for (j = 2; j < ny - 2; j++)
        for (i = 3; i < nx - 2; i++) {
          long s = (long)nx * j + i, sl = s + nx, sr = s;
          double val[5], vac[5], a0, a1, b0, b1;
          int m;
          for (m = 0; m < 5; m++) {
            val[m] = fx[8 * (s + (long)(m - 2) * nx) + 5];
            vac[m] = fc[s + (long)(m - 2) * nx];
          }
          s_lcal_flr(odr, val, &a0, &b0);
          s_lcal_flr(odr, vac, &a1, &b1);
          cl[2 * sl + 0] = a0; cr[2 * sr + 0] = b0; cl[2 * sl + 1] = a1; cr[2 * sr + 1] = b1;
        }
#pragma omp for
      for (j = 3; j < ny - 2; j++)
        for (i = 3; i < nx - 2; i++) {
          long s = (long)nx * j + i;
          ez[s] += -0.5 * ((cl[2 * s + 0] + cr[2 * s + 0]) + (1.0 - ctx[s]) * (cl[2 * s + 1] + cr[2 * s + 1]));
        }
      /* Y reconstruction (:404-463) */
#pragma omp for
      for (j = 2; j < ny - 2; j++)
        for (i = 0; i < nx; i++) {
          long s = (long)nx * j + i, sl = s + nx, sr = s;
          double vl[7], vr[7];
          int kind = (mx_(m2[s - nx], mx_(m2[s], m2[s + nx])) >= 1.0) ? 0 : 1;
          recon_cell(odr, kind, ro + s, vy + s, vz + s, vx + s, bz + s, cx + s, pr + s, nx, cy[s], gamma, vl, vr);
          ul[8 * sl + 0] = vl[0]; ul[8 * sl + 1] = vl[1]; ul[8 * sl + 2] = vl[2]; ul[8 * sl + 3] = vl[3];
          ul[8 * sl + 4] = by[sl]; ul[8 * sl + 5] = vl[4]; ul[8 * sl + 6] = vl[5]; ul[8 * sl + 7] = vl[6];
          ur[8 * sr + 0] = vr[0]; ur[8 * sr + 1] = vr[1]; ur[8 * sr + 2] = vr[2]; ur[8 * sr + 3] = vr[3];
          ur[8 * sr + 4] = by[sr]; ur[8 * sr + 5] = vr[4]; ur[8 * sr + 6] = vr[5]; ur[8 * sr + 7] = vr[6];
          double val[5], vxs[5], a0, a1, b0, b1;
          int m;
          for (m = 0; m < 5; m++) {
            long o = s + (long)(m - 2) * nx;
            val[m] = cx[o] * vy[o];
            vxs[m] = vx[o];
          }
          s_lcal_flr(odr, val, &a0, &b0);
          s_lcal_flr(odr, vxs, &a1, &b1);
          ql[sl] = a0 - by[sl] * a1;
          qr[sr] = b0 - by[sr] * b1;
        }
      /* Y Riemann (:464-493) */
#pragma omp for
      for (j = 3; j < ny - 2; j++)
        for (i = 0; i < nx; i++) {
          long s = (long)nx * j + i;
          double f[7];
          double bn = 0.5 * (ul[8 * s + 4] + ur[8 * s + 4]);
          double dv[2] = {(vy[s] - vy[s - nx]), mn_(dvx[s - nx], dvx[s])};
          const double *a = ul + 8 * s, *b = ur + 8 * s;
          flux7(fk, a[0], a[1], a[2], a[3], a[5], a[6], a[7], b[0], b[1], b[2], b[3], b[5], b[6], b[7], bn,
                gamma, dv, f);
          fy[8 * s + 0] = f[0]; fy[8 * s + 2] = f[1]; fy[8 * s + 3] = f[2]; fy[8 * s + 1] = f[3];
          fy[8 * s + 5] = 0; fy[8 * s + 6] = f[4]; fy[8 * s + 4] = f[5]; fy[8 * s + 7] = f[6];
          fc[s] = 0.5 * (ql[s] + qr[s]);
          fy[8 * s + 4] -= fc[s];
        }
      /* Bx flux to corners along X (:495-533) */
#pragma omp for
      for (j = 3; j < ny - 2; j++)
        for (i = 2; i < nx - 2; i++) {
          long s = (long)nx * j + i, sl = s + 1, sr = s;
          double val[5], vac[5], a0, a1, b0, b1;
          int m;
          for (m = 0; m < 5; m++) {
            val[m] = fy[8 * (s + m - 2) + 4];
            vac[m] = fc[s + m - 2];
          }
          s_lcal_flr(odr, val, &a0, &b0);
          s_lcal_flr(odr, vac, &a1, &b1);
          cl[2 * sl + 0] = a0; cr[2 * sr + 0] = b0; cl[2 * sl + 1] = a1; cr[2 * sr + 1] = b1;
        }
#pragma omp for
      for (j = 3; j < ny - 2; j++)
        for (i = 3; i < nx - 2; i++) {
          long s = (long)nx * j + i;
          ez[s] += +0.5 * ((cl[2 * s + 0] + cr[2 * s + 0]) + ctx[s] * (cl[2 * s + 1] + cr[2 * s + 1]));
        }
      /* cell-centre update + RK combine (:535-592) */
#pragma omp for
      for (j = yoff; j < ny - yoff; j++)
        for (i = xoff; i < nx - xoff; i++) {
          long s = (long)nx * j + i;
          double du0[8], du1[8];
          static const int vs[6] = {0, 1, 2, 3, 6, 7};
          int q;
          for (q = 0; q < 6; q++) {
            int n = vs[q];
            double a[4] = {fx[8 * (s - 1) + n], fx[8 * s + n], fx[8 * (s + 1) + n], fx[8 * (s + 2) + n]};
            double b[4] = {fy[8 * (s - nx) + n], fy[8 * s + n], fy[8 * (s + nx) + n], fy[8 * (s + 2 * nx) + n]};
            du0[n] = s_calc_df(odr, a);
            du1[n] = s_calc_df(odr, b);
          }
          ro[s] = ra * ut[8 * s + 0] + rb * (ro[s] - dtdx * du0[0] - dtdy * du1[0]);
          mx[s] = ra * ut[8 * s + 1] + rb * (mx[s] - dtdx * du0[1] - dtdy * du1[1]);
          my[s] = ra * ut[8 * s + 2] + rb * (my[s] - dtdx * du0[2] - dtdy * du1[2]);
          mz[s] = ra * ut[8 * s + 3] + rb * (mz[s] - dtdx * du0[3] - dtdy * du1[3]);
          bz[s] = ra * ut[8 * s + 6] + rb * (bz[s] - dtdx * du0[6] - dtdy * du1[6]);
          en[s] = ra * ut[8 * s + 7] + rb * (en[s] - dtdx * du0[7] - dtdy * du1[7]);
        }
      /* CT Bx (:593-612) */
#pragma omp for
      for (j = yoff; j < ny - yoff; j++)
        for (i = xoff; i < nx - xoff + 1; i++) {
          long s = (long)nx * j + i;
          double v[4] = {ez[s - nx], ez[s], ez[s + nx], ez[s + 2 * nx]};
          double du = s_calc_df(odr, v);
          bx[s] = ra * ut[8 * s + 4] + rb * (bx[s] - dtdy * du);
        }
      /* CT By (:613-632) */
#pragma omp for
      for (j = yoff; j < ny - yoff + 1; j++)
        for (i = xoff; i < nx - xoff; i++) {
          long s = (long)nx * j + i;
          double du = s_calc_df(odr, &ez[s - 1]);
          by[s] = ra * ut[8 * s + 5] + rb * (by[s] + dtdx * du);
        }
    }
    {
      double *p[8] = {ro, mx, my, mz, en, bx, by, bz};
      int n;
      for (n = 0; n < 8; n++) oc_bc2d(p[n], nx, ny, xoff, yoff, stxs[n], dnxs[n], stys[n], dnys[n]);
    }
  }
  free(ut); free(ul); free(ur); free(cl); free(cr); free(ql); free(qr); free(fx); free(fy);
  free(vx); free(vy); free(vz); free(pr); free(cx); free(cy); free(ez); free(ctx); free(fc);
  free(m2); free(dvx); free(dvy);
}

/* 2D/SERIAL/OTvortex/cflcheck.hpp:1-32 with fmode from myfunc.h:8-13 */
double oc_cfl_vmax_2d(const double *ro, const double *mx, const double *my, const double *mz, const double *en,
                      const double *bx, const double *by, const double *bz,
                      int nx, int ny, int xoff, int yoff, double gamma, double vfloor)
{
  double vmax = vfloor;
  int i, j;
  for (j = yoff; j < ny - yoff; j++)
    for (i = xoff; i < nx - xoff; i++) {
      long s = (long)nx * j + i;
      double iro = 1.0 / ro[s];
      double vx = mx[s] * iro, vy = my[s] * iro, vz = mz[s] * iro;
      double bxc = 0.5 * (bx[s] + bx[s + 1]);
      double byc = 0.5 * (by[s] + by[s + nx]);
      double bzc = bz[s];
      double v2 = vx * vx + vy * vy + vz * vz;
      double b2 = bxc * bxc + byc * byc + bzc * bzc;
      double pr = (gamma - 1) * (en[s] - 0.5 * (ro[s] * v2 + b2));
      double cf = sqrt((0.5 * gamma * (2.0 * pr) + b2) / (ro[s]));
      double vtmp = sqrt(vx * vx + vy * vy) + cf;
      if (vtmp > vmax) vmax = vtmp;
    }
  return vmax;
}

/* ------------------------------------------------------------------ 3D -- */

#define I3(i, j, k) ((long)nx * ((long)ny * (k) + (j)) + (i))

/* 3D/MPI/OTvortex/boundary.c:91-104,150-163,209-222 (single-rank periodic copies) */
static void per3d(double **f, int nn, int nx, int ny, int nz, int xoff, int yoff, int zoff, int dnx, int dny,
                  int dnz)
{
  int n, i, j, k;
  if (dnx == 0)
    for (n = 0; n < nn; n++)
      for (k = 0; k < nz; k++)
        for (j = 0; j < ny; j++)
          for (i = 0; i < xoff; i++) {
            f[n][I3(nx - 1 - i, j, k)] = f[n][I3(2 * xoff - 1 - i, j, k)];
            f[n][I3(i, j, k)] = f[n][I3(nx - 2 * xoff + i, j, k)];
          }
  if (dny == 0)
    for (n = 0; n < nn; n++)
      for (k = 0; k < nz; k++)
        for (j = 0; j < yoff; j++)
          for (i = 0; i < nx; i++) {
            f[n][I3(i, ny - 1 - j, k)] = f[n][I3(i, 2 * yoff - 1 - j, k)];
            f[n][I3(i, j, k)] = f[n][I3(i, ny - 2 * yoff + j, k)];
          }
  if (dnz == 0)
    for (n = 0; n < nn; n++)
      for (k = 0; k < zoff; k++)
        for (j = 0; j < ny; j++)
          for (i = 0; i < nx; i++) {
            f[n][I3(i, j, nz - 1 - k)] = f[n][I3(i, j, 2 * zoff - 1 - k)];
            f[n][I3(i, j, k)] = f[n][I3(i, j, nz - 2 * zoff + k)];
          }
}
/* 3D/MPI/OTvortex/boundary.c:227-309 (single rank: both edges are physical) */
static void xbc3d(double *f, int nx, int ny, int nz, int xoff, int st, int dn)
{
  int i, j, k;
  if (dn == 0) return;
  for (k = 0; k < nz; k++)
    for (j = 0; j < ny; j++)
      for (i = 0; i < xoff; i++) f[I3(i, j, k)] = dn * f[I3((2 * xoff - 1 + st) - i, j, k)];
  for (k = 0; k < nz; k++)
    for (j = 0; j < ny; j++)
      for (i = 0; i < xoff - st; i++) f[I3(nx - 1 - i, j, k)] = dn * f[I3((nx - 2 * xoff + st) + i, j, k)];
}
static void ybc3d(double *f, int nx, int ny, int nz, int yoff, int st, int dn)
{
  int i, j, k;
  if (dn == 0) return;
  for (k = 0; k < nz; k++)
    for (j = 0; j < yoff; j++)
      for (i = 0; i < nx; i++) f[I3(i, j, k)] = dn * f[I3(i, (2 * yoff - 1 + st) - j, k)];
  for (k = 0; k < nz; k++)
    for (j = 0; j < yoff - st; j++)
      for (i = 0; i < nx; i++) f[I3(i, ny - 1 - j, k)] = dn * f[I3(i, (ny - 2 * yoff + st) + j, k)];
}
static void zbc3d(double *f, int nx, int ny, int nz, int zoff, int st, int dn)
{
  int i, j, k;
  if (dn == 0) return;
  for (k = 0; k < zoff; k++)
    for (j = 0; j < ny; j++)
      for (i = 0; i < nx; i++) f[I3(i, j, k)] = dn * f[I3(i, j, (2 * zoff - 1 + st) - k)];
  for (k = 0; k < zoff - st; k++)
    for (j = 0; j < ny; j++)
      for (i = 0; i < nx; i++) f[I3(i, j, nz - 1 - k)] = dn * f[I3(i, j, (nz - 2 * zoff + st) + k)];
}

/* 3D/MPI/OTvortex/mhd_fd_ct_3d.c:1161-1175 */
void oc_bc3d_state(double **f, int nn, int nx, int ny, int nz, int xoff, int yoff, int zoff,
                   const int *dnxs, const int *dnys, const int *dnzs,
                   const int *stxs, const int *stys, const int *stzs)
{
  int n;
  per3d(f, nn, nx, ny, nz, xoff, yoff, zoff, dnxs[0], dnys[0], dnzs[0]);
  for (n = 0; n < nn; n++) {
    xbc3d(f[n], nx, ny, nz, xoff, stxs[n], dnxs[n]);
    ybc3d(f[n], nx, ny, nz, yoff, stys[n], dnys[n]);
    zbc3d(f[n], nx, ny, nz, zoff, stzs[n], dnzs[n]);
  }
}

/* 3D/MPI/OTvortex/mhd_fd_ct_3d.c:99-1198 (single rank) */
void oc_step_3d(double *ro, double *mx, double *my, double *mz, double *en, double *bx, double *by, double *bz,
                double dt, double dx, double dy, double dz, int nx, int ny, int nz, int xoff, int yoff, int zoff,
                double gamma, const oc_scheme *sc, const int *dnxs, const int *dnys, const int *dnzs,
                const int *stxs, const int *stys, const int *stzs)
{
  const int odr = sc->odr, R_K = sc->rk, fk = sc->flux, ctw = sc->ctw;
  const double rkf[3][2] = {{0.0, 1.0}, {0.5 + (R_K - 2) * 0.25, 0.5 - (R_K - 2) * 0.25}, {1. / 3., 2. / 3.}};
  const long nxyz = (long)nx * ny * nz;
  const double dtdx = dt / dx, dtdy = dt / dy, dtdz = dt / dz;
  const double eps = 1e-6;
  const long sx = 1, sy = nx, sz = (long)nx * ny;
  double *ut = (double *)malloc(sizeof(double) * 8 * nxyz);
  double *fx = (double *)malloc(sizeof(double) * 8 * nxyz);
  double *fy = (double *)malloc(sizeof(double) * 8 * nxyz);
  double *fz = (double *)malloc(sizeof(double) * 8 * nxyz);
  double *vx = (double *)malloc(sizeof(double) * nxyz);
  double *vy = (double *)malloc(sizeof(double) * nxyz);
  double *vz = (double *)malloc(sizeof(double) * nxyz);
  double *pr = (double *)malloc(sizeof(double) * nxyz);
  double *cx = (double *)malloc(sizeof(double) * nxyz);
  double *cy = (double *)malloc(sizeof(double) * nxyz);
  double *cz = (double *)malloc(sizeof(double) * nxyz);
  double *ex = (double *)malloc(sizeof(double) * nxyz);
  double *ey = (double *)malloc(sizeof(double) * nxyz);
  double *ez = (double *)malloc(sizeof(double) * nxyz);
  double *ct = (double *)malloc(sizeof(double) * 3 * nxyz);
  double *fc = (double *)malloc(sizeof(double) * 2 * nxyz);
  double *m2 = (double *)malloc(sizeof(double) * nxyz);
  double *dvx = (double *)malloc(sizeof(double) * nxyz);
  double *dvy = (double *)malloc(sizeof(double) * nxyz);
  double *dvz = (double *)malloc(sizeof(double) * nxyz);
  /* per-sweep face buffers (the reference mallocs one line per thread) */
  double *ul = (double *)malloc(sizeof(double) * 8 * nxyz);
  double *ur = (double *)malloc(sizeof(double) * 8 * nxyz);
  double *ql = (double *)malloc(sizeof(double) * 2 * nxyz);
  double *qr = (double *)malloc(sizeof(double) * 2 * nxyz);
  double *cl = (double *)malloc(sizeof(double) * 2 * nxyz);
  double *cr = (double *)malloc(sizeof(double) * 2 * nxyz);
  long s;
  int rk;
  for (s = 0; s < nxyz; s++) {
    ut[8 * s + 0] = ro[s]; ut[8 * s + 1] = mx[s]; ut[8 * s + 2] = my[s]; ut[8 * s + 3] = mz[s];
    ut[8 * s + 4] = bx[s]; ut[8 * s + 5] = by[s]; ut[8 * s + 6] = bz[s]; ut[8 * s + 7] = en[s];
  }
  for (rk = 0; rk < R_K; rk++) {
    const double ra = rkf[rk][0], rb = rkf[rk][1];
    double *ct0 = ct, *ct1 = ct + nxyz, *ct2 = ct + 2 * nxyz;
#pragma omp parallel
    {
      int i, j, k;
      /* B at cell centre (:186-220) */
#pragma omp for
      for (k = 0; k < nz; k++) {
        for (j = 0; j < ny; j++)
          for (i = 1; i < nx - 2; i++) cx[I3(i, j, k)] = s_calc_bc(odr, &bx[I3(i - 1, j, k)]);
        for (j = 1; j < ny - 2; j++)
          for (i = 0; i < nx; i++) {
            double v[4] = {by[I3(i, j - 1, k)], by[I3(i, j, k)], by[I3(i, j + 1, k)], by[I3(i, j + 2, k)]};
            cy[I3(i, j, k)] = s_calc_bc(odr, v);
          }
      }
#pragma omp for
      for (k = 1; k < nz - 2; k++)
        for (j = 0; j < ny; j++)
          for (i = 0; i < nx; i++) {
            double v[4] = {bz[I3(i, j, k - 1)], bz[I3(i, j, k)], bz[I3(i, j, k + 1)], bz[I3(i, j, k + 2)]};
            cz[I3(i, j, k)] = s_calc_bc(odr, v);
          }
#pragma omp single
      {
        double *p[3] = {cx, cy, cz};
        per3d(p, 3, nx, ny, nz, xoff, yoff, zoff, dnxs[0], dnys[0], dnzs[0]);
        xbc3d(cx, nx, ny, nz, xoff, 0, dnxs[5]); xbc3d(cy, nx, ny, nz, xoff, 0, dnxs[6]);
        xbc3d(cz, nx, ny, nz, xoff, 0, dnxs[7]);
        ybc3d(cx, nx, ny, nz, yoff, 0, dnys[5]); ybc3d(cy, nx, ny, nz, yoff, 0, dnys[6]);
        ybc3d(cz, nx, ny, nz, yoff, 0, dnys[7]);
        zbc3d(cx, nx, ny, nz, zoff, 0, dnzs[5]); zbc3d(cy, nx, ny, nz, zoff, 0, dnzs[6]);
        zbc3d(cz, nx, ny, nz, zoff, 0, dnzs[7]);
      }
      /* primitives (:240-262) */
#pragma omp for
      for (k = 0; k < nz; k++)
        for (j = 0; j < ny; j++)
          for (i = 0; i < nx; i++) {
            long q = I3(i, j, k);
            double iro = 1.0 / ro[q], v2, b2;
            vx[q] = mx[q] * iro;
            vy[q] = my[q] * iro;
            vz[q] = mz[q] * iro;
            v2 = (vx[q] * vx[q] + vy[q] * vy[q] + vz[q] * vz[q]);
            b2 = (cx[q] * cx[q] + cy[q] * cy[q] + cz[q] * cz[q]);
            pr[q] = (gamma - 1) * (en[q] - 0.5 * (ro[q] * v2 + b2));
            m2[q] = v2 / (gamma * pr[q] * iro);
            ex[q] = ey[q] = ez[q] = 0.0;
            ct0[q] = ct1[q] = ct2[q] = 0.5;
          }
      /* CT upwind weights (:264-340) */
      if (ctw) {
#pragma omp for
        for (k = 0; k < nz; k++)
          for (j = 1; j < ny; j++)
            for (i = 1; i < nx; i++) {
              long q = I3(i, j, k), a = q - sy - sx, b = q - sy, c = q - sx;
              double vxc = 0.25 * (vx[a] + vx[b] + vx[c] + vx[q]);
              double vyc = 0.25 * (vy[a] + vy[b] + vy[c] + vy[q]);
              double roc = 0.25 * (ro[a] + ro[b] + ro[c] + ro[q]);
              roc = 1.0 / sqrt(roc);
              double bxc = 0.5 * (bx[q - sy] + bx[q]);
              double byc = 0.5 * (by[q - sx] + by[q]);
              vxc = fabs(vxc) + fabs(bxc * roc);
              vyc = fabs(vyc) + fabs(byc * roc);
              ct0[q] = (vxc + 0.5 * eps) / (vxc + vyc + eps);
            }
#pragma omp for
        for (k = 1; k < nz; k++) {
          for (j = 1; j < ny; j++)
            for (i = 0; i < nx; i++) {
              long q = I3(i, j, k), a = q - sz - sy, b = q - sz, c = q - sy;
              double vyc = 0.25 * (vy[a] + vy[b] + vy[c] + vy[q]);
              double vzc = 0.25 * (vz[a] + vz[b] + vz[c] + vz[q]);
              double roc = 0.25 * (ro[a] + ro[b] + ro[c] + ro[q]);
              roc = 1.0 / sqrt(roc);
              double byc = 0.5 * (by[q - sz] + by[q]);
              double bzc = 0.5 * (bz[q - sy] + bz[q]);
              vyc = fabs(vyc) + fabs(byc * roc);
              vzc = fabs(vzc) + fabs(bzc * roc);
              ct1[q] = (vyc + 0.5 * eps) / (vyc + vzc + eps);
            }
          for (j = 0; j < ny; j++)
            for (i = 1; i < nx; i++) {
              long q = I3(i, j, k), a = q - sz - sx, b = q - sz, c = q - sx;
              double vzc = 0.25 * (vz[a] + vz[b] + vz[c] + vz[q]);
              double vxc = 0.25 * (vx[a] + vx[b] + vx[c] + vx[q]);
              double roc = 0.25 * (ro[a] + ro[b] + ro[c] + ro[q]);
              roc = 1.0 / sqrt(roc);
              double bzc = 0.5 * (bz[q - sx] + bz[q]);
              double bxc = 0.5 * (bx[q - sz] + bx[q]);
              vzc = fabs(vzc) + fabs(bzc * roc);
              vxc = fabs(vxc) + fabs(bxc * roc);
              ct2[q] = (vzc + 0.5 * eps) / (vzc + vxc + eps);
            }
        }
      }
      /* shock sensors (:342-377) */
#pragma omp for
      for (k = 1; k < nz - 1; k++)
        for (j = 1; j < ny - 1; j++)
          for (i = 1; i < nx - 1; i++) {
            long q = I3(i, j, k);
            dvx[q] = mn_((vx[q + sx] - vx[q]), (vx[q] - vx[q - sx]));
            dvy[q] = mn_((vy[q + sy] - vy[q]), (vy[q] - vy[q - sy]));
            dvz[q] = mn_((vz[q + sz] - vz[q]), (vz[q] - vz[q - sz]));
          }
#pragma omp single
      {
        double *p[3] = {dvx, dvy, dvz};
        per3d(p, 3, nx, ny, nz, xoff, yoff, zoff, dnxs[0], dnys[0], dnzs[0]);
        xbc3d(dvx, nx, ny, nz, xoff, 0, -dnxs[1]); xbc3d(dvy, nx, ny, nz, xoff, 0, +dnxs[2]);
        xbc3d(dvz, nx, ny, nz, xoff, 0, +dnxs[3]);
        ybc3d(dvx, nx, ny, nz, yoff, 0, +dnys[1]); ybc3d(dvy, nx, ny, nz, yoff, 0, -dnys[2]);
        ybc3d(dvz, nx, ny, nz, yoff, 0, +dnys[3]);
        zbc3d(dvx, nx, ny, nz, zoff, 0, +dnzs[1]); zbc3d(dvy, nx, ny, nz, zoff, 0, +dnzs[2]);
        zbc3d(dvz, nx, ny, nz, zoff, 0, -dnzs[3]);
      }

      /* ---- X sweep (:379-493) ---- */
#pragma omp for
      for (k = 0; k < nz; k++)
        for (j = 0; j < ny; j++) {
          for (i = 2; i < nx - 2; i++) {
            long q = I3(i, j, k), ql_ = q + sx, qr_ = q;
            double vl[7], vr[7], v5[5], a0, a1, b0, b1;
            int m, kind = (mx_(m2[q - sx], mx_(m2[q], m2[q + sx])) >= 1.0) ? 0 : 1;
            recon_cell(odr, kind, ro + q, vx + q, vy + q, vz + q, cy + q, cz + q, pr + q, sx, cx[q], gamma, vl, vr);
            ul[8 * ql_ + 0] = vl[0]; ul[8 * ql_ + 1] = vl[1]; ul[8 * ql_ + 2] = vl[2]; ul[8 * ql_ + 3] = vl[3];
            ul[8 * ql_ + 4] = bx[ql_]; ul[8 * ql_ + 5] = vl[4]; ul[8 * ql_ + 6] = vl[5]; ul[8 * ql_ + 7] = vl[6];
            ur[8 * qr_ + 0] = vr[0]; ur[8 * qr_ + 1] = vr[1]; ur[8 * qr_ + 2] = vr[2]; ur[8 * qr_ + 3] = vr[3];
            ur[8 * qr_ + 4] = bx[qr_]; ur[8 * qr_ + 5] = vr[4]; ur[8 * qr_ + 6] = vr[5]; ur[8 * qr_ + 7] = vr[6];
            for (m = 0; m < 5; m++) v5[m] = cy[q + (m - 2) * sx] * vx[q + (m - 2) * sx];
            s_lcal_flr(odr, v5, &a0, &b0);
            s_lcal_flr(odr, vy + q - 2 * sx, &a1, &b1);
            ql[2 * ql_ + 0] = a0 - bx[ql_] * a1;
            qr[2 * qr_ + 0] = b0 - bx[qr_] * b1;
            for (m = 0; m < 5; m++) v5[m] = cz[q + (m - 2) * sx] * vx[q + (m - 2) * sx];
            s_lcal_flr(odr, v5, &a0, &b0);
            s_lcal_flr(odr, vz + q - 2 * sx, &a1, &b1);
            ql[2 * ql_ + 1] = a0 - bx[ql_] * a1;
            qr[2 * qr_ + 1] = b0 - bx[qr_] * b1;
          }
          for (i = 3; i < nx - 2; i++) {
            long q = I3(i, j, k);
            double f[7];
            double bn = 0.5 * (ul[8 * q + 4] + ur[8 * q + 4]);
            double dvtm = mn_(mn_(dvy[q - sx], dvy[q]), mn_(dvz[q - sx], dvz[q]));
            double dv[2] = {(vx[q] - vx[q - sx]), dvtm};
            const double *a = ul + 8 * q, *b = ur + 8 * q;
            flux7(fk, a[0], a[1], a[2], a[3], a[5], a[6], a[7], b[0], b[1], b[2], b[3], b[5], b[6], b[7], bn,
                  gamma, dv, f);
            fx[8 * q + 0] = f[0]; fx[8 * q + 1] = f[1]; fx[8 * q + 2] = f[2]; fx[8 * q + 3] = f[3];
            fx[8 * q + 4] = 0; fx[8 * q + 5] = f[4]; fx[8 * q + 6] = f[5]; fx[8 * q + 7] = f[6];
            fc[2 * q + 0] = 0.5 * (ql[2 * q + 0] + qr[2 * q + 0]);
            fx[8 * q + 5] -= fc[2 * q + 0];
            fc[2 * q + 1] = 0.5 * (ql[2 * q + 1] + qr[2 * q + 1]);
            fx[8 * q + 6] -= fc[2 * q + 1];
          }
        }
      /* By flux to corners along Y -> Ez (:494-540) */
#pragma omp for
      for (k = 0; k < nz; k++)
        for (i = 3; i < nx - 2; i++) {
          for (j = 2; j < ny - 2; j++) {
            long q = I3(i, j, k);
            double val[5], vac[5];
            int m;
            for (m = 0; m < 5; m++) {
              val[m] = fx[8 * (q + (m - 2) * sy) + 5];
              vac[m] = fc[2 * (q + (m - 2) * sy) + 0];
            }
            s_lcal_flr(odr, val, &cl[2 * (q + sy) + 0], &cr[2 * q + 0]);
            s_lcal_flr(odr, vac, &cl[2 * (q + sy) + 1], &cr[2 * q + 1]);
          }
          for (j = 3; j < ny - 2; j++) {
            long q = I3(i, j, k);
            ez[q] += -0.5 * ((cl[2 * q + 0] + cr[2 * q + 0]) + (1.0 - ct0[q]) * (cl[2 * q + 1] + cr[2 * q + 1]));
          }
        }
      /* Bz flux to corners along Z -> Ey (:541-585) */
#pragma omp for
      for (j = 0; j < ny; j++)
        for (i = 3; i < nx - 2; i++) {
          for (k = 2; k < nz - 2; k++) {
            long q = I3(i, j, k);
            double val[5], vac[5];
            int m;
            for (m = 0; m < 5; m++) {
              val[m] = fx[8 * (q + (m - 2) * sz) + 6];
              vac[m] = fc[2 * (q + (m - 2) * sz) + 1];
            }
            /* separate corner buffers per axis keep this loop independent of the Ez loop */
            s_lcal_flr(odr, val, &ul[8 * (q + sz) + 0], &ur[8 * q + 0]);
            s_lcal_flr(odr, vac, &ul[8 * (q + sz) + 1], &ur[8 * q + 1]);
          }
          for (k = 3; k < nz - 2; k++) {
            long q = I3(i, j, k);
            ey[q] += +0.5 * ((ul[8 * q + 0] + ur[8 * q + 0]) + ct2[q] * (ul[8 * q + 1] + ur[8 * q + 1]));
          }
        }

      /* ---- Y sweep (:587-701) ---- */
#pragma omp for
      for (k = 0; k < nz; k++)
        for (i = 0; i < nx; i++) {
          for (j = 2; j < ny - 2; j++) {
            long q = I3(i, j, k), ql_ = q + sy, qr_ = q;
            double vl[7], vr[7], v5[5], w5[5], a0, a1, b0, b1;
            int m, kind = (mx_(m2[q - sy], mx_(m2[q], m2[q + sy])) >= 1.0) ? 0 : 1;
            recon_cell(odr, kind, ro + q, vy + q, vz + q, vx + q, cz + q, cx + q, pr + q, sy, cy[q], gamma, vl, vr);
            ul[8 * ql_ + 0] = vl[0]; ul[8 * ql_ + 1] = vl[1]; ul[8 * ql_ + 2] = vl[2]; ul[8 * ql_ + 3] = vl[3];
            ul[8 * ql_ + 4] = by[ql_]; ul[8 * ql_ + 5] = vl[4]; ul[8 * ql_ + 6] = vl[5]; ul[8 * ql_ + 7] = vl[6];
            ur[8 * qr_ + 0] = vr[0]; ur[8 * qr_ + 1] = vr[1]; ur[8 * qr_ + 2] = vr[2]; ur[8 * qr_ + 3] = vr[3];
            ur[8 * qr_ + 4] = by[qr_]; ur[8 * qr_ + 5] = vr[4]; ur[8 * qr_ + 6] = vr[5]; ur[8 * qr_ + 7] = vr[6];
            for (m = 0; m < 5; m++) { v5[m] = cz[q + (m - 2) * sy] * vy[q + (m - 2) * sy]; w5[m] = vz[q + (m - 2) * sy]; }
            s_lcal_flr(odr, v5, &a0, &b0);
            s_lcal_flr(odr, w5, &a1, &b1);
            ql[2 * ql_ + 0] = a0 - by[ql_] * a1;
            qr[2 * qr_ + 0] = b0 - by[qr_] * b1;
            for (m = 0; m < 5; m++) { v5[m] = cx[q + (m - 2) * sy] * vy[q + (m - 2) * sy]; w5[m] = vx[q + (m - 2) * sy]; }
            s_lcal_flr(odr, v5, &a0, &b0);
            s_lcal_flr(odr, w5, &a1, &b1);
            ql[2 * ql_ + 1] = a0 - by[ql_] * a1;
            qr[2 * qr_ + 1] = b0 - by[qr_] * b1;
          }
          for (j = 3; j < ny - 2; j++) {
            long q = I3(i, j, k);
            double f[7];
            double bn = 0.5 * (ul[8 * q + 4] + ur[8 * q + 4]);
            double dvtm = mn_(mn_(dvz[q - sy], dvz[q]), mn_(dvx[q - sy], dvx[q]));
            double dv[2] = {(vy[q] - vy[q - sy]), dvtm};
            const double *a = ul + 8 * q, *b = ur + 8 * q;
            flux7(fk, a[0], a[1], a[2], a[3], a[5], a[6], a[7], b[0], b[1], b[2], b[3], b[5], b[6], b[7], bn,
                  gamma, dv, f);
            fy[8 * q + 0] = f[0]; fy[8 * q + 2] = f[1]; fy[8 * q + 3] = f[2]; fy[8 * q + 1] = f[3];
            fy[8 * q + 5] = 0; fy[8 * q + 6] = f[4]; fy[8 * q + 4] = f[5]; fy[8 * q + 7] = f[6];
            fc[2 * q + 0] = 0.5 * (ql[2 * q + 0] + qr[2 * q + 0]);
            fy[8 * q + 6] -= fc[2 * q + 0];
            fc[2 * q + 1] = 0.5 * (ql[2 * q + 1] + qr[2 * q + 1]);
            fy[8 * q + 4] -= fc[2 * q + 1];
          }
        }
      /* Bz flux to corners along Z -> Ex (:702-747) */
#pragma omp for
      for (j = 3; j < ny - 2; j++)
        for (i = 0; i < nx; i++) {
          for (k = 2; k < nz - 2; k++) {
            long q = I3(i, j, k);
            double val[5], vac[5];
            int m;
            for (m = 0; m < 5; m++) {
              val[m] = fy[8 * (q + (m - 2) * sz) + 6];
              vac[m] = fc[2 * (q + (m - 2) * sz) + 0];
            }
            s_lcal_flr(odr, val, &cl[2 * (q + sz) + 0], &cr[2 * q + 0]);
            s_lcal_flr(odr, vac, &cl[2 * (q + sz) + 1], &cr[2 * q + 1]);
          }
          for (k = 3; k < nz - 2; k++) {
            long q = I3(i, j, k);
            ex[q] += -0.5 * ((cl[2 * q + 0] + cr[2 * q + 0]) + (1.0 - ct1[q]) * (cl[2 * q + 1] + cr[2 * q + 1]));
          }
        }
      /* Bx flux to corners along X -> Ez (:748-793) */
#pragma omp for
      for (k = 0; k < nz; k++)
        for (j = 3; j < ny - 2; j++) {
          for (i = 2; i < nx - 2; i++) {
            long q = I3(i, j, k);
            double val[5], vac[5];
            int m;
            for (m = 0; m < 5; m++) {
              val[m] = fy[8 * (q + (m - 2) * sx) + 4];
              vac[m] = fc[2 * (q + (m - 2) * sx) + 1];
            }
            s_lcal_flr(odr, val, &ul[8 * (q + sx) + 0], &ur[8 * q + 0]);
            s_lcal_flr(odr, vac, &ul[8 * (q + sx) + 1], &ur[8 * q + 1]);
          }
          for (i = 3; i < nx - 2; i++) {
            long q = I3(i, j, k);
            ez[q] += +0.5 * ((ul[8 * q + 0] + ur[8 * q + 0]) + ct0[q] * (ul[8 * q + 1] + ur[8 * q + 1]));
          }
        }

      /* ---- Z sweep (:795-909) ---- */
#pragma omp for
      for (j = 0; j < ny; j++)
        for (i = 0; i < nx; i++) {
          for (k = 2; k < nz - 2; k++) {
            long q = I3(i, j, k), ql_ = q + sz, qr_ = q;
            double vl[7], vr[7], v5[5], w5[5], a0, a1, b0, b1;
            int m, kind = (mx_(m2[q - sz], mx_(m2[q], m2[q + sz])) >= 1.0) ? 0 : 1;
            recon_cell(odr, kind, ro + q, vz + q, vx + q, vy + q, cx + q, cy + q, pr + q, sz, cz[q], gamma, vl, vr);
            ul[8 * ql_ + 0] = vl[0]; ul[8 * ql_ + 1] = vl[1]; ul[8 * ql_ + 2] = vl[2]; ul[8 * ql_ + 3] = vl[3];
            ul[8 * ql_ + 4] = bz[ql_]; ul[8 * ql_ + 5] = vl[4]; ul[8 * ql_ + 6] = vl[5]; ul[8 * ql_ + 7] = vl[6];
            ur[8 * qr_ + 0] = vr[0]; ur[8 * qr_ + 1] = vr[1]; ur[8 * qr_ + 2] = vr[2]; ur[8 * qr_ + 3] = vr[3];
            ur[8 * qr_ + 4] = bz[qr_]; ur[8 * qr_ + 5] = vr[4]; ur[8 * qr_ + 6] = vr[5]; ur[8 * qr_ + 7] = vr[6];
            for (m = 0; m < 5; m++) { v5[m] = cx[q + (m - 2) * sz] * vz[q + (m - 2) * sz]; w5[m] = vx[q + (m - 2) * sz]; }
            s_lcal_flr(odr, v5, &a0, &b0);
            s_lcal_flr(odr, w5, &a1, &b1);
            ql[2 * ql_ + 0] = a0 - bz[ql_] * a1;
            qr[2 * qr_ + 0] = b0 - bz[qr_] * b1;
            for (m = 0; m < 5; m++) { v5[m] = cy[q + (m - 2) * sz] * vz[q + (m - 2) * sz]; w5[m] = vy[q + (m - 2) * sz]; }
            s_lcal_flr(odr, v5, &a0, &b0);
            s_lcal_flr(odr, w5, &a1, &b1);
            ql[2 * ql_ + 1] = a0 - bz[ql_] * a1;
            qr[2 * qr_ + 1] = b0 - bz[qr_] * b1;
          }
          for (k = 3; k < nz - 2; k++) {
            long q = I3(i, j, k);
            double f[7];
            double bn = 0.5 * (ul[8 * q + 4] + ur[8 * q + 4]);
            double dvtm = mn_(mn_(dvx[q - sz], dvx[q]), mn_(dvy[q - sz], dvy[q]));
            double dv[2] = {(vz[q] - vz[q - sz]), dvtm};
            const double *a = ul + 8 * q, *b = ur + 8 * q;
            flux7(fk, a[0], a[1], a[2], a[3], a[5], a[6], a[7], b[0], b[1], b[2], b[3], b[5], b[6], b[7], bn,
                  gamma, dv, f);
            fz[8 * q + 0] = f[0]; fz[8 * q + 3] = f[1]; fz[8 * q + 1] = f[2]; fz[8 * q + 2] = f[3];
            fz[8 * q + 6] = 0; fz[8 * q + 4] = f[4]; fz[8 * q + 5] = f[5]; fz[8 * q + 7] = f[6];
            fc[2 * q + 0] = 0.5 * (ql[2 * q + 0] + qr[2 * q + 0]);
            fz[8 * q + 4] -= fc[2 * q + 0];
            fc[2 * q + 1] = 0.5 * (ql[2 * q + 1] + qr[2 * q + 1]);
            fz[8 * q + 5] -= fc[2 * q + 1];
          }
        }
      /* Bx flux to corners along X -> Ey (:910-955) */
#pragma omp for
      for (k = 3; k < nz - 2; k++)
        for (j = 0; j < ny; j++) {
          for (i = 2; i < nx - 2; i++) {
            long q = I3(i, j, k);
            double val[5], vac[5];
            int m;
            for (m = 0; m < 5; m++) {
              val[m] = fz[8 * (q + (m - 2) * sx) + 4];
              vac[m] = fc[2 * (q + (m - 2) * sx) + 0];
            }
            s_lcal_flr(odr, val, &cl[2 * (q + sx) + 0], &cr[2 * q + 0]);
            s_lcal_flr(odr, vac, &cl[2 * (q + sx) + 1], &cr[2 * q + 1]);
          }
          for (i = 3; i < nx - 2; i++) {
            long q = I3(i, j, k);
            ey[q] += -0.5 * ((cl[2 * q + 0] + cr[2 * q + 0]) + (1.0 - ct2[q]) * (cl[2 * q + 1] + cr[2 * q + 1]));
          }
        }
      /* By flux to corners along Y -> Ex (:956-1001) */
#pragma omp for
      for (k = 3; k < nz - 2; k++)
        for (i = 0; i < nx; i++) {
          for (j = 2; j < ny - 2; j++) {
            long q = I3(i, j, k);
            double val[5], vac[5];
            int m;
            for (m = 0; m < 5; m++) {
              val[m] = fz[8 * (q + (m - 2) * sy) + 5];
              vac[m] = fc[2 * (q + (m - 2) * sy) + 1];
            }
            s_lcal_flr(odr, val, &ul[8 * (q + sy) + 0], &ur[8 * q + 0]);
            s_lcal_flr(odr, vac, &ul[8 * (q + sy) + 1], &ur[8 * q + 1]);
          }
          for (j = 3; j < ny - 2; j++) {
            long q = I3(i, j, k);
            ex[q] += +0.5 * ((ul[8 * q + 0] + ur[8 * q + 0]) + ct1[q] * (ul[8 * q + 1] + ur[8 * q + 1]));
          }
        }

      /* cell-centre update + RK combine (:1003-1078) */
#pragma omp for
      for (k = zoff; k < nz - zoff; k++)
        for (j = yoff; j < ny - yoff; j++)
          for (i = xoff; i < nx - xoff; i++) {
            long q = I3(i, j, k);
            static const int vs[5] = {0, 1, 2, 3, 7};
            double d0[8], d1[8], d2[8];
            int n, t;
            for (t = 0; t < 5; t++) {
              n = vs[t];
              double a[4] = {fx[8 * (q - sx) + n], fx[8 * q + n], fx[8 * (q + sx) + n], fx[8 * (q + 2 * sx) + n]};
              double b[4] = {fy[8 * (q - sy) + n], fy[8 * q + n], fy[8 * (q + sy) + n], fy[8 * (q + 2 * sy) + n]};
              double c[4] = {fz[8 * (q - sz) + n], fz[8 * q + n], fz[8 * (q + sz) + n], fz[8 * (q + 2 * sz) + n]};
              d0[n] = s_calc_df(odr, a);
              d1[n] = s_calc_df(odr, b);
              d2[n] = s_calc_df(odr, c);
            }
            ro[q] = ra * ut[8 * q + 0] + rb * (ro[q] - dtdx * d0[0] - dtdy * d1[0] - dtdz * d2[0]);
            mx[q] = ra * ut[8 * q + 1] + rb * (mx[q] - dtdx * d0[1] - dtdy * d1[1] - dtdz * d2[1]);
            my[q] = ra * ut[8 * q + 2] + rb * (my[q] - dtdx * d0[2] - dtdy * d1[2] - dtdz * d2[2]);
            mz[q] = ra * ut[8 * q + 3] + rb * (mz[q] - dtdx * d0[3] - dtdy * d1[3] - dtdz * d2[3]);
            en[q] = ra * ut[8 * q + 7] + rb * (en[q] - dtdx * d0[7] - dtdy * d1[7] - dtdz * d2[7]);
          }
      /* CT Bx, By (:1079-1128) */
#pragma omp for
      for (k = zoff; k < nz - zoff; k++) {
        for (j = yoff; j < ny - yoff; j++)
          for (i = xoff; i < nx - xoff + 1; i++) {
            long q = I3(i, j, k);
            double a[4] = {ez[q - sy], ez[q], ez[q + sy], ez[q + 2 * sy]};
            double b[4] = {ey[q - sz], ey[q], ey[q + sz], ey[q + 2 * sz]};
            double du0 = s_calc_df(odr, a), du1 = s_calc_df(odr, b);
            bx[q] = ra * ut[8 * q + 4] + rb * (bx[q] - dtdy * du0 + dtdz * du1);
          }
        for (j = yoff; j < ny - yoff + 1; j++)
          for (i = xoff; i < nx - xoff; i++) {
            long q = I3(i, j, k);
            double a[4] = {ex[q - sz], ex[q], ex[q + sz], ex[q + 2 * sz]};
            double du0 = s_calc_df(odr, a), du1 = s_calc_df(odr, &ez[q - sx]);
            by[q] = ra * ut[8 * q + 5] + rb * (by[q] - dtdz * du0 + dtdx * du1);
          }
      }
      /* CT Bz (:1129-1157) */
#pragma omp for
      for (k = zoff; k < nz - zoff + 1; k++)
        for (j = yoff; j < ny - yoff; j++)
          for (i = xoff; i < nx - xoff; i++) {
            long q = I3(i, j, k);
            double b[4] = {ex[q - sy], ex[q], ex[q + sy], ex[q + 2 * sy]};
            double du0 = s_calc_df(odr, &ey[q - sx]), du1 = s_calc_df(odr, b);
            bz[q] = ra * ut[8 * q + 6] + rb * (bz[q] - dtdx * du0 + dtdy * du1);
          }
    }
    {
      double *p[8] = {ro, mx, my, mz, en, bx, by, bz};
      oc_bc3d_state(p, 8, nx, ny, nz, xoff, yoff, zoff, dnxs, dnys, dnzs, stxs, stys, stzs);
    }
  }
  free(ut); free(fx); free(fy); free(fz); free(vx); free(vy); free(vz); free(pr); free(cx); free(cy);
  free(cz); free(ex); free(ey); free(ez); free(ct); free(fc); free(m2); free(dvx); free(dvy); free(dvz);
  free(ul); free(ur); free(ql); free(qr); free(cl); free(cr);
}

/* 3D/MPI/OTvortex/cflcheck.hpp:1-39 (single rank, so the allreduce is the identity) */
double oc_cfl_vmax_3d(const double *ro, const double *mx, const double *my, const double *mz, const double *en,
                      const double *bx, const double *by, const double *bz,
                      int nx, int ny, int nz, int xoff, int yoff, int zoff, double gamma, double vfloor)
{
  double vmax = vfloor;
  int i, j, k;
  for (k = zoff; k < nz - zoff; k++)
    for (j = yoff; j < ny - yoff; j++)
      for (i = xoff; i < nx - xoff; i++) {
        long s = I3(i, j, k);
        double iro = 1.0 / ro[s];
        double vx = mx[s] * iro, vy = my[s] * iro, vz = mz[s] * iro;
        double bxc = 0.5 * (bx[s] + bx[s + 1]);
        double byc = 0.5 * (by[s] + by[s + nx]);
        double bzc = 0.5 * (bz[s] + bz[s + (long)nx * ny]);
        double v2 = vx * vx + vy * vy + vz * vz;
        double b2 = bxc * bxc + byc * byc + bzc * bzc;
        double pr = (gamma - 1) * (en[s] - 0.5 * (ro[s] * v2 + b2));
        double cf = sqrt((0.5 * gamma * (2.0 * pr) + b2) / (ro[s]));
        double vtmp = sqrt(v2) + cf;
        if (vtmp > vmax) vmax = vtmp;
      }
  return vmax;
}
