// mlau_3d_body.h -- device bodies of the 3D ideal-MHD + CT stage kernels,
// shared by the single-grid kernels (mlau_3d.hip) and the batched ensemble
// kernels (mlau_ens3d.hip).  Reference: 3D/MPI/OTvortex/mhd_fd_ct_3d.c:178-1176.
#ifndef MLAU_3D_BODY_H
#define MLAU_3D_BODY_H
#include "mlau_dev_mixed.h"
#include "mlau_internal.h"

namespace mlau {

struct U3 {
  double *f[8];  // ro,mx,my,mz,en,bx,by,bz
};
struct G3 {
  int nx, ny, nz, xoff, yoff, zoff;
  double gamma;
};
// rotated view of one sweep direction
struct Sw3 {
  const double *ro, *vn, *vt, *vu, *bt, *bu, *pr, *m2;
  const double *bnc;    // cell-centred normal B (reconstruction)
  const double *bface;  // face-centred normal B (state)
  const double *dvt, *dvu;
  double *out[7];
  double *fc[2];
};

// ------------------------------------------------------------------- prim3d
// Cell (i, j, k), any cell of the grid.
template <int ODR>
__device__ __forceinline__ void prim3d_body(U3 u, double *vx_, double *vy_, double *vz_, double *pr_, double *m2_,
                                            double *cx_, double *cy_, double *cz_, G3 g, const BcFlags &bc,
                                            int i, int j, int k)
{
  const long nx = g.nx, nxy = (long)g.nx * g.ny;
  const long s = k * nxy + j * nx + i;
  double c[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    double sg = 1.0;
    const int fld = 5 + a;
    const int si = bc_src(i, g.nx, g.xoff, 0, bc.dnx[fld], &sg);
    const int sj = bc_src(j, g.ny, g.yoff, 0, bc.dny[fld], &sg);
    const int sk = bc_src(k, g.nz, g.zoff, 0, bc.dnz[fld], &sg);
    const long st = (a == 0) ? 1 : (a == 1) ? nx : nxy;
    const double *b = u.f[fld] + sk * nxy + sj * nx + si;
    c[a] = sg * bc_centre<ODR>(b[-st], b[0], b[st], b[2 * st]);
  }
  const double ro = u.f[0][s];
  const double iro = 1.0 / ro;
  const double vx = u.f[1][s] * iro, vy = u.f[2][s] * iro, vz = u.f[3][s] * iro;
  const double v2 = (vx * vx + vy * vy + vz * vz);
  const double b2 = (c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  const double pr = (g.gamma - 1.0) * (u.f[4][s] - 0.5 * (ro * v2 + b2));
  vx_[s] = vx;
  vy_[s] = vy;
  vz_[s] = vz;
  pr_[s] = pr;
  m2_[s] = v2 / (g.gamma * pr * iro);
  cx_[s] = c[0];
  cy_[s] = c[1];
  cz_[s] = c[2];
}

// -------------------------------------------------------------------- aux3d
// ct weight on the corner of a plane spanned by axes (a,b): offsets sa, sb
__device__ __forceinline__ double ct_weight(const double *va, const double *vb, const double *ro, const double *ba,
                                            const double *bb, long s, long sa, long sb)
{
  double vac = 0.25 * (va[s - sb - sa] + va[s - sb] + va[s - sa] + va[s]);
  double vbc = 0.25 * (vb[s - sb - sa] + vb[s - sb] + vb[s - sa] + vb[s]);
  double roc = 0.25 * (ro[s - sb - sa] + ro[s - sb] + ro[s - sa] + ro[s]);
  roc = 1.0 / sqrt(roc);
  const double bac = 0.5 * (ba[s - sb] + ba[s]);
  const double bbc = 0.5 * (bb[s - sa] + bb[s]);
  vac = fabs(vac) + fabs(bac * roc);
  vbc = fabs(vbc) + fabs(bbc * roc);
  const double eps = 1e-6;
  return (vac + 0.5 * eps) / (vac + vbc + eps);
}

// Cell (i, j, k), any cell of the grid.
template <int CTW>
__device__ __forceinline__ void aux3d_body(U3 u, const double *vx, const double *vy, const double *vz, double *ct0,
                                           double *ct1, double *ct2, double *dvx, double *dvy, double *dvz, G3 g,
                                           const BcFlags &bc, int i, int j, int k)
{
  const long nx = g.nx, nxy = (long)g.nx * g.ny;
  const long s = k * nxy + j * nx + i;
  double w0 = 0.5, w1 = 0.5, w2 = 0.5;
  if (CTW) {
    // xy plane (:268-294): axes x (a) and y (b), Bx averaged along y, By along x
    if (i >= 1 && j >= 1) w0 = ct_weight(vx, vy, u.f[0], u.f[5], u.f[6], s, 1, nx);
    // yz plane (:299-320): a = y, b = z
    if (k >= 1 && j >= 1) w1 = ct_weight(vy, vz, u.f[0], u.f[6], u.f[7], s, nx, nxy);
    // zx plane (:321-338): a = z ... written with a = x, b = z after reordering below
    if (k >= 1 && i >= 1) {
      double vzc = 0.25 * (vz[s - nxy - 1] + vz[s - nxy] + vz[s - 1] + vz[s]);
      double vxc = 0.25 * (vx[s - nxy - 1] + vx[s - nxy] + vx[s - 1] + vx[s]);
      double roc = 0.25 * (u.f[0][s - nxy - 1] + u.f[0][s - nxy] + u.f[0][s - 1] + u.f[0][s]);
      roc = 1.0 / sqrt(roc);
      const double bzc = 0.5 * (u.f[7][s - 1] + u.f[7][s]);
      const double bxc = 0.5 * (u.f[5][s - nxy] + u.f[5][s]);
      vzc = fabs(vzc) + fabs(bzc * roc);
      vxc = fabs(vxc) + fabs(bxc * roc);
      const double eps = 1e-6;
      w2 = (vzc + 0.5 * eps) / (vzc + vxc + eps);
    }
  }
  ct0[s] = w0;
  ct1[s] = w1;
  ct2[s] = w2;
  const double *vv[3] = {vx, vy, vz};
  double *dd[3] = {dvx, dvy, dvz};
  const long st[3] = {1, nx, nxy};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    // dv_a boundary factors (:361-375): the axis' own flag is negated
    double sg = 1.0;
    const int fl = 1 + a;
    const int si = bc_src(i, g.nx, g.xoff, 0, (a == 0 ? -1 : 1) * bc.dnx[fl], &sg);
    const int sj = bc_src(j, g.ny, g.yoff, 0, (a == 1 ? -1 : 1) * bc.dny[fl], &sg);
    const int sk = bc_src(k, g.nz, g.zoff, 0, (a == 2 ? -1 : 1) * bc.dnz[fl], &sg);
    const double *v = vv[a] + sk * nxy + sj * nx + si;
    dd[a][s] = sg * dmin((v[st[a]] - v[0]), (v[0] - v[-st[a]]));
  }
}

// ct_weight on the four corner values of a plane (a: first axis, b: second,
// r: density, in ct_weight's order s-sb-sa, s-sb, s-sa, s) and the two face-B
// averages; the same operations as ct_weight.
__device__ __forceinline__ double ct_w4(double a0, double a1, double a2, double a3, double b0, double b1, double b2,
                                        double b3, double r0, double r1, double r2, double r3, double bac, double bbc)
{
  double vac = 0.25 * (a0 + a1 + a2 + a3);
  double vbc = 0.25 * (b0 + b1 + b2 + b3);
  double roc = 0.25 * (r0 + r1 + r2 + r3);
  roc = 1.0 / sqrt(roc);
  vac = fabs(vac) + fabs(bac * roc);
  vbc = fabs(vbc) + fabs(bbc * roc);
  const double eps = 1e-6;
  return (vac + 0.5 * eps) / (vac + vbc + eps);
}

// aux3d_body on cells (i, j, kc*CZ ..): one thread walks CZ cells along z and
// keeps the values of the plane below that the CT weights and the z shock
// sensor read (velocities and density at (i,j), (i-1,j), (i,j-1); Bx, By at
// (i,j)), so that plane is not read again.  Same arithmetic as aux3d_body.
template <int CTW, int CZ>
__device__ __forceinline__ void aux3d_zm_body(U3 u, const double *vx, const double *vy, const double *vz, double *ct0,
                                              double *ct1, double *ct2, double *dvx, double *dvy, double *dvz, G3 g,
                                              const BcFlags &bc, int i, int j, int kc)
{
  const long nx = g.nx, nxy = (long)g.nx * g.ny;
  const int k0 = kc * CZ, k1 = (k0 + CZ - 1 < g.nz - 1) ? k0 + CZ - 1 : g.nz - 1;
  const bool hw = i >= 1, hs = j >= 1;
  const double *ro = u.f[0];
  // previous plane: (i,j): x,y,z,r; (i,j-1): y,z,r; (i-1,j): z,x,r; Bx, By at (i,j)
  double cx0 = 0, cy0 = 0, cz0 = 0, cr0 = 0, sy0 = 0, sz0 = 0, sr0 = 0, wz0 = 0, wx0 = 0, wr0 = 0, bx0 = 0, by0 = 0;
  if (k0 >= 1) {
    const long q = (long)(k0 - 1) * nxy + j * nx + i;
    cx0 = vx[q]; cy0 = vy[q]; cz0 = vz[q]; cr0 = ro[q];
    if (CTW) {
      if (hs) { sy0 = vy[q - nx]; sz0 = vz[q - nx]; sr0 = ro[q - nx]; }
      if (hw) { wz0 = vz[q - 1]; wx0 = vx[q - 1]; wr0 = ro[q - 1]; }
      bx0 = u.f[5][q];
      by0 = u.f[6][q];
    }
  }
  for (int k = k0; k <= k1; k++) {
    const long s = k * nxy + j * nx + i;
    const double cx = vx[s], cy = vy[s], cz = vz[s], cr = ro[s];
    const double bxs = u.f[5][s], bys = u.f[6][s];
    double w0 = 0.5, w1 = 0.5, w2 = 0.5;
    double sy = 0, sz = 0, sr = 0, wz = 0, wx = 0, wr = 0;
    if (CTW) {
      if (hs) { sy = vy[s - nx]; sz = vz[s - nx]; sr = ro[s - nx]; }
      if (hw) { wz = vz[s - 1]; wx = vx[s - 1]; wr = ro[s - 1]; }
      if (hw && hs) {  // xy plane (:268-294): a = x, b = y
        const double sx = vx[s - nx], wy = vy[s - 1];
        w0 = ct_w4(vx[s - nx - 1], sx, wx, cx, vy[s - nx - 1], sy, wy, cy, ro[s - nx - 1], sr, wr, cr,
                   0.5 * (u.f[5][s - nx] + bxs), 0.5 * (u.f[6][s - 1] + bys));
      }
      if (k >= 1 && hs)  // yz plane (:299-320): a = y, b = z
        w1 = ct_w4(sy0, cy0, sy, cy, sz0, cz0, sz, cz, sr0, cr0, sr, cr, 0.5 * (by0 + bys),
                   0.5 * (u.f[7][s - nx] + u.f[7][s]));
      if (k >= 1 && hw)  // zx plane (:321-338): a = z, b = x
        w2 = ct_w4(wz0, cz0, wz, cz, wx0, cx0, wx, cx, wr0, cr0, wr, cr, 0.5 * (u.f[7][s - 1] + u.f[7][s]),
                   0.5 * (bx0 + bxs));
    }
    ct0[s] = w0;
    ct1[s] = w1;
    ct2[s] = w2;
    const double *vv[3] = {vx, vy, vz};
    double *dd[3] = {dvx, dvy, dvz};
    const long st[3] = {1, nx, nxy};
#pragma unroll
    for (int a = 0; a < 3; a++) {
      // dv_a boundary factors (:361-375): the axis' own flag is negated
      double sg = 1.0;
      const int fl = 1 + a;
      const int si = bc_src(i, g.nx, g.xoff, 0, (a == 0 ? -1 : 1) * bc.dnx[fl], &sg);
      const int sj = bc_src(j, g.ny, g.yoff, 0, (a == 1 ? -1 : 1) * bc.dny[fl], &sg);
      const int sk = bc_src(k, g.nz, g.zoff, 0, (a == 2 ? -1 : 1) * bc.dnz[fl], &sg);
      const long q = sk * nxy + sj * nx + si;
      const double *v = vv[a] + q;
      // an interior cell (q == s; then k >= 1 and the plane below is at hand)
      // takes vz of its own and of the lower cell from registers
      const bool here = (a == 2) && (q == s);
      const double v0 = here ? cz : v[0];
      const double vm = here ? cz0 : v[-st[a]];
      dd[a][s] = sg * dmin((v[st[a]] - v0), (v0 - vm));
    }
    cx0 = cx; cy0 = cy; cz0 = cz; cr0 = cr;
    sy0 = sy; sz0 = sz; sr0 = sr;
    wz0 = wz; wx0 = wx; wr0 = wr;
    bx0 = bxs; by0 = bys;
  }
}

// ------------------------------------------------------------- sweep cores
template <int ODR>
__device__ __forceinline__ void central_lr(const double (&bn)[5], const double (&vn)[5], const double (&vt)[5],
                                           double bl, double br, double &ql, double &qr)
{
  double p[5];
#pragma unroll
  for (int m = 0; m < 5; m++) p[m] = bn[m] * vn[m];
  const double a0 = lcal_l<ODR>(p[0], p[1], p[2], p[3], p[4]);
  const double b0 = lcal_r<ODR>(p[0], p[1], p[2], p[3], p[4]);
  const double a1 = lcal_l<ODR>(vt[0], vt[1], vt[2], vt[3], vt[4]);
  const double b1 = lcal_r<ODR>(vt[0], vt[1], vt[2], vt[3], vt[4]);
  ql = a0 - bl * a1;
  qr = b0 - br * b1;
}

template <int FLUX, int MIX>
__device__ __forceinline__ void face_flux(const Sw3 &w, long q, long st, const double (&L)[7], const double (&R)[7],
                                          double qlt, double qlu, double qrt, double qru, double gamma)
{
  const double bf = w.bface[q];
  const double bn = 0.5 * (bf + bf);
  const double dvtm = dmin(dmin(w.dvt[q - st], w.dvt[q]), dmin(w.dvu[q - st], w.dvu[q]));
  const double dv0 = w.vn[q] - w.vn[q - st];
  double f[7];
  riemann_p<FLUX, MIX>(L, R, bn, gamma, dv0, dvtm, f);
  const double fct = 0.5 * (qlt + qrt);
  const double fcu = 0.5 * (qlu + qru);
  w.out[0][q] = f[0];
  w.out[1][q] = f[1];
  w.out[2][q] = f[2];
  w.out[3][q] = f[3];
  w.out[4][q] = f[4] - fct;
  w.out[5][q] = f[5] - fcu;
  w.out[6][q] = f[6];
  w.fc[0][q] = fct;
  w.fc[1][q] = fcu;
}

template <int ODR>
__device__ __forceinline__ void load_stencil(const Sw3 &w, long q, long st, Stencil &S, double (&bn5)[5])
{
#pragma unroll
  for (int m = 0; m < 5; m++) {
    const long o = q + (m - 2) * st;
    S.q[0][m] = w.ro[o];
    S.q[1][m] = w.vn[o];
    S.q[2][m] = w.vt[o];
    S.q[3][m] = w.vu[o];
    S.q[4][m] = w.bt[o];
    S.q[5][m] = w.bu[o];
    S.q[6][m] = w.pr[o];
    bn5[m] = 0.0;
  }
}

// X sweep: wave-shuffle of the left face state over cells 2..nx-3 of all
// rows (j,k) flattened, as xsweep2d_body (xsweep_waves wavefronts).
template <int ODR, int FLUX, int MIX>
__device__ __forceinline__ void xsweep3d_body(const Sw3 &w, G3 g, long wave, int lane)
{
  const int nx = g.nx, rl = nx - 4;
  const long total = (long)g.ny * g.nz * rl;
  long q = 63 * wave + lane;
  const bool live = q < total;
  if (!live) q = total - 1;
  const int c = 2 + (int)(q % rl);
  const long s = (q / rl) * nx + c;
  Stencil S;
  double dummy[5];
  load_stencil<ODR>(w, s, 1, S, dummy);
  const bool sup = dmax(w.m2[s - 1], dmax(w.m2[s], w.m2[s + 1])) >= 1.0;
  double vl[7], vr[7];
  recon_p<ODR, MIX>(S, sup, w.bnc[s], g.gamma, vl, vr);
  double qlt, qrt, qlu, qru;
  central_lr<ODR>(S.q[4], S.q[1], S.q[2], w.bface[s + 1], w.bface[s], qlt, qrt);
  central_lr<ODR>(S.q[5], S.q[1], S.q[3], w.bface[s + 1], w.bface[s], qlu, qru);
  double L[7];
#pragma unroll
  for (int k = 0; k < 7; k++) L[k] = shfl_up1(vl[k]);
  const double qltL = shfl_up1(qlt), qluL = shfl_up1(qlu);
  if (lane == 0 || !live || c == 2) return;
  face_flux<FLUX, MIX>(w, s, 1, L, vr, qltL, qluL, qrt, qru, g.gamma);
}

// Y/Z sweep: one lane per column (i, bz) with bz the third coordinate (k for
// Y, j for Z); the lane marches faces f0..f0+CH-1 (f0 = 3+CH*by) along the
// sweep axis.
template <int ODR, int FLUX, int CH, int AX, int MIX>
__device__ __forceinline__ void msweep3d_body(const Sw3 &w, G3 g, int i, int by, int bz)
{
  const long nx = g.nx, nxy = (long)g.nx * g.ny;
  const int n = (AX == 1) ? g.ny : g.nz;
  const long st = (AX == 1) ? nx : nxy;
  const long base = (AX == 1) ? (long)bz * nxy + i : (long)bz * nx + i;
  const int f0 = 3 + CH * by;
  if (f0 > n - 3) return;
  const int f1 = (f0 + CH - 1 < n - 3) ? f0 + CH - 1 : n - 3;
  // window rows r-2..r+2 of cell r; bit m of supb: m2 >= 1 in window row m
  // (the reconstruction selector takes the maximum of rows r-1..r+1; for
  // finite values that is the same test)
  double W[7][5];
  unsigned supb = 0;
  auto load = [&](int r, int m) {
    const long q = base + r * st;
    W[0][m] = w.ro[q];
    W[1][m] = w.vn[q];
    W[2][m] = w.vt[q];
    W[3][m] = w.vu[q];
    W[4][m] = w.bt[q];
    W[5][m] = w.bu[q];
    W[6][m] = w.pr[q];
    supb = (supb & ~(1u << m)) | ((w.m2[q] >= 1.0 ? 1u : 0u) << m);
  };
  auto cell = [&](int r, double (&vl)[7], double (&vr)[7], double &qlt, double &qrt, double &qlu, double &qru) {
    Stencil S;
#pragma unroll
    for (int a = 0; a < 7; a++)
#pragma unroll
      for (int m = 0; m < 5; m++) S.q[a][m] = W[a][m];
    const bool sup = (supb & 14u) != 0;
    const long q = base + r * st;
    recon_p<ODR, MIX>(S, sup, w.bnc[q], g.gamma, vl, vr);
    central_lr<ODR>(W[4], W[1], W[2], w.bface[q + st], w.bface[q], qlt, qrt);
    central_lr<ODR>(W[5], W[1], W[3], w.bface[q + st], w.bface[q], qlu, qru);
  };
#pragma unroll
  for (int m = 0; m < 5; m++) load(f0 - 3 + m, m);
  double vlp[7], vr[7], qltp, qlup, qrt, qru;
  cell(f0 - 1, vlp, vr, qltp, qrt, qlup, qru);
  for (int r = f0; r <= f1; r++) {
#pragma unroll
    for (int m = 0; m < 4; m++) {
#pragma unroll
      for (int a = 0; a < 7; a++) W[a][m] = W[a][m + 1];
    }
    supb >>= 1;
    load(r + 2, 4);
    double vl[7], qlt, qlu;
    cell(r, vl, vr, qlt, qrt, qlu, qru);
    face_flux<FLUX, MIX>(w, base + r * st, st, vlp, vr, qltp, qlup, qrt, qru, g.gamma);
#pragma unroll
    for (int a = 0; a < 7; a++) vlp[a] = vl[a];
    qltp = qlt;
    qlup = qlu;
  }
}

// ------------------------------------------------------------------- edge3d
struct Edge3 {
  const double *fxt, *fxu, *fcx0, *fcx1;  // x sweep: By upwind (t), Bz upwind (u), centrals
  const double *fyt, *fyu, *fcy0, *fcy1;  // y sweep: Bz upwind (t), Bx upwind (u)
  const double *fzt, *fzu, *fcz0, *fcz1;  // z sweep: Bx upwind (t), By upwind (u)
  const double *ct0, *ct1, *ct2;
  double *ex, *ey, *ez;
};

// corner value of a face quantity interpolated along one axis: left value of
// the stencil centred one cell below plus right value of the stencil centred here
template <int ODR>
__device__ __forceinline__ double corner_sum(const double *f, long st)
{
  const double l = lcal_l<ODR>(f[-3 * st], f[-2 * st], f[-st], f[0], f[st]);
  const double r = lcal_r<ODR>(f[-2 * st], f[-st], f[0], f[st], f[2 * st]);
  return l + r;
}

template <int ODR>
__device__ __forceinline__ void edge3d_body(const Edge3 &e, G3 g, unsigned bx, unsigned jj, unsigned kk, unsigned tid)
{
  const int i = bx * 256 + tid;
  const int j = jj, k = kk;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (i >= nx) return;
  const long NX = nx, NXY = (long)nx * ny;
  const long s = k * NXY + j * NX + i;
  const bool ix = (i >= 3 && i <= nx - 3), jy = (j >= 3 && j <= ny - 3), kz = (k >= 3 && k <= nz - 3);
  if (ix && jy) {  // Ez (:494-540 then :748-793)
    const double c = e.ct0[s];
    double v = 0.0;
    v += -0.5 * (corner_sum<ODR>(e.fxt + s, NX) + (1.0 - c) * corner_sum<ODR>(e.fcx0 + s, NX));
    v += +0.5 * (corner_sum<ODR>(e.fyu + s, 1) + c * corner_sum<ODR>(e.fcy1 + s, 1));
    e.ez[s] = v;
  }
  if (ix && kz) {  // Ey (:541-585 then :910-955)
    const double c = e.ct2[s];
    double v = 0.0;
    v += +0.5 * (corner_sum<ODR>(e.fxu + s, NXY) + c * corner_sum<ODR>(e.fcx1 + s, NXY));
    v += -0.5 * (corner_sum<ODR>(e.fzt + s, 1) + (1.0 - c) * corner_sum<ODR>(e.fcz0 + s, 1));
    e.ey[s] = v;
  }
  if (jy && kz) {  // Ex (:702-747 then :956-1001)
    const double c = e.ct1[s];
    double v = 0.0;
    v += -0.5 * (corner_sum<ODR>(e.fyt + s, NXY) + (1.0 - c) * corner_sum<ODR>(e.fcy0 + s, NXY));
    v += +0.5 * (corner_sum<ODR>(e.fzu + s, NX) + c * corner_sum<ODR>(e.fcz1 + s, NX));
    e.ex[s] = v;
  }
}

// ----------------------------------------------------------------- update3d
struct Fl3 {
  const double *x[7], *y[7], *z[7];
};
// Block (bx, jj, kk) covers cells xoff+256*bx.. of row (yoff+jj, zoff+kk).
template <int ODR>
__device__ __forceinline__ void update3d_body(U3 in, U3 ut, U3 out, const Fl3 &fl, const double *ex,
                                              const double *ey, const double *ez, G3 g, double dtdx, double dtdy,
                                              double dtdz, double ra, double rb, unsigned bx, unsigned jj,
                                              unsigned kk, unsigned tid)
{
  const int i = g.xoff + bx * 256 + tid;
  const int j = g.yoff + jj, k = g.zoff + kk;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (i > nx - g.xoff) return;
  const long NX = nx, NXY = (long)nx * ny;
  const long s = k * NXY + j * NX + i;
  const bool ci = i < nx - g.xoff, cj = j < ny - g.yoff, ck = k < nz - g.zoff;
  if (ci && cj && ck) {
    // ro, mx, my, mz, en: component index in each sweep's rotated order
    const int fld[5] = {0, 1, 2, 3, 4};
    const int cx_[5] = {0, 1, 2, 3, 6}, cy_[5] = {0, 3, 1, 2, 6}, cz_[5] = {0, 2, 3, 1, 6};
#pragma unroll
    for (int t = 0; t < 5; t++) {
      const double *a = fl.x[cx_[t]] + s, *b = fl.y[cy_[t]] + s, *c = fl.z[cz_[t]] + s;
      const double d0 = diff4<ODR>(a[-1], a[0], a[1], a[2]);
      const double d1 = diff4<ODR>(b[-NX], b[0], b[NX], b[2 * NX]);
      const double d2 = diff4<ODR>(c[-NXY], c[0], c[NXY], c[2 * NXY]);
      out.f[fld[t]][s] = ra * ut.f[fld[t]][s] + rb * (in.f[fld[t]][s] - dtdx * d0 - dtdy * d1 - dtdz * d2);
    }
  }
  if (cj && ck) {  // Bx (:1083-1105)
    const double du0 = diff4<ODR>(ez[s - NX], ez[s], ez[s + NX], ez[s + 2 * NX]);
    const double du1 = diff4<ODR>(ey[s - NXY], ey[s], ey[s + NXY], ey[s + 2 * NXY]);
    out.f[5][s] = ra * ut.f[5][s] + rb * (in.f[5][s] - dtdy * du0 + dtdz * du1);
  }
  if (ci && ck) {  // By (:1106-1127)
    const double du0 = diff4<ODR>(ex[s - NXY], ex[s], ex[s + NXY], ex[s + 2 * NXY]);
    const double du1 = diff4<ODR>(ez[s - 1], ez[s], ez[s + 1], ez[s + 2]);
    out.f[6][s] = ra * ut.f[6][s] + rb * (in.f[6][s] - dtdz * du0 + dtdx * du1);
  }
  if (ci && cj) {  // Bz (:1129-1157)
    const double du0 = diff4<ODR>(ey[s - 1], ey[s], ey[s + 1], ey[s + 2]);
    const double du1 = diff4<ODR>(ex[s - NX], ex[s], ex[s + NX], ex[s + 2 * NX]);
    out.f[7][s] = ra * ut.f[7][s] + rb * (in.f[7][s] - dtdx * du0 + dtdy * du1);
  }
}

// ------------------------------------------------------------- plane tiles
// Cell kernels tile each plane (or the face-inclusive interior of a plane) in
// 256-cell pieces of the flattened (j, i) index, so rows narrower than a
// block do not leave blocks nearly empty.
__device__ __forceinline__ bool plane_cell(G3 g, unsigned piece, unsigned tid, int &i, int &j)
{
  const long p = (long)piece * 256 + tid;
  if (p >= (long)g.nx * g.ny) return false;
  i = (int)(p % g.nx);
  j = (int)(p / g.nx);
  return true;
}
__device__ __forceinline__ bool plane_face_cell(G3 g, unsigned piece, unsigned tid, int &i, int &j)
{
  const int ni = g.nx - 2 * g.xoff + 1, nj = g.ny - 2 * g.yoff + 1;
  const long p = (long)piece * 256 + tid;
  if (p >= (long)ni * nj) return false;
  i = g.xoff + (int)(p % ni);
  j = g.yoff + (int)(p / ni);
  return true;
}
__host__ __device__ __forceinline__ unsigned plane_pieces(int nx, int ny) { return (unsigned)(((long)nx * ny + 255) / 256); }

// ------------------------------------------------------- z-marching variants
// The same per-cell arithmetic as edge3d_body / update3d_body, but one thread
// walks a chunk of CZ cells along z and keeps the z-stencil values of the
// previous cells in registers, so each plane of a z-differenced array is read
// once per chunk instead of four to six times.

// corner_sum along z from a window w[0..5] = f[k-3..k+2]
template <int ODR>
__device__ __forceinline__ double corner_sum_w(const double (&w)[6])
{
  const double l = lcal_l<ODR>(w[0], w[1], w[2], w[3], w[4]);
  const double r = lcal_r<ODR>(w[1], w[2], w[3], w[4], w[5]);
  return l + r;
}

// Column (i, j), planes kc*CZ.. kc*CZ+CZ-1 (any column of the grid).
template <int ODR, int CZ>
__device__ __forceinline__ void edge3d_zm_body(const Edge3 &e, G3 g, int i, int j, int kc)
{
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const long NX = nx, NXY = (long)nx * ny;
  const int k0 = kc * CZ, k1 = (k0 + CZ - 1 < nz - 1) ? k0 + CZ - 1 : nz - 1;
  const bool ix = (i >= 3 && i <= nx - 3), jy = (j >= 3 && j <= ny - 3);
  double wxu[6], wcx[6], wyt[6], wcy[6];  // fxu, fcx1, fyt, fcy0 at planes k-3..k+2
  int kw = -10;                           // plane whose window is loaded
  for (int k = k0; k <= k1; k++) {
    const long s = k * NXY + j * NX + i;
    if (ix && jy) {  // Ez (:494-540 then :748-793)
      const double c = e.ct0[s];
      double v = 0.0;
      v += -0.5 * (corner_sum<ODR>(e.fxt + s, NX) + (1.0 - c) * corner_sum<ODR>(e.fcx0 + s, NX));
      v += +0.5 * (corner_sum<ODR>(e.fyu + s, 1) + c * corner_sum<ODR>(e.fcy1 + s, 1));
      e.ez[s] = v;
    }
    const bool kz = (k >= 3 && k <= nz - 3);
    if (!kz || !(ix || jy)) continue;
    if (kw == k - 1) {
#pragma unroll
      for (int m = 0; m < 5; m++) {
        wxu[m] = wxu[m + 1];
        wcx[m] = wcx[m + 1];
        wyt[m] = wyt[m + 1];
        wcy[m] = wcy[m + 1];
      }
      const long q = s + 2 * NXY;
      wxu[5] = ix ? e.fxu[q] : 0.0;
      wcx[5] = ix ? e.fcx1[q] : 0.0;
      wyt[5] = jy ? e.fyt[q] : 0.0;
      wcy[5] = jy ? e.fcy0[q] : 0.0;
    } else {
#pragma unroll
      for (int m = 0; m < 6; m++) {
        const long q = s + (m - 3) * NXY;
        wxu[m] = ix ? e.fxu[q] : 0.0;
        wcx[m] = ix ? e.fcx1[q] : 0.0;
        wyt[m] = jy ? e.fyt[q] : 0.0;
        wcy[m] = jy ? e.fcy0[q] : 0.0;
      }
    }
    kw = k;
    if (ix) {  // Ey (:541-585 then :910-955)
      const double c = e.ct2[s];
      double v = 0.0;
      v += +0.5 * (corner_sum_w<ODR>(wxu) + c * corner_sum_w<ODR>(wcx));
      v += -0.5 * (corner_sum<ODR>(e.fzt + s, 1) + (1.0 - c) * corner_sum<ODR>(e.fcz0 + s, 1));
      e.ey[s] = v;
    }
    if (jy) {  // Ex (:702-747 then :956-1001)
      const double c = e.ct1[s];
      double v = 0.0;
      v += -0.5 * (corner_sum_w<ODR>(wyt) + (1.0 - c) * corner_sum_w<ODR>(wcy));
      v += +0.5 * (corner_sum<ODR>(e.fzu + s, NX) + c * corner_sum<ODR>(e.fcz1 + s, NX));
      e.ex[s] = v;
    }
  }
}

// Column (i, j) with xoff <= i <= nx-xoff, yoff <= j <= ny-yoff, planes
// zoff+kc*CZ.. up to nz-zoff (one extra face per axis for the CT update).
template <int ODR, int CZ>
__device__ __forceinline__ void update3d_zm_body(U3 in, U3 ut, U3 out, const Fl3 &fl, const double *ex,
                                                 const double *ey, const double *ez, G3 g, double dtdx,
                                                 double dtdy, double dtdz, double ra, double rb, int i, int j, int kc)
{
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const long NX = nx, NXY = (long)nx * ny;
  const int k0 = g.zoff + kc * CZ;
  const int k1 = (k0 + CZ - 1 < nz - g.zoff) ? k0 + CZ - 1 : nz - g.zoff;
  const bool ci = i < nx - g.xoff, cj = j < ny - g.yoff;
  const int cz_[5] = {0, 2, 3, 1, 6};
  // windows over planes k-1..k+2: z fluxes of ro,mx,my,mz,en; ey (for Bx); ex (for By)
  double wz[5][4], wey[4], wex[4];
  {
    const long s0 = (long)k0 * NXY + j * NX + i;
#pragma unroll
    for (int m = 0; m < 3; m++) {
      const long q = s0 + (m - 1) * NXY;
#pragma unroll
      for (int t = 0; t < 5; t++) wz[t][m + 1] = (ci && cj) ? fl.z[cz_[t]][q] : 0.0;
      wey[m + 1] = cj ? ey[q] : 0.0;
      wex[m + 1] = ci ? ex[q] : 0.0;
    }
  }
  for (int k = k0; k <= k1; k++) {
    const long s = k * NXY + j * NX + i;
    {
      const long q = s + 2 * NXY;
#pragma unroll
      for (int t = 0; t < 5; t++) {
#pragma unroll
        for (int m = 0; m < 3; m++) wz[t][m] = wz[t][m + 1];
        wz[t][3] = (ci && cj) ? fl.z[cz_[t]][q] : 0.0;
      }
#pragma unroll
      for (int m = 0; m < 3; m++) {
        wey[m] = wey[m + 1];
        wex[m] = wex[m + 1];
      }
      wey[3] = cj ? ey[q] : 0.0;
      wex[3] = ci ? ex[q] : 0.0;
    }
    const bool ck = k < nz - g.zoff;
    if (ci && cj && ck) {
      const int fld[5] = {0, 1, 2, 3, 4};
      const int cx_[5] = {0, 1, 2, 3, 6}, cy_[5] = {0, 3, 1, 2, 6};
#pragma unroll
      for (int t = 0; t < 5; t++) {
        const double *a = fl.x[cx_[t]] + s, *b = fl.y[cy_[t]] + s;
        const double d0 = diff4<ODR>(a[-1], a[0], a[1], a[2]);
        const double d1 = diff4<ODR>(b[-NX], b[0], b[NX], b[2 * NX]);
        const double d2 = diff4<ODR>(wz[t][0], wz[t][1], wz[t][2], wz[t][3]);
        out.f[fld[t]][s] = ra * ut.f[fld[t]][s] + rb * (in.f[fld[t]][s] - dtdx * d0 - dtdy * d1 - dtdz * d2);
      }
    }
    if (cj && ck) {  // Bx (:1083-1105)
      const double du0 = diff4<ODR>(ez[s - NX], ez[s], ez[s + NX], ez[s + 2 * NX]);
      const double du1 = diff4<ODR>(wey[0], wey[1], wey[2], wey[3]);
      out.f[5][s] = ra * ut.f[5][s] + rb * (in.f[5][s] - dtdy * du0 + dtdz * du1);
    }
    if (ci && ck) {  // By (:1106-1127)
      const double du0 = diff4<ODR>(wex[0], wex[1], wex[2], wex[3]);
      const double du1 = diff4<ODR>(ez[s - 1], ez[s], ez[s + 1], ez[s + 2]);
      out.f[6][s] = ra * ut.f[6][s] + rb * (in.f[6][s] - dtdz * du0 + dtdx * du1);
    }
    if (ci && cj) {  // Bz (:1129-1157)
      const double du0 = diff4<ODR>(ey[s - 1], ey[s], ey[s + 1], ey[s + 2]);
      const double du1 = diff4<ODR>(ex[s - NX], ex[s], ex[s + NX], ex[s + 2 * NX]);
      out.f[7][s] = ra * ut.f[7][s] + rb * (in.f[7][s] - dtdx * du0 + dtdy * du1);
    }
  }
}

// --------------------------------------------------------------------- bc3d
// Three ghost slabs per field (x: 2*xoff x ny x nz, y: nx x 2*yoff x nz,
// z: nx x ny x 2*zoff).  A cell in several slabs gets the same value from each.
__device__ __forceinline__ long bc3d_points(G3 g)
{
  return (long)2 * g.xoff * g.ny * g.nz + (long)g.nx * 2 * g.yoff * g.nz + (long)g.nx * g.ny * 2 * g.zoff;
}
__device__ __forceinline__ void bc3d_body(U3 u, G3 g, const BcFlags &bc, int fld, long t)
{
  const int nx = g.nx, ny = g.ny, nz = g.nz, xo = g.xoff, yo = g.yoff, zo = g.zoff;
  const long sx = (long)2 * xo * ny * nz, sy = (long)nx * 2 * yo * nz, sz = (long)nx * ny * 2 * zo;
  int i, j, k;
  if (t < sx) {
    const int a = (int)(t % (2 * xo));
    const long r = t / (2 * xo);
    j = (int)(r % ny);
    k = (int)(r / ny);
    i = (a < xo) ? a : nx - 2 * xo + a;
  } else if ((t -= sx) < sy) {
    i = (int)(t % nx);
    const long r = t / nx;
    const int a = (int)(r % (2 * yo));
    k = (int)(r / (2 * yo));
    j = (a < yo) ? a : ny - 2 * yo + a;
  } else if ((t -= sy) < sz) {
    i = (int)(t % nx);
    const long r = t / nx;
    j = (int)(r % ny);
    const int a = (int)(r / ny);
    k = (a < zo) ? a : nz - 2 * zo + a;
  } else {
    return;
  }
  double sg = 1.0;
  const int si = bc_src(i, nx, xo, bc.stx[fld], bc.dnx[fld], &sg);
  const int sj = bc_src(j, ny, yo, bc.sty[fld], bc.dny[fld], &sg);
  const int sk = bc_src(k, nz, zo, bc.stz[fld], bc.dnz[fld], &sg);
  if (si == i && sj == j && sk == k) return;
  double *f = u.f[fld];
  const long NX = nx, NXY = (long)nx * ny;
  f[k * NXY + j * NX + i] = sg * f[sk * NXY + sj * NX + si];
}

// -------------------------------------------------------------------- cfl3d
// Signal speed |v|+c_f of interior cell t (3D/MPI/OTvortex/cflcheck.hpp:1-39).
__device__ __forceinline__ double cfl3d_cell(U3 u, G3 g, long t)
{
  const int ni = g.nx - 2 * g.xoff, nj = g.ny - 2 * g.yoff;
  const long NX = g.nx, NXY = (long)g.nx * g.ny;
  const int i = g.xoff + (int)(t % ni);
  const long r = t / ni;
  const int j = g.yoff + (int)(r % nj), k = g.zoff + (int)(r / nj);
  const long s = k * NXY + j * NX + i;
  const double iro = 1.0 / u.f[0][s];
  const double vx = u.f[1][s] * iro, vy = u.f[2][s] * iro, vz = u.f[3][s] * iro;
  const double bxc = 0.5 * (u.f[5][s] + u.f[5][s + 1]);
  const double byc = 0.5 * (u.f[6][s] + u.f[6][s + NX]);
  const double bzc = 0.5 * (u.f[7][s] + u.f[7][s + NXY]);
  const double v2 = vx * vx + vy * vy + vz * vz;
  const double b2 = bxc * bxc + byc * byc + bzc * bzc;
  const double pr = (g.gamma - 1.0) * (u.f[4][s] - 0.5 * (u.f[0][s] * v2 + b2));
  const double cf = sqrt((0.5 * g.gamma * (2.0 * pr) + b2) / (u.f[0][s]));
  return sqrt(v2) + cf;
}

// Ensemble form: +Inf when the cell's density or energy is not finite (the
// member then fails), otherwise the speed of cfl3d_cell.
__device__ __forceinline__ double cfl3d_cell_checked(U3 u, G3 g, long t)
{
  const int ni = g.nx - 2 * g.xoff, nj = g.ny - 2 * g.yoff;
  const long r = t / ni;
  const long s = ((long)(g.zoff + (int)(r / nj)) * g.ny + g.yoff + (int)(r % nj)) * g.nx + g.xoff + (int)(t % ni);
  if (!isfinite(u.f[0][s]) || !isfinite(u.f[4][s])) return __builtin_huge_val();
  return cfl3d_cell(u, g, t);
}

}  // namespace mlau
#endif
