// mlau_2d_body.h -- device bodies of the 2D ideal-MHD + CT stage kernels.
//
// Each body does the work of one logical thread block; the single-grid
// kernels (mlau_2d.hip) and the batched ensemble kernels (mlau_ens2d.hip)
// launch the same bodies, so a grid gives identical bits alone or in a batch.
// Reference: 2D/SERIAL/OTvortex/mhd_fd_ct_2d.c:165-639 (one RK stage).
#ifndef MLAU_2D_BODY_H
#define MLAU_2D_BODY_H
#include "mlau_dev_mixed.h"
#include "mlau_internal.h"

namespace mlau {

struct U2 {
  double *f[8];  // ro,mx,my,mz,en,bx,by,bz
};
struct W2 {
  double *vx, *vy, *vz, *pr, *m2, *cx, *cy;
  double *ctx, *dvx, *dvy;
  double *fx[7], *fcx;  // ro,mx,my,mz,by(upwind part),bz,en at x-faces
  double *fy[7], *fcy;  // ro,mx,my,mz,bx(upwind part),bz,en at y-faces
  double *ez;
};
struct G2 {
  int nx, ny, xoff, yoff;
  double gamma;
};

// ------------------------------------------------------------ region tiles
// The cell kernels tile a w x h region (origin i0, j0) in 256-cell pieces of
// the flattened (j, i) index, so a row's last block is not left nearly empty.
__device__ __forceinline__ bool region_cell(int w, int h, int i0, int j0, unsigned piece, unsigned tid, int &i,
                                            int &j)
{
  const long p = (long)piece * 256 + tid;
  if (p >= (long)w * h) return false;
  i = i0 + (int)(p % w);
  j = j0 + (int)(p / w);
  return true;
}
__host__ __device__ __forceinline__ unsigned region_pieces(int w, int h) { return (unsigned)(((long)w * h + 255) / 256); }

// ------------------------------------------------------------------- prim2d
// Primitive variables of any cell (i, j) of the grid, formed as the reference
// does (mhd_fd_ct_2d.c:172-223): face B to the centre (cx, cy; a ghost cell
// takes its boundary source's value times the boundary factor), v = m/rho,
// p = (gamma-1)(E - 1/2 rho v^2 - 1/2 B^2), m2 = v^2 / (gamma p / rho).
struct Prim2 {
  double ro, vx, vy, vz, pr, cx, cy, bz, v2, iro;
};
template <int ODR, bool CX = true, bool CY = true>
__device__ __forceinline__ Prim2 prim2d_point(const U2 &u, const G2 &g, const BcFlags &bc, int i, int j)
{
  const int nx = g.nx;
  const long s = (long)j * nx + i;
  Prim2 P;
  P.cx = 0.0;
  P.cy = 0.0;
  double sg;
  int si, sj;
  if (CX) {
    sg = 1.0;
    si = bc_src(i, nx, g.xoff, 0, bc.dnx[5], &sg);
    sj = bc_src(j, g.ny, g.yoff, 0, bc.dny[5], &sg);
    const double *bx = u.f[5] + (long)sj * nx + si;
    P.cx = sg * bc_centre<ODR>(bx[-1], bx[0], bx[1], bx[2]);
  }
  if (CY) {
    sg = 1.0;
    si = bc_src(i, nx, g.xoff, 0, bc.dnx[6], &sg);
    sj = bc_src(j, g.ny, g.yoff, 0, bc.dny[6], &sg);
    const double *by = u.f[6] + (long)sj * nx + si;
    P.cy = sg * bc_centre<ODR>(by[-nx], by[0], by[nx], by[2 * nx]);
  }
  P.ro = u.f[0][s];
  P.iro = 1.0 / P.ro;
  P.vx = u.f[1][s] * P.iro;
  P.vy = u.f[2][s] * P.iro;
  P.vz = u.f[3][s] * P.iro;
  P.bz = u.f[7][s];
  P.v2 = (P.vx * P.vx + P.vy * P.vy + P.vz * P.vz);
  return P;
}
// pressure and Mach^2 need cx and cy
__device__ __forceinline__ double prim2d_pr(const U2 &u, const G2 &g, const Prim2 &P, long s)
{
  const double b2 = (P.cx * P.cx + P.cy * P.cy + P.bz * P.bz);
  return (g.gamma - 1.0) * (u.f[4][s] - 0.5 * (P.ro * P.v2 + b2));
}
__device__ __forceinline__ double prim2d_m2(const G2 &g, const Prim2 &P) { return P.v2 / (g.gamma * P.pr * P.iro); }

// Any cell (i, j) of the grid.
template <int ODR>
__device__ __forceinline__ void prim2d_body(U2 u, W2 w, G2 g, BcFlags bc, int i, int j)
{
  const long s = (long)j * g.nx + i;
  Prim2 P = prim2d_point<ODR>(u, g, bc, i, j);
  P.pr = prim2d_pr(u, g, P, s);
  w.vx[s] = P.vx;
  w.vy[s] = P.vy;
  w.vz[s] = P.vz;
  w.pr[s] = P.pr;
  w.m2[s] = prim2d_m2(g, P);
  w.cx[s] = P.cx;
  w.cy[s] = P.cy;
}

// Shock sensors of cell (i, j) from the conserved state, as aux2d_body forms
// them from the primitive arrays (mhd_fd_ct_2d.c:251-271): the velocity of
// any cell q is u[1 or 2][q] * (1/rho[q]), the expression prim2d uses.
__device__ __forceinline__ double dvx2d_point(const U2 &u, const G2 &g, const BcFlags &bc, int i, int j)
{
  double sg = 1.0;
  const int si = bc_src(i, g.nx, g.xoff, 0, -bc.dnx[1], &sg);
  const int sj = bc_src(j, g.ny, g.yoff, 0, +bc.dny[1], &sg);
  const long q = (long)sj * g.nx + si;
  const double vm = u.f[1][q - 1] * (1.0 / u.f[0][q - 1]);
  const double v0 = u.f[1][q] * (1.0 / u.f[0][q]);
  const double vp = u.f[1][q + 1] * (1.0 / u.f[0][q + 1]);
  return sg * dmin((vp - v0), (v0 - vm));
}
__device__ __forceinline__ double dvy2d_point(const U2 &u, const G2 &g, const BcFlags &bc, int i, int j)
{
  double sg = 1.0;
  const int si = bc_src(i, g.nx, g.xoff, 0, +bc.dnx[2], &sg);
  const int sj = bc_src(j, g.ny, g.yoff, 0, -bc.dny[2], &sg);
  const long q = (long)sj * g.nx + si, N = g.nx;
  const double vm = u.f[2][q - N] * (1.0 / u.f[0][q - N]);
  const double v0 = u.f[2][q] * (1.0 / u.f[0][q]);
  const double vp = u.f[2][q + N] * (1.0 / u.f[0][q + N]);
  return sg * dmin((vp - v0), (v0 - vm));
}

// -------------------------------------------------------------------- aux2d
// Any cell (i, j) of the grid.
template <int CTW>
__device__ __forceinline__ void aux2d_body(U2 u, W2 w, G2 g, BcFlags bc, int i, int j)
{
  const int nx = g.nx;
  const long s = (long)j * nx + i;
  double ctx = 0.5;
  if (CTW && i >= 1 && j >= 1) {
    const long a = s - nx, c = s;
    double vxc = 0.25 * (w.vx[a - 1] + w.vx[a] + w.vx[c - 1] + w.vx[c]);
    double vyc = 0.25 * (w.vy[a - 1] + w.vy[a] + w.vy[c - 1] + w.vy[c]);
    double roc = 0.25 * (u.f[0][a - 1] + u.f[0][a] + u.f[0][c - 1] + u.f[0][c]);
    roc = 1.0 / sqrt(roc);
    const double bxc = 0.5 * (u.f[5][a] + u.f[5][c]);
    const double byc = 0.5 * (u.f[6][c - 1] + u.f[6][c]);
    vxc = fabs(vxc) + fabs(bxc * roc);
    vyc = fabs(vyc) + fabs(byc * roc);
    const double eps = 1e-6;
    ctx = (vxc + 0.5 * eps) / (vxc + vyc + eps);
  }
  w.ctx[s] = ctx;
  double sg = 1.0;
  int si = bc_src(i, nx, g.xoff, 0, -bc.dnx[1], &sg);
  int sj = bc_src(j, g.ny, g.yoff, 0, +bc.dny[1], &sg);
  const double *vx = w.vx + (long)sj * nx + si;
  w.dvx[s] = sg * dmin((vx[1] - vx[0]), (vx[0] - vx[-1]));
  sg = 1.0;
  si = bc_src(i, nx, g.xoff, 0, +bc.dnx[2], &sg);
  sj = bc_src(j, g.ny, g.yoff, 0, -bc.dny[2], &sg);
  const double *vy = w.vy + (long)sj * nx + si;
  w.dvy[s] = sg * dmin((vy[nx] - vy[0]), (vy[0] - vy[-nx]));
}

// ---------------------------------------------------------------- primaux2d
// prim2d_body and aux2d_body in one pass over the cells: the velocities aux
// needs at neighbouring cells are formed from the conserved state with the
// same expressions prim uses (u[1 or 2] * (1/rho)), so the outputs are the
// bits the two kernels give; the primitive arrays are not read back.
__device__ __forceinline__ double vel2d(const U2 &u, long q, int comp)
{
  const double iro = 1.0 / u.f[0][q];
  return u.f[comp][q] * iro;
}

template <int ODR, int CTW>
__device__ __forceinline__ void primaux2d_body(U2 u, W2 w, G2 g, const BcFlags &bc, int i, int j)
{
  const int nx = g.nx;
  const long s = (long)j * nx + i;
  double sg = 1.0;
  int si = bc_src(i, nx, g.xoff, 0, bc.dnx[5], &sg);
  int sj = bc_src(j, g.ny, g.yoff, 0, bc.dny[5], &sg);
  const double *bxp = u.f[5] + (long)sj * nx + si;
  const double cx = sg * bc_centre<ODR>(bxp[-1], bxp[0], bxp[1], bxp[2]);
  sg = 1.0;
  si = bc_src(i, nx, g.xoff, 0, bc.dnx[6], &sg);
  sj = bc_src(j, g.ny, g.yoff, 0, bc.dny[6], &sg);
  const double *byp = u.f[6] + (long)sj * nx + si;
  const double cy = sg * bc_centre<ODR>(byp[-nx], byp[0], byp[nx], byp[2 * nx]);
  const double ro = u.f[0][s];
  const double iro = 1.0 / ro;
  const double vx = u.f[1][s] * iro;
  const double vy = u.f[2][s] * iro;
  const double vz = u.f[3][s] * iro;
  const double bz = u.f[7][s];
  const double v2 = (vx * vx + vy * vy + vz * vz);
  const double b2 = (cx * cx + cy * cy + bz * bz);
  const double pr = (g.gamma - 1.0) * (u.f[4][s] - 0.5 * (ro * v2 + b2));
  w.vx[s] = vx;
  w.vy[s] = vy;
  w.vz[s] = vz;
  w.pr[s] = pr;
  w.m2[s] = v2 / (g.gamma * pr * iro);
  w.cx[s] = cx;
  w.cy[s] = cy;
  // aux (mhd_fd_ct_2d.c:224-271)
  double ctx = 0.5;
  if (CTW && i >= 1 && j >= 1) {
    const long a = s - nx, c = s;
    double vxc = 0.25 * (vel2d(u, a - 1, 1) + vel2d(u, a, 1) + vel2d(u, c - 1, 1) + vx);
    double vyc = 0.25 * (vel2d(u, a - 1, 2) + vel2d(u, a, 2) + vel2d(u, c - 1, 2) + vy);
    double roc = 0.25 * (u.f[0][a - 1] + u.f[0][a] + u.f[0][c - 1] + u.f[0][c]);
    roc = 1.0 / sqrt(roc);
    const double bxc = 0.5 * (u.f[5][a] + u.f[5][c]);
    const double byc = 0.5 * (u.f[6][c - 1] + u.f[6][c]);
    vxc = fabs(vxc) + fabs(bxc * roc);
    vyc = fabs(vyc) + fabs(byc * roc);
    const double eps = 1e-6;
    ctx = (vxc + 0.5 * eps) / (vxc + vyc + eps);
  }
  w.ctx[s] = ctx;
  sg = 1.0;
  si = bc_src(i, nx, g.xoff, 0, -bc.dnx[1], &sg);
  sj = bc_src(j, g.ny, g.yoff, 0, +bc.dny[1], &sg);
  long q = (long)sj * nx + si;
  w.dvx[s] = sg * dmin((vel2d(u, q + 1, 1) - vel2d(u, q, 1)), (vel2d(u, q, 1) - vel2d(u, q - 1, 1)));
  sg = 1.0;
  si = bc_src(i, nx, g.xoff, 0, +bc.dnx[2], &sg);
  sj = bc_src(j, g.ny, g.yoff, 0, -bc.dny[2], &sg);
  q = (long)sj * nx + si;
  w.dvy[s] = sg * dmin((vel2d(u, q + nx, 2) - vel2d(u, q, 2)), (vel2d(u, q, 2) - vel2d(u, q - nx, 2)));
}

// ----------------------------------------------------------------- xsweep2d
// Cells 2..nx-3 of every row, flattened (q = j*(nx-4) + c-2), are covered by
// wavefronts of 64 lanes at a stride of 63 (neighbouring wavefronts share one
// cell): lane l >= 1 owns face c (between cells c-1 and c) unless c is the
// first cell of its row, so faces 3..nx-3 of every row come out exactly once
// and no lanes idle at row ends.  xsweep_waves() gives the wavefront count.
template <int ODR, int FLUX, int MIX>
__device__ __forceinline__ void xsweep2d_body(U2 u, W2 w, G2 g, long wave, int lane)
{
  const int nx = g.nx, rl = nx - 4;
  const long total = (long)g.ny * rl;
  long q = 63 * wave + lane;
  const bool live = q < total;
  if (!live) q = total - 1;
  const int c = 2 + (int)(q % rl);
  const long s = (q / rl) * nx + c;
  Stencil S;
#pragma unroll
  for (int m = 0; m < 5; m++) {
    const long q = s - 2 + m;
    S.q[0][m] = u.f[0][q];
    S.q[1][m] = w.vx[q];
    S.q[2][m] = w.vy[q];
    S.q[3][m] = w.vz[q];
    S.q[4][m] = w.cy[q];
    S.q[5][m] = u.f[7][q];
    S.q[6][m] = w.pr[q];
  }
  const bool sup = dmax(w.m2[s - 1], dmax(w.m2[s], w.m2[s + 1])) >= 1.0;
  double vl[7], vr[7];
  recon_p<ODR, MIX>(S, sup, w.cx[s], g.gamma, vl, vr);
  // central part of the By flux, linearly interpolated (:325-330)
  double p[5];
#pragma unroll
  for (int m = 0; m < 5; m++) p[m] = w.cy[s - 2 + m] * S.q[1][m];
  const double a0 = lcal_l<ODR>(p[0], p[1], p[2], p[3], p[4]);
  const double b0 = lcal_r<ODR>(p[0], p[1], p[2], p[3], p[4]);
  const double a1 = lcal_l<ODR>(S.q[2][0], S.q[2][1], S.q[2][2], S.q[2][3], S.q[2][4]);
  const double b1 = lcal_r<ODR>(S.q[2][0], S.q[2][1], S.q[2][2], S.q[2][3], S.q[2][4]);
  const double bxr = u.f[5][s];
  const double ql = a0 - u.f[5][s + 1] * a1;  // face c+1
  const double qr = b0 - bxr * b1;            // face c
  double L[7];
#pragma unroll
  for (int k = 0; k < 7; k++) L[k] = shfl_up1(vl[k]);
  const double qlL = shfl_up1(ql);
  if (lane == 0 || !live || c == 2) return;
  const double bn = 0.5 * (bxr + bxr);
  const double dv0 = w.vx[s] - w.vx[s - 1];
  const double dv1 = dmin(w.dvy[s - 1], w.dvy[s]);
  double f[7];
  riemann_p<FLUX, MIX>(L, vr, bn, g.gamma, dv0, dv1, f);
  const double fc = 0.5 * (qlL + qr);
  w.fx[0][s] = f[0];
  w.fx[1][s] = f[1];
  w.fx[2][s] = f[2];
  w.fx[3][s] = f[3];
  w.fx[4][s] = f[4] - fc;
  w.fx[5][s] = f[5];
  w.fx[6][s] = f[6];
  w.fcx[s] = fc;
}

// ----------------------------------------------------------------- ysweep2d
// One lane per column i; each lane walks faces f0..f0+CH-1 (rows), carrying
// the left state of the face below from the previous iteration.
// A 64-thread block (bx, by) holds columns 64*bx.. and faces 3+CH*by...
template <int ODR, int FLUX, int CH, int MIX>
__device__ __forceinline__ void ysweep2d_body(U2 u, W2 w, G2 g, unsigned bx, unsigned by, unsigned tid)
{
  const int i = bx * 64 + tid;
  const int nx = g.nx, ny = g.ny;
  if (i >= nx) return;
  const int f0 = 3 + CH * by;
  if (f0 > ny - 3) return;
  const int f1 = (f0 + CH - 1 < ny - 3) ? f0 + CH - 1 : ny - 3;
  // window rows r-2..r+2 of the rotated primitives (ro,vy,vz,vx,bz,cx,pr), m2
  double W[7][5], M[5];
  auto load = [&](int r, int m) {
    const long q = (long)r * nx + i;
    W[0][m] = u.f[0][q];
    W[1][m] = w.vy[q];
    W[2][m] = w.vz[q];
    W[3][m] = w.vx[q];
    W[4][m] = u.f[7][q];
    W[5][m] = w.cx[q];
    W[6][m] = w.pr[q];
    M[m] = w.m2[q];
  };
  auto cell = [&](int r, double (&vl)[7], double (&vr)[7], double &ql, double &qr) {
    Stencil S;
#pragma unroll
    for (int k = 0; k < 7; k++)
#pragma unroll
      for (int m = 0; m < 5; m++) S.q[k][m] = W[k][m];
    const bool sup = dmax(M[1], dmax(M[2], M[3])) >= 1.0;
    const long q = (long)r * nx + i;
    recon_p<ODR, MIX>(S, sup, w.cy[q], g.gamma, vl, vr);
    double p[5];
#pragma unroll
    for (int m = 0; m < 5; m++) p[m] = W[5][m] * W[1][m];
    const double a0 = lcal_l<ODR>(p[0], p[1], p[2], p[3], p[4]);
    const double b0 = lcal_r<ODR>(p[0], p[1], p[2], p[3], p[4]);
    const double a1 = lcal_l<ODR>(W[3][0], W[3][1], W[3][2], W[3][3], W[3][4]);
    const double b1 = lcal_r<ODR>(W[3][0], W[3][1], W[3][2], W[3][3], W[3][4]);
    ql = a0 - u.f[6][q + nx] * a1;
    qr = b0 - u.f[6][q] * b1;
  };
#pragma unroll
  for (int m = 0; m < 5; m++) load(f0 - 3 + m, m);
  double vlp[7], vr[7], qlp, qr;
  cell(f0 - 1, vlp, vr, qlp, qr);
  for (int r = f0; r <= f1; r++) {
#pragma unroll
    for (int m = 0; m < 4; m++) {
#pragma unroll
      for (int k = 0; k < 7; k++) W[k][m] = W[k][m + 1];
      M[m] = M[m + 1];
    }
    load(r + 2, 4);
    double vl[7], ql;
    cell(r, vl, vr, ql, qr);
    const long q = (long)r * nx + i;
    const double byf = u.f[6][q];
    const double bn = 0.5 * (byf + byf);
    const double dv0 = w.vy[q] - w.vy[q - nx];
    const double dv1 = dmin(w.dvx[q - nx], w.dvx[q]);
    double f[7];
    riemann_p<FLUX, MIX>(vlp, vr, bn, g.gamma, dv0, dv1, f);
    const double fc = 0.5 * (qlp + qr);
    w.fy[0][q] = f[0];
    w.fy[1][q] = f[3];
    w.fy[2][q] = f[1];
    w.fy[3][q] = f[2];
    w.fy[4][q] = f[5] - fc;
    w.fy[5][q] = f[4];
    w.fy[6][q] = f[6];
    w.fcy[q] = fc;
#pragma unroll
    for (int k = 0; k < 7; k++) vlp[k] = vl[k];
    qlp = ql;
  }
}

// --------------------------------------------------------------------- ez2d
// Corner (i, j) with 3 <= i <= nx-3, 3 <= j <= ny-3.
template <int ODR>
__device__ __forceinline__ void ez2d_body(W2 w, G2 g, int i, int j)
{
  const int nx = g.nx;
  const long s = (long)j * nx + i;
  const double ctx = w.ctx[s];
  // By upwind/central parts along Y at column i (:364-402)
  const double *u4 = w.fx[4] + s, *uc = w.fcx + s;
  const long N = nx;
  double a0 = lcal_l<ODR>(u4[-3 * N], u4[-2 * N], u4[-N], u4[0], u4[N]);
  double b0 = lcal_r<ODR>(u4[-2 * N], u4[-N], u4[0], u4[N], u4[2 * N]);
  double a1 = lcal_l<ODR>(uc[-3 * N], uc[-2 * N], uc[-N], uc[0], uc[N]);
  double b1 = lcal_r<ODR>(uc[-2 * N], uc[-N], uc[0], uc[N], uc[2 * N]);
  double ez = 0.0;
  ez += -0.5 * ((a0 + b0) + (1.0 - ctx) * (a1 + b1));
  // Bx upwind/central parts along X at row j (:495-533)
  const double *v4 = w.fy[4] + s, *vc = w.fcy + s;
  a0 = lcal_l<ODR>(v4[-3], v4[-2], v4[-1], v4[0], v4[1]);
  b0 = lcal_r<ODR>(v4[-2], v4[-1], v4[0], v4[1], v4[2]);
  a1 = lcal_l<ODR>(vc[-3], vc[-2], vc[-1], vc[0], vc[1]);
  b1 = lcal_r<ODR>(vc[-2], vc[-1], vc[0], vc[1], vc[2]);
  ez += +0.5 * ((a0 + b0) + ctx * (a1 + b1));
  w.ez[s] = ez;
}

// ----------------------------------------------------------------- update2d
// Cell (i, j) with xoff <= i <= nx-xoff, yoff <= j <= ny-yoff (one extra face
// column/row for the CT update).  PER (periodic grid whose interior spans at
// least two ghost widths per axis; interior cells only): the cell also writes
// its periodic images into the ghost frame, the face B at nx-xoff and ny-yoff
// included, as the boundary fill would (the update reads the stage input only
// at its own cell, so no ghost is read while they are written).
template <int ODR, bool FIRST, bool PER = false>
__device__ __forceinline__ void update2d_body(U2 in, U2 ut, U2 out, W2 w, G2 g,
                                              double dtdx, double dtdy, double ra, double rb, int i, int j)
{
  const int nx = g.nx, ny = g.ny;
  const long s = (long)j * nx + i;
  const long N = nx;
  if constexpr (PER) {
    const int xo = g.xoff, yo = g.yoff, NX = nx - 2 * xo, NY = ny - 2 * yo;
    const int gi = (i < 2 * xo) ? i + NX : (i >= nx - 2 * xo ? i - NX : -1);
    const int gj = (j < 2 * yo) ? j + NY : (j >= ny - 2 * yo ? j - NY : -1);
    auto put = [&](double *f, double v) {
      f[s] = v;
      if (gi >= 0) f[(long)j * nx + gi] = v;
      if (gj >= 0) f[(long)gj * nx + i] = v;
      if (gi >= 0 && gj >= 0) f[(long)gj * nx + gi] = v;
    };
    const int fld[6] = {0, 1, 2, 3, 7, 4};
    const int cmp[6] = {0, 1, 2, 3, 5, 6};
    for (int t = 0; t < 6; t++) {
      const double *a = w.fx[cmp[t]] + s;
      const double *b = w.fy[cmp[t]] + s;
      const double d0 = diff4<ODR>(a[-1], a[0], a[1], a[2]);
      const double d1 = diff4<ODR>(b[-N], b[0], b[N], b[2 * N]);
      const double upd = rb * (in.f[fld[t]][s] - dtdx * d0 - dtdy * d1);
      put(out.f[fld[t]], FIRST ? upd : ra * ut.f[fld[t]][s] + upd);
    }
    const double *ez = w.ez + s;
    double du = diff4<ODR>(ez[-N], ez[0], ez[N], ez[2 * N]);
    double upd = rb * (in.f[5][s] - dtdy * du);
    put(out.f[5], FIRST ? upd : ra * ut.f[5][s] + upd);
    du = diff4<ODR>(ez[-1], ez[0], ez[1], ez[2]);
    upd = rb * (in.f[6][s] + dtdx * du);
    put(out.f[6], FIRST ? upd : ra * ut.f[6][s] + upd);
    return;
  }
  if (i < nx - g.xoff && j < ny - g.yoff) {
    // field index in U2 for ro,mx,my,mz,bz,en and flux component in fx/fy
    const int fld[6] = {0, 1, 2, 3, 7, 4};
    const int cmp[6] = {0, 1, 2, 3, 5, 6};
#pragma unroll
    for (int t = 0; t < 6; t++) {
      const double *a = w.fx[cmp[t]] + s;
      const double *b = w.fy[cmp[t]] + s;
      const double d0 = diff4<ODR>(a[-1], a[0], a[1], a[2]);
      const double d1 = diff4<ODR>(b[-N], b[0], b[N], b[2 * N]);
      const double upd = rb * (in.f[fld[t]][s] - dtdx * d0 - dtdy * d1);
      out.f[fld[t]][s] = FIRST ? upd : ra * ut.f[fld[t]][s] + upd;
    }
  }
  const double *ez = w.ez + s;
  if (j < ny - g.yoff) {  // CT Bx on face i-1/2 (:593-612)
    const double du = diff4<ODR>(ez[-N], ez[0], ez[N], ez[2 * N]);
    const double upd = rb * (in.f[5][s] - dtdy * du);
    out.f[5][s] = FIRST ? upd : ra * ut.f[5][s] + upd;
  }
  if (i < nx - g.xoff) {  // CT By on face j-1/2 (:613-632)
    const double du = diff4<ODR>(ez[-1], ez[0], ez[1], ez[2]);
    const double upd = rb * (in.f[6][s] + dtdx * du);
    out.f[6][s] = FIRST ? upd : ra * ut.f[6][s] + upd;
  }
}

// --------------------------------------------------------------------- bc2d
// Ghost frame of field fld, frame point t: first the 2*xoff ghost columns of
// every row, then the 2*yoff ghost rows.
__device__ __forceinline__ long bc2d_points(G2 g) { return (long)g.ny * 2 * g.xoff + (long)2 * g.yoff * g.nx; }
__device__ __forceinline__ void bc2d_body(U2 u, G2 g, const BcFlags &bc, int fld, long t)
{
  const int nx = g.nx, ny = g.ny, xo = g.xoff, yo = g.yoff;
  const long nxs = (long)ny * 2 * xo;
  int i, j;
  if (t < nxs) {
    j = (int)(t / (2 * xo));
    const int k = (int)(t % (2 * xo));
    i = (k < xo) ? k : nx - 2 * xo + k;
  } else {
    const long r = t - nxs;
    if (r >= (long)2 * yo * nx) return;
    const int k = (int)(r / nx);
    i = (int)(r % nx);
    j = (k < yo) ? k : ny - 2 * yo + k;
  }
  double sg = 1.0;
  const int si = bc_src(i, nx, xo, bc.stx[fld], bc.dnx[fld], &sg);
  const int sj = bc_src(j, ny, yo, bc.sty[fld], bc.dny[fld], &sg);
  if (si == i && sj == j) return;  // a face the solver itself updated
  double *f = u.f[fld];
  f[(long)j * nx + i] = sg * f[(long)sj * nx + si];
}

// -------------------------------------------------------------------- cfl2d
// Signal speed |v_perp|+c_f of interior cell t (row-major over the interior)
// as 2D/SERIAL/OTvortex/cflcheck.hpp:1-32 computes it.
__device__ __forceinline__ double cfl2d_cell(U2 u, G2 g, long t)
{
  const int ni = g.nx - 2 * g.xoff;
  const int nx = g.nx;
  const int i = g.xoff + (int)(t % ni), j = g.yoff + (int)(t / ni);
  const long s = (long)j * nx + i;
  const double iro = 1.0 / u.f[0][s];
  const double vx = u.f[1][s] * iro, vy = u.f[2][s] * iro, vz = u.f[3][s] * iro;
  const double bxc = 0.5 * (u.f[5][s] + u.f[5][s + 1]);
  const double byc = 0.5 * (u.f[6][s] + u.f[6][s + nx]);
  const double bzc = u.f[7][s];
  const double v2 = vx * vx + vy * vy + vz * vz;
  const double b2 = bxc * bxc + byc * byc + bzc * bzc;
  const double pr = (g.gamma - 1.0) * (u.f[4][s] - 0.5 * (u.f[0][s] * v2 + b2));
  const double cf = sqrt((0.5 * g.gamma * (2.0 * pr) + b2) / (u.f[0][s]));
  return sqrt(vx * vx + vy * vy) + cf;
}

// Ensemble form: +Inf when the cell's density or energy is not finite (the
// member then fails), otherwise the speed of cfl2d_cell.
__device__ __forceinline__ double cfl2d_cell_checked(U2 u, G2 g, long t)
{
  const int ni = g.nx - 2 * g.xoff;
  const long s = (long)(g.yoff + (int)(t / ni)) * g.nx + g.xoff + (int)(t % ni);
  if (!isfinite(u.f[0][s]) || !isfinite(u.f[4][s])) return __builtin_huge_val();
  return cfl2d_cell(u, g, t);
}

}  // namespace mlau
#endif
