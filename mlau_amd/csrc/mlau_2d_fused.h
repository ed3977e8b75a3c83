// mlau_2d_fused.h -- the 2D stage of periodic grids without primitive arrays:
//
//   xsweep2d_inl_body  X reconstruction + Riemann solver, the primitives and
//                      shock sensors of the stencil formed from the conserved
//                      state in the sweep
//   ysweep2d_inl_body  Y reconstruction + Riemann solver, likewise, and the
//                      CT weights of the corners on its faces
//
// followed by the corner E-field (ez2d_body) and the periodic update
// (update2d_body with PER), so that neither primitives nor shock sensors make
// a round trip through HBM and the two cell passes (prim2d, aux2d) disappear.  Every value is formed by the
// expressions of the separate kernels (mlau_2d_body.h) from the same inputs,
// so the results are the same bits.  Reference:
// 2D/SERIAL/OTvortex/mhd_fd_ct_2d.c:165-639 (one RK stage).
#ifndef MLAU_2D_FUSED_H
#define MLAU_2D_FUSED_H
#include "mlau_2d_body.h"

namespace mlau {

// ------------------------------------------------------------ xsweep2d_inl
// As xsweep2d_body (cells 2..nx-3 of every row, flattened, 63 faces per
// wavefront), the primitives of the five stencil cells formed here
// (prim2d_point); the shock sensor dvy of cell c here and of cell c-1 from
// the lane below.  Writes fx[0..6] and fcx only.  Periodic grids.
template <int ODR, int FLUX, int MIX>
__device__ __forceinline__ void xsweep2d_inl_body(const U2 &u, const W2 &w, const G2 &g, long wave, int lane)
{
  constexpr BcFlags bc{};  // periodic grids only: every boundary flag 0
  const int nx = g.nx, rl = nx - 4;
  const long total = (long)g.ny * rl;
  long q = 63 * wave + lane;
  const bool live = q < total;
  if (!live) q = total - 1;
  const int c = 2 + (int)(q % rl);
  const int j = (int)(q / rl);
  const long s = (long)j * nx + c;
  Stencil S;
  double m2[3], cxs = 0.0;
#pragma unroll
  for (int m = 0; m < 5; m++) {
    Prim2 P = prim2d_point<ODR>(u, g, bc, c - 2 + m, j);
    P.pr = prim2d_pr(u, g, P, s - 2 + m);
    S.q[0][m] = P.ro;
    S.q[1][m] = P.vx;
    S.q[2][m] = P.vy;
    S.q[3][m] = P.vz;
    S.q[4][m] = P.cy;
    S.q[5][m] = P.bz;
    S.q[6][m] = P.pr;
    if (m >= 1 && m <= 3) m2[m - 1] = prim2d_m2(g, P);
    if (m == 2) cxs = P.cx;
  }
  const bool sup = dmax(m2[0], dmax(m2[1], m2[2])) >= 1.0;
  double vl[7], vr[7];
  recon_p<ODR, MIX>(S, sup, cxs, g.gamma, vl, vr);
  double p[5];
#pragma unroll
  for (int m = 0; m < 5; m++) p[m] = S.q[4][m] * S.q[1][m];
  const double a0 = lcal_l<ODR>(p[0], p[1], p[2], p[3], p[4]);
  const double b0 = lcal_r<ODR>(p[0], p[1], p[2], p[3], p[4]);
  const double a1 = lcal_l<ODR>(S.q[2][0], S.q[2][1], S.q[2][2], S.q[2][3], S.q[2][4]);
  const double b1 = lcal_r<ODR>(S.q[2][0], S.q[2][1], S.q[2][2], S.q[2][3], S.q[2][4]);
  const double bxr = u.f[5][s];
  const double ql = a0 - u.f[5][s + 1] * a1;  // face c+1
  const double qr = b0 - bxr * b1;            // face c
  const double dvy = dvy2d_point(u, g, bc, c, j);
  double L[7];
#pragma unroll
  for (int k = 0; k < 7; k++) L[k] = shfl_up1(vl[k]);
  const double qlL = shfl_up1(ql);
  const double dvyL = shfl_up1(dvy);
  if (lane == 0 || !live || c == 2) return;
  const double bn = 0.5 * (bxr + bxr);
  const double dv0 = S.q[1][2] - S.q[1][1];
  const double dv1 = dmin(dvyL, dvy);
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

// ------------------------------------------------------------ ysweep2d_inl
// As ysweep2d_body (one lane per column i marching the y-faces f0..f1 of its
// chunk, the 5-row window of rotated primitives in registers), the window
// rows formed from the conserved state here (prim2d_point) and the shock
// sensor dvx of rows f-1, f from the neighbouring lanes' velocities (wave
// shuffles; the wavefront's two outermost lanes and ghost cells form it from
// the state directly), and, with CTW, the CT weight of corner (i, f) (the
// corners' rows are the faces').  Writes fy[0..6], fcy and ctx.  Periodic
// grids.
template <int ODR, int FLUX, int CH, int MIX, int CTW>
__device__ __forceinline__ void ysweep2d_inl_body(const U2 &u, const W2 &w, const G2 &g, unsigned bx, unsigned by,
                                                  unsigned lane)
{
  constexpr BcFlags bc{};  // periodic grids only: every boundary flag 0
  const int nx = g.nx, ny = g.ny;
  const long N = nx;
  const int f0 = 3 + CH * by;
  if (f0 > ny - 3) return;  // whole wavefront
  const int f1 = (f0 + CH - 1 < ny - 3) ? f0 + CH - 1 : ny - 3;
  const int i = bx * 64 + lane;
  const bool live = i < nx;
  const int ic = live ? i : nx - 1;  // lanes past the grid take part in the shuffles only
  // window rows r-2..r+2 of cell r: rotated primitives (ro,vy,vz,vx,bz,cx,pr);
  // bit m of supb: m2 >= 1 in window row m (the reconstruction selector takes
  // the maximum of rows r-1..r+1; for finite values that is the same test)
  double W[7][5];
  unsigned supb = 0;
  auto load = [&](int R, int m) {
    Prim2 P = prim2d_point<ODR>(u, g, bc, ic, R);
    P.pr = prim2d_pr(u, g, P, (long)R * N + ic);
    W[0][m] = P.ro;
    W[1][m] = P.vy;
    W[2][m] = P.vz;
    W[3][m] = P.vx;
    W[4][m] = P.bz;
    W[5][m] = P.cx;
    W[6][m] = P.pr;
    supb = (supb & ~(1u << m)) | ((prim2d_m2(g, P) >= 1.0 ? 1u : 0u) << m);
  };
  // shock sensor dvx of window row m (row R): aux2d_body's expression
  auto dvx = [&](int R, int m) {
    const double vl = __shfl_up(W[3][m], 1, 64), vr = __shfl_down(W[3][m], 1, 64);
    const bool inner = ic >= g.xoff && ic < nx - g.xoff && R >= g.yoff && R < ny - g.yoff;
    if (inner && lane >= 1 && lane < 63) return dmin((vr - W[3][m]), (W[3][m] - vl));
    return dvx2d_point(u, g, bc, ic, R);
  };
  auto cell = [&](int r, double (&vl)[7], double (&vr)[7], double &ql, double &qr) {
    Stencil S;
#pragma unroll
    for (int k = 0; k < 7; k++)
#pragma unroll
      for (int m = 0; m < 5; m++) S.q[k][m] = W[k][m];
    const bool sup = (supb & 14u) != 0;
    const long q = (long)r * N + ic;
    double sg = 1.0;  // cy of the cell (prim2d_point's centring of face By)
    const int si = bc_src(ic, nx, g.xoff, 0, 0, &sg);
    const int sj = bc_src(r, ny, g.yoff, 0, 0, &sg);
    const double *byp = u.f[6] + (long)sj * nx + si;
    const double cy = sg * bc_centre<ODR>(byp[-nx], byp[0], byp[nx], byp[2 * nx]);
    recon_p<ODR, MIX>(S, sup, cy, g.gamma, vl, vr);
    double p[5];
#pragma unroll
    for (int m = 0; m < 5; m++) p[m] = W[5][m] * W[1][m];
    const double a0 = lcal_l<ODR>(p[0], p[1], p[2], p[3], p[4]);
    const double b0 = lcal_r<ODR>(p[0], p[1], p[2], p[3], p[4]);
    const double a1 = lcal_l<ODR>(W[3][0], W[3][1], W[3][2], W[3][3], W[3][4]);
    const double b1 = lcal_r<ODR>(W[3][0], W[3][1], W[3][2], W[3][3], W[3][4]);
    ql = a0 - u.f[6][q + N] * a1;
    qr = b0 - u.f[6][q] * b1;
  };
#pragma unroll
  for (int m = 0; m < 5; m++) load(f0 - 3 + m, m);
  double vlp[7], vr[7], qlp, qr;
  cell(f0 - 1, vlp, vr, qlp, qr);
  double dvp = dvx(f0 - 1, 2);  // dvx of the row below the face (window row 1 after the shift)
  for (int f = f0; f <= f1; f++) {
#pragma unroll
    for (int m = 0; m < 4; m++) {
#pragma unroll
      for (int k = 0; k < 7; k++) W[k][m] = W[k][m + 1];
    }
    supb >>= 1;
    load(f + 2, 4);
    const double dvc = dvx(f, 2);
    const double dv1 = dmin(dvp, dvc);
    dvp = dvc;
    if (CTW) {  // CT weight of corner (i, f) (aux2d_body), from rows f-1, f of columns i-1, i
      double lvx1 = __shfl_up(W[3][1], 1, 64), lvy1 = __shfl_up(W[1][1], 1, 64), lro1 = __shfl_up(W[0][1], 1, 64);
      double lvx2 = __shfl_up(W[3][2], 1, 64), lvy2 = __shfl_up(W[1][2], 1, 64), lro2 = __shfl_up(W[0][2], 1, 64);
      const long q = (long)f * N + ic;
      if (lane == 0 && ic >= 1) {  // column i-1 belongs to another wavefront: from the state
        lvx1 = vel2d(u, q - N - 1, 1);
        lvy1 = vel2d(u, q - N - 1, 2);
        lro1 = u.f[0][q - N - 1];
        lvx2 = vel2d(u, q - 1, 1);
        lvy2 = vel2d(u, q - 1, 2);
        lro2 = u.f[0][q - 1];
      }
      if (live && i >= 3 && i <= nx - 3) {
        double vxc = 0.25 * (lvx1 + W[3][1] + lvx2 + W[3][2]);
        double vyc = 0.25 * (lvy1 + W[1][1] + lvy2 + W[1][2]);
        double roc = 0.25 * (lro1 + W[0][1] + lro2 + W[0][2]);
        roc = 1.0 / sqrt(roc);
        const double bxc = 0.5 * (u.f[5][q - N] + u.f[5][q]);
        const double byc = 0.5 * (u.f[6][q - 1] + u.f[6][q]);
        vxc = fabs(vxc) + fabs(bxc * roc);
        vyc = fabs(vyc) + fabs(byc * roc);
        const double eps = 1e-6;
        w.ctx[q] = (vxc + 0.5 * eps) / (vxc + vyc + eps);
      }
    }
    double vl[7], ql;
    cell(f, vl, vr, ql, qr);
    const long q = (long)f * N + ic;
    const double byf = u.f[6][q];
    const double bn = 0.5 * (byf + byf);
    const double dv0 = W[1][2] - W[1][1];
    double fl[7];
    riemann_p<FLUX, MIX>(vlp, vr, bn, g.gamma, dv0, dv1, fl);
    const double fc = 0.5 * (qlp + qr);
    if (live) {
      w.fy[0][q] = fl[0];
      w.fy[1][q] = fl[3];
      w.fy[2][q] = fl[1];
      w.fy[3][q] = fl[2];
      w.fy[4][q] = fl[5] - fc;
      w.fy[5][q] = fl[4];
      w.fy[6][q] = fl[6];
      w.fcy[q] = fc;
    }
#pragma unroll
    for (int k = 0; k < 7; k++) vlp[k] = vl[k];
    qlp = ql;
  }
}

}  // namespace mlau
#endif
