// mlau_1d_body.h -- device bodies of the 1D ideal-MHD stage kernels, shared
// by the single-grid kernels (mlau_1d.hip) and the batched ensemble kernels
// (mlau_ens1d.hip).  Reference: 1D/SERIAL/shock/mhd_fd4c_1d.c:110-257.
#ifndef MLAU_1D_BODY_H
#define MLAU_1D_BODY_H
#include "mlau_dev_mixed.h"
#include "mlau_internal.h"

namespace mlau {

struct U1d {
  double *f[7];
};
struct W1d {
  double *vx, *vy, *vz, *pr;
  double *fx[7];  // flux per field, argument order
};
struct G1d {
  int nx, xoff;
  double gamma, bx;
};

// A 256-thread block bx holds tiles 4*bx..4*bx+3 of 63 faces each.  The
// stencil's primitives vx,vy,vz,pr are formed here from the conserved fields
// as 1D/SERIAL/shock/mhd_fd4c_1d.c:121-131 forms them, so a stage needs no
// separate primitive pass.
template <int ODR, int FLUX, int MIX>
__device__ __forceinline__ void sweep1d_body(U1d u, W1d w, G1d g, unsigned bx, unsigned tid)
{
  const int lane = tid & 63;
  const int tile = bx * 4 + (tid >> 6);
  const int nx = g.nx;
  const int c0 = 2 + 63 * tile;
  if (c0 > nx - 4) return;
  const int c = c0 + lane;
  const int cc = (c < nx - 3) ? c : nx - 3;
  const double bx2 = g.bx * g.bx;
  Stencil S;
#pragma unroll
  for (int m = 0; m < 5; m++) {
    const int q = cc - 2 + m;
    const double ro = u.f[0][q];
    const double iro = 1.0 / ro;
    const double vx = u.f[1][q] * iro, vy = u.f[2][q] * iro, vz = u.f[3][q] * iro;
    const double by = u.f[5][q], bz = u.f[6][q];
    const double v2 = vx * vx + vy * vy + vz * vz;
    const double b2 = bx2 + by * by + bz * bz;
    S.q[0][m] = ro;
    S.q[1][m] = vx;
    S.q[2][m] = vy;
    S.q[3][m] = vz;
    S.q[4][m] = by;
    S.q[5][m] = bz;
    S.q[6][m] = (g.gamma - 1.0) * (u.f[4][q] - 0.5 * (ro * v2 + b2));
  }
  // supersonic switch of the 1D solver (mhd_fd4c_1d.c:155-166)
  double r[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double cs2 = g.gamma * S.q[6][k + 1] / S.q[0][k + 1];
    const double vv2 = S.q[1][k + 1] * S.q[1][k + 1] + S.q[2][k + 1] * S.q[2][k + 1] + S.q[3][k + 1] * S.q[3][k + 1];
    r[k] = vv2 / cs2;
  }
  const bool sup = dmax(r[0], dmax(r[1], r[2])) >= 1.0;
  double vl[7], vr[7];
  recon_p<ODR, MIX>(S, sup, g.bx, g.gamma, vl, vr);
  double L[7];
#pragma unroll
  for (int k = 0; k < 7; k++) L[k] = shfl_up1(vl[k]);
  if (lane == 0 || c > nx - 3) return;
  const double dv0 = S.q[1][2] - S.q[1][1];  // vx[c] - vx[c-1] (cc == c here)
  double f[7];
  riemann_p<FLUX, MIX>(L, vr, g.bx, g.gamma, dv0, 0.0, f);
  w.fx[0][c] = f[0];
  w.fx[1][c] = f[1];
  w.fx[2][c] = f[2];
  w.fx[3][c] = f[3];
  w.fx[4][c] = f[6];
  w.fx[5][c] = f[4];
  w.fx[6][c] = f[5];
}

// Full Athena eigensystem with the reference's complete 7x7 products
// (include/mhd_eigen.h:24-192), for the two boundary cells of the open BC.
static __device__ void eigen_st_full(const double p[8], double gamma, double al[7], double L[7][7], double R[7][7])
{
  const double ro = p[0], vx = p[1], bx = p[4], by = p[5], bz = p[6], pr = p[7];
  const double eps = 1e-12;
  const double sqrt2 = sqrt(2.0);
  double iro = 1.0 / ro;
  double sro = sqrt(ro);
  double isro = 1.0 / sro;
  double bxsr = bx * isro, bysr = by * isro, bzsr = bz * isro;
  double bx2 = bxsr * bxsr;
  double bt2 = bysr * bysr + bzsr * bzsr;
  double b2 = bx2 + bt2;
  double a2 = fabs(gamma * pr / ro);
  double aa = sqrt(a2);
  double ca2 = bx2;
  double ca = sqrt(ca2);
  double junk = sqrt(fabs((a2 + b2) * (a2 + b2) - 4.0 * a2 * ca2));
  double cf2 = fabs(0.5 * (a2 + b2 + junk));
  double cf = sqrt(cf2);
  double cs2 = fabs(0.5 * (a2 + b2 - junk));
  double cs = sqrt(cs2);
  bx2 = bx * bx;
  bt2 = by * by + bz * bz;
  b2 = bx2 + bt2;
  double den = 1.0 / sqrt(bt2);
  int flag = (bt2 > eps * b2);
  double betay = flag ? (by * den) : (0.5 * sqrt2);
  double betaz = flag ? (bz * den) : (0.5 * sqrt2);
  den = 1.0 / sqrt(fabs(cf2 - cs2));
  flag = ((bt2 > eps * b2) || (fabs(gamma * pr - bx2) > eps * gamma * pr));
  double alphaf = flag ? (sqrt(fabs(a2 - cs2)) * den) : 1.0;
  double alphas = flag ? (sqrt(fabs(cf2 - a2)) * den) : 0.0;
  double sgnbx = (bx > 0) ? (1.0) : (-1.0);
  double cfalf = cf * alphaf, csals = cs * alphas, aaalf = aa * alphaf, aaals = aa * alphas;
  den = 0.5 / a2;
  al[0] = vx - cf; al[1] = vx - ca; al[2] = vx - cs; al[3] = vx; al[4] = vx + cs; al[5] = vx + ca; al[6] = vx + cf;
  const double sp[7] = {+1.0, -1.0, -1.0, -1.0, +1.0, +1.0, +1.0};  // mirror pattern of l7/l6/l5, r7/r6/r5
  double l1[7] = {0, -cfalf * den, +csals * den * sgnbx * betay, +csals * den * sgnbx * betaz,
                  +aaals * den * isro * betay, +aaals * den * isro * betaz, alphaf * den * iro};
  double l2[7] = {0, 0, -0.5 * betaz, +0.5 * betay, -0.5 * isro * betaz * sgnbx, +0.5 * isro * betay * sgnbx, 0};
  double l3[7] = {0, -csals * den, -cfalf * den * sgnbx * betay, -cfalf * den * sgnbx * betaz,
                  -aaalf * den * isro * betay, -aaalf * den * isro * betaz, alphas * den * iro};
  double r1[7] = {ro * alphaf, -cfalf, +csals * sgnbx * betay, +csals * sgnbx * betaz, +aaals * sro * betay,
                  +aaals * sro * betaz, alphaf * gamma * pr};
  double r2[7] = {0, 0, -betaz, +betay, -sro * betaz * sgnbx, +sro * betay * sgnbx, 0};
  double r3[7] = {ro * alphas, -csals, -cfalf * sgnbx * betay, -cfalf * sgnbx * betaz, -aaalf * sro * betay,
                  -aaalf * sro * betaz, alphas * gamma * pr};
  for (int n = 0; n < 7; n++) {
    L[0][n] = l1[n]; L[1][n] = l2[n]; L[2][n] = l3[n];
    L[4][n] = (sp[n] > 0) ? +l3[n] : -l3[n];
    L[5][n] = (sp[n] > 0) ? +l2[n] : -l2[n];
    L[6][n] = (sp[n] > 0) ? +l1[n] : -l1[n];
    L[3][n] = 0;
    R[0][n] = r1[n]; R[1][n] = r2[n]; R[2][n] = r3[n];
    R[4][n] = (sp[n] > 0) ? +r3[n] : -r3[n];
    R[5][n] = (sp[n] > 0) ? +r2[n] : -r2[n];
    R[6][n] = (sp[n] > 0) ? +r1[n] : -r1[n];
    R[3][n] = 0;
  }
  L[3][0] = 1.0;
  L[3][6] = -1.0 / a2;
  R[3][0] = 1.0;
}

// openbc1d_mhd with istt=0, iend=1 (1D/SERIAL/shock/openbc1d_mhd.c:21-175);
// thread 0 = left edge, thread 1 = right edge.  Reads the stage input and
// writes the boundary ghosts of every buffer in `dst` (the stage output and
// the step-start state, so later stages and the final state both see them).
__device__ __forceinline__ void openbc1d_body(U1d in, U1d dst0, U1d dst1, G1d g, double dtdx, unsigned side)
{
  if (side > 1) return;
  const int nx = g.nx, xoff = g.xoff;
  const int sgn = side == 0 ? 1 : -1;
  const int si = side == 0 ? xoff : nx - xoff - 1;
  const int so = si - sgn;
  const double bx = g.bx, bx2 = bx * bx, gamma = g.gamma;
  double rob[2], vxb[2], vyb[2], vzb[2], bxb[2], byb[2], bzb[2], prb[2];
  const int idx[2] = {si, so};
  for (int j = 0; j < 2; j++) {
    const int ss = idx[j];
    rob[j] = in.f[0][ss];
    const double iro = 1.0 / rob[j];
    vxb[j] = in.f[1][ss] * iro;
    vyb[j] = in.f[2][ss] * iro;
    vzb[j] = in.f[3][ss] * iro;
    bxb[j] = bx;
    byb[j] = in.f[5][ss];
    bzb[j] = in.f[6][ss];
    const double v2 = vxb[j] * vxb[j] + vyb[j] * vyb[j] + vzb[j] * vzb[j];
    const double b2 = bxb[j] * bxb[j] + byb[j] * byb[j] + bzb[j] * bzb[j];
    prb[j] = (gamma - 1.0) * (in.f[4][ss] - 0.5 * (rob[j] * v2 + b2));
  }
  double p[8] = {0.5 * (rob[0] + rob[1]), 0.5 * (vxb[0] + vxb[1]), 0.5 * (vyb[0] + vyb[1]), 0.5 * (vzb[0] + vzb[1]),
                 0.5 * (bxb[0] + bxb[1]), 0.5 * (byb[0] + byb[1]), 0.5 * (bzb[0] + bzb[1]), 0.5 * (prb[0] + prb[1])};
  double al[7], L[7][7], R[7][7], w[7][2];
  eigen_st_full(p, gamma, al, L, R);
  for (int j = 0; j < 2; j++)
    for (int n = 0; n < 7; n++)
      w[n][j] = (+L[n][0] * rob[j] + L[n][1] * vxb[j] + L[n][2] * vyb[j] + L[n][3] * vzb[j] + L[n][4] * byb[j] +
                 L[n][5] * bzb[j] + L[n][6] * prb[j]);
  for (int n = 0; n < 7; n++) {
    const double sa = sgn * al[n];
    w[n][1] = (sa > 0) ? w[n][1] : (1 + sa * dtdx) * w[n][1] - sa * dtdx * w[n][0];
  }
  double q[7];
  for (int n = 0; n < 7; n++)
    q[n] = R[0][n] * w[0][1] + R[1][n] * w[1][1] + R[2][n] * w[2][1] + R[3][n] * w[3][1] + R[4][n] * w[4][1] +
           R[5][n] * w[5][1] + R[6][n] * w[6][1];
  // q = (ro, vx, vy, vz, by, bz, pr) -> conservative
  const double ro = q[0];
  double mx = q[1], my = q[2], mz = q[3];
  const double by = q[4], bz = q[5];
  double en = q[6];
  const double v2 = mx * mx + my * my + mz * mz;
  const double b2 = bx2 + by * by + bz * bz;
  en = en / (gamma - 1.0) + 0.5 * (ro * v2 + b2);
  mx *= ro;
  my *= ro;
  mz *= ro;
  const double vals[7] = {ro, mx, my, mz, en, by, bz};
  U1d *dsts[2] = {&dst0, &dst1};
  for (int d = 0; d < 2; d++) {
    for (int n = 0; n < 7; n++) {
      double *f = dsts[d]->f[n];
      f[so] = vals[n];
      for (int j = 0; j < xoff - 1; j++) f[side * (nx - 1) + sgn * j] = vals[n];
    }
  }
}

// Logical block b covers cells 4+256*b..  With PER, a cell that is the
// periodic image of a ghost cell also writes that ghost with the same value
// (prdc_1d, 1D/SERIAL/wave/prdc_1d.c:18-25: f[j] = f[nx-2xoff+j] and
// f[nx-1-j] = f[2xoff-1-j] for j < xoff), so a stage needs no ghost pass.
template <int ODR, bool PER>
__device__ __forceinline__ void update1d_body(U1d in, U1d ut, U1d out, W1d w, G1d g, double dtdx, double ra,
                                              double rb, unsigned b, unsigned tid)
{
  const int i = 4 + b * 256 + tid;
  const int nx = g.nx, xoff = g.xoff;
  if (i >= nx - 4) return;
#pragma unroll
  for (int n = 0; n < 7; n++) {
    const double *a = w.fx[n] + i;
    const double du = diff4<ODR>(a[-1], a[0], a[1], a[2]);
    const double v = ra * ut.f[n][i] + rb * (in.f[n][i] - dtdx * du);
    out.f[n][i] = v;
    if (PER) {
      if (i >= nx - 2 * xoff && i < nx - xoff) out.f[n][i - (nx - 2 * xoff)] = v;  // f[j] = f[nx-2xoff+j]
      if (i >= xoff && i < 2 * xoff) out.f[n][nx - 2 * xoff + i] = v;              // f[nx-1-j] = f[2xoff-1-j]
    }
  }
}

// One RK stage in one pass for ODR <= 2 (update from faces i and i+1 only):
// wavefront `wave` holds cells c0-1..c0+62 (c0 = 4 + 62*wave) one per lane,
// forms the face fluxes as sweep1d_body does (left state from the lane below)
// and updates cells c0..c0+61 as update1d_body does, taking the flux of face
// i+1 from the lane above.  The stage input is only read, so `out` must not
// be `in`.  With copy_ghosts the ghost cells of `in` are copied to `out` (an
// open boundary's ghosts stay fixed over the later stages of a step).
template <int ODR, int FLUX, int MIX, bool PER>
__device__ __forceinline__ void stage1d_body(U1d in, U1d ut, U1d out, G1d g, double dtdx, double ra, double rb,
                                             unsigned wave, unsigned lane, bool copy_ghosts)
{
  static_assert(ODR <= 2, "the fused 1D stage takes the update from two faces");
  const int nx = g.nx, xoff = g.xoff;
  const int c0 = 4 + 62 * (int)wave;
  if (c0 >= nx - 4) return;  // whole wavefront
  const int c = c0 - 1 + (int)lane;
  const int cc = (c < nx - 4) ? c : nx - 4;
  const double bx2 = g.bx * g.bx;
  Stencil S;
#pragma unroll
  for (int m = 0; m < 5; m++) {
    const int q = cc - 2 + m;
    const double ro = in.f[0][q];
    const double iro = 1.0 / ro;
    const double vx = in.f[1][q] * iro, vy = in.f[2][q] * iro, vz = in.f[3][q] * iro;
    const double by = in.f[5][q], bz = in.f[6][q];
    const double v2 = vx * vx + vy * vy + vz * vz;
    const double b2 = bx2 + by * by + bz * bz;
    S.q[0][m] = ro;
    S.q[1][m] = vx;
    S.q[2][m] = vy;
    S.q[3][m] = vz;
    S.q[4][m] = by;
    S.q[5][m] = bz;
    S.q[6][m] = (g.gamma - 1.0) * (in.f[4][q] - 0.5 * (ro * v2 + b2));
  }
  double r[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double cs2 = g.gamma * S.q[6][k + 1] / S.q[0][k + 1];
    const double vv2 = S.q[1][k + 1] * S.q[1][k + 1] + S.q[2][k + 1] * S.q[2][k + 1] + S.q[3][k + 1] * S.q[3][k + 1];
    r[k] = vv2 / cs2;
  }
  const bool sup = dmax(r[0], dmax(r[1], r[2])) >= 1.0;
  double vl[7], vr[7];
  recon_p<ODR, MIX>(S, sup, g.bx, g.gamma, vl, vr);
  double L[7];
#pragma unroll
  for (int k = 0; k < 7; k++) L[k] = shfl_up1(vl[k]);
  const double dv0 = S.q[1][2] - S.q[1][1];
  double f[7];
  riemann_p<FLUX, MIX>(L, vr, g.bx, g.gamma, dv0, 0.0, f);  // lane 0's face is not used
  // flux per field in argument order (ro,mx,my,mz,en,by,bz), then face c+1's
  const double fa[7] = {f[0], f[1], f[2], f[3], f[6], f[4], f[5]};
  double fb[7];
#pragma unroll
  for (int n = 0; n < 7; n++) fb[n] = __shfl_down(fa[n], 1, 64);
  if (copy_ghosts) {
    if (wave == 0 && lane < (unsigned)xoff)
      for (int n = 0; n < 7; n++) out.f[n][lane] = in.f[n][lane];
    if (c0 + 62 >= nx - 4 && lane < (unsigned)xoff)
      for (int n = 0; n < 7; n++) out.f[n][nx - 1 - lane] = in.f[n][nx - 1 - lane];
  }
  if (lane == 0 || lane > 62 || c >= nx - 4) return;
#pragma unroll
  for (int n = 0; n < 7; n++) {
    const double du = diff4<ODR>(0.0, fa[n], fb[n], 0.0);
    const double v = ra * ut.f[n][c] + rb * (in.f[n][c] - dtdx * du);
    out.f[n][c] = v;
    if (PER) {
      if (c >= nx - 2 * xoff && c < nx - xoff) out.f[n][c - (nx - 2 * xoff)] = v;  // f[j] = f[nx-2xoff+j]
      if (c >= xoff && c < 2 * xoff) out.f[n][nx - 2 * xoff + c] = v;              // f[nx-1-j] = f[2xoff-1-j]
    }
  }
}

// Signal speed |vx|+c_f of cell i for the driver's one-off dt estimate
// (1D/SERIAL/shock/main.cpp:114-128), taken over all cells.
__device__ __forceinline__ double cfl1d_cell(U1d u, G1d g, int i)
{
  const double iro = 1.0 / u.f[0][i];
  const double vx = u.f[1][i] * iro, vy = u.f[2][i] * iro, vz = u.f[3][i] * iro;
  const double v2 = vx * vx + vy * vy + vz * vz;
  const double b2 = g.bx * g.bx + u.f[5][i] * u.f[5][i] + u.f[6][i] * u.f[6][i];
  const double pr = (g.gamma - 1.0) * (u.f[4][i] - 0.5 * (u.f[0][i] * v2 + b2));
  const double cf = sqrt((g.gamma * pr + b2) * iro);
  return fabs(vx) + cf;
}

// Ensemble form: +Inf when the cell's density or energy is not finite (the
// member then fails), otherwise the speed of cfl1d_cell.
__device__ __forceinline__ double cfl1d_cell_checked(U1d u, G1d g, int i)
{
  if (!isfinite(u.f[0][i]) || !isfinite(u.f[4][i])) return __builtin_huge_val();
  return cfl1d_cell(u, g, i);
}

}  // namespace mlau
#endif
