// mlau_2d.hip -- 2D ideal-MHD + CT step on gfx950 (FP64).
//
// One RK stage (2D/SERIAL/OTvortex/mhd_fd_ct_2d.c:165-639) is six kernels:
//   k_prim2d    face B -> centre (cx,cy) and primitives (vx,vy,vz,pr,m2)   :172-223
//   k_aux2d     CT upwind weight ctx and shock sensors dvx,dvy            :224-271
//   k_xsweep2d  X reconstruction + Riemann solver, one wavefront per 63
//               faces of a row; each lane reconstructs one cell and hands
//               its left face state to the next lane by a wave shuffle    :273-362
//   k_ysweep2d  Y reconstruction + Riemann solver, one lane per column
//               marching up a chunk of rows with the 5-row stencil held in
//               registers (coalesced rows, no reconstruction twice)      :404-493
//   k_ez2d      corner E-field from both sweeps' induction fluxes       :364-402,495-533
//   k_update2d  cell-centre update, CT face-B update, SSP-RK combine    :535-632
// followed by k_bc2d (ghost cells, :636-638).  The reference's halo fills of
// the intermediate fields (bc2d on cx,cy,dvx,dvy) are folded into the
// kernels that produce them: a ghost value is computed at its source cell and
// multiplied by the boundary factor.
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

// ---------------------------------------------------------------- k_prim2d
template <int ODR>
__global__ __launch_bounds__(256) void k_prim2d(U2 u, W2 w, G2 g, BcFlags bc)
{
  const unsigned gx = (g.nx + 255) / 256;
  const unsigned b = xcd_band(blockIdx.x, gridDim.x);
  const int i = (b % gx) * 256 + threadIdx.x;
  const int j = b / gx;
  if (i >= g.nx) return;
  const int nx = g.nx;
  const long s = (long)j * nx + i;
  double sg = 1.0;
  int si = bc_src(i, nx, g.xoff, 0, bc.dnx[5], &sg);
  int sj = bc_src(j, g.ny, g.yoff, 0, bc.dny[5], &sg);
  const double *bx = u.f[5] + (long)sj * nx + si;
  const double cx = sg * bc_centre<ODR>(bx[-1], bx[0], bx[1], bx[2]);
  sg = 1.0;
  si = bc_src(i, nx, g.xoff, 0, bc.dnx[6], &sg);
  sj = bc_src(j, g.ny, g.yoff, 0, bc.dny[6], &sg);
  const double *by = u.f[6] + (long)sj * nx + si;
  const double cy = sg * bc_centre<ODR>(by[-nx], by[0], by[nx], by[2 * nx]);
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
}

// ----------------------------------------------------------------- k_aux2d
template <int CTW>
__global__ __launch_bounds__(256) void k_aux2d(U2 u, W2 w, G2 g, BcFlags bc)
{
  const unsigned gx = (g.nx + 255) / 256;
  const unsigned b = xcd_band(blockIdx.x, gridDim.x);
  const int i = (b % gx) * 256 + threadIdx.x;
  const int j = b / gx;
  if (i >= g.nx) return;
  const int nx = g.nx;
  const long s = (long)j * nx + i;
  double ctx = 0.5;
  if (CTW && i >= 1 && j >= 1) {
    const long a = s - nx, b = s;
    double vxc = 0.25 * (w.vx[a - 1] + w.vx[a] + w.vx[b - 1] + w.vx[b]);
    double vyc = 0.25 * (w.vy[a - 1] + w.vy[a] + w.vy[b - 1] + w.vy[b]);
    double roc = 0.25 * (u.f[0][a - 1] + u.f[0][a] + u.f[0][b - 1] + u.f[0][b]);
    roc = 1.0 / sqrt(roc);
    const double bxc = 0.5 * (u.f[5][a] + u.f[5][b]);
    const double byc = 0.5 * (u.f[6][b - 1] + u.f[6][b]);
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

// -------------------------------------------------------------- k_xsweep2d
// One wavefront covers cells c = 2+63*t+lane of row j; lane l>=1 owns face c
// (between cells c-1 and c).  Faces 3..nx-3 are produced.
template <int ODR, int FLUX, int WPE>
__global__ __launch_bounds__(256, WPE) void k_xsweep2d(U2 u, W2 w, G2 g)
{
  const int lane = threadIdx.x & 63;
  const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int j = blockIdx.y;
  const int nx = g.nx;
  const int c0 = 2 + 63 * tile;
  if (c0 > nx - 4) return;  // no face of this tile is in range (uniform per wave)
  const int c = c0 + lane;
  const int cc = (c < nx - 3) ? c : nx - 3;
  const long row = (long)j * nx;
  const long s = row + cc;
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
  recon<ODR>(S, sup, w.cx[s], g.gamma, vl, vr);
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
  if (lane == 0 || c > nx - 3) return;
  const double bn = 0.5 * (bxr + bxr);
  const double dv0 = w.vx[s] - w.vx[s - 1];
  const double dv1 = dmin(w.dvy[s - 1], w.dvy[s]);
  double f[7];
  riemann<FLUX>(L, vr, bn, g.gamma, dv0, dv1, f);
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

// -------------------------------------------------------------- k_ysweep2d
// One lane per column i; each lane walks faces f0..f0+CH-1 (rows), carrying
// the left state of the face below from the previous iteration.
template <int ODR, int FLUX, int CH, int WPE>
__global__ __launch_bounds__(64, WPE) void k_ysweep2d(U2 u, W2 w, G2 g)
{
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int nx = g.nx, ny = g.ny;
  if (i >= nx) return;
  const int f0 = 3 + CH * blockIdx.y;
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
    recon<ODR>(S, sup, w.cy[q], g.gamma, vl, vr);
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
    riemann<FLUX>(vlp, vr, bn, g.gamma, dv0, dv1, f);
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

// ------------------------------------------------------------------ k_ez2d
template <int ODR>
__global__ __launch_bounds__(256) void k_ez2d(W2 w, G2 g)
{
  const unsigned gx = (g.nx - 5 + 255) / 256;
  const unsigned b = xcd_band(blockIdx.x, gridDim.x);
  const int i = 3 + (b % gx) * 256 + threadIdx.x;
  const int j = 3 + b / gx;
  const int nx = g.nx;
  if (i > nx - 3) return;
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

// -------------------------------------------------------------- k_update2d
// Cells xoff..nx-xoff (one extra face column/row for the CT update).
template <int ODR, bool FIRST>
__global__ __launch_bounds__(256) void k_update2d(U2 in, U2 ut, U2 out, W2 w, G2 g, const StepInfo *info,
                                                  double ra, double rb)
{
  const unsigned gx = (g.nx - 2 * g.xoff + 1 + 255) / 256;
  const unsigned b = xcd_band(blockIdx.x, gridDim.x);
  const int i = g.xoff + (b % gx) * 256 + threadIdx.x;
  const int j = g.yoff + b / gx;
  const int nx = g.nx, ny = g.ny;
  if (i > nx - g.xoff) return;
  const long s = (long)j * nx + i;
  const double dtdx = info->dtdx, dtdy = info->dtdy;
  const long N = nx;
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

// ------------------------------------------------------------------ k_bc2d
// Ghost frame of each field (blockIdx.y = field): first the 2*xoff ghost
// columns of every row, then the 2*yoff ghost rows.
__global__ __launch_bounds__(256) void k_bc2d(U2 u, G2 g, BcFlags bc)
{
  const int fld = blockIdx.y;
  const int nx = g.nx, ny = g.ny, xo = g.xoff, yo = g.yoff;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
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

// ----------------------------------------------------------------- k_cfl2d
// vmax over the interior (2D/SERIAL/OTvortex/cflcheck.hpp:1-32), grid-stride
// with one atomic per block.
__global__ __launch_bounds__(256) void k_cfl2d(U2 u, G2 g, StepInfo *info)
{
  const int ni = g.nx - 2 * g.xoff, nj = g.ny - 2 * g.yoff;
  const long total = (long)ni * nj;
  const int nx = g.nx;
  double v = 0.0;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
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
    const double vt = sqrt(vx * vx + vy * vy) + cf;
    if (vt > v) v = vt;
  }
  block_max_atomic<256>(v, &info->vmax_bits);
}

// ------------------------------------------------------------------ driver
static void make_views(Handle *h, U2 &u0, U2 &u1, W2 &w, G2 &g)
{
  for (int n = 0; n < 8; n++) {
    u0.f[n] = h->U0[n];
    u1.f[n] = h->U1[n];
  }
  w.vx = h->prim[0]; w.vy = h->prim[1]; w.vz = h->prim[2]; w.pr = h->prim[3]; w.m2 = h->prim[4];
  w.cx = h->prim[5]; w.cy = h->prim[6];
  w.ctx = h->aux[0]; w.dvx = h->aux[3]; w.dvy = h->aux[4];
  for (int k = 0; k < 7; k++) {
    w.fx[k] = h->fx[k];
    w.fy[k] = h->fy[k];
  }
  w.fcx = h->fcx[0];
  w.fcy = h->fcy[0];
  w.ez = h->e[2];
  g.nx = h->g.nx; g.ny = h->g.ny; g.xoff = h->g.xoff; g.yoff = h->g.yoff; g.gamma = h->g.gamma;
}

template <int ODR, int FLUX, int CH, int WPE>
static void launch_ysweep(Handle *h, const U2 &in, const W2 &w, const G2 &g)
{
  const int chunks = (g.ny - 5 + CH - 1) / CH;
  MLAU_LAUNCH(h, K_YSWEEP2D, (k_ysweep2d<ODR, FLUX, CH, WPE>), dim3((g.nx + 63) / 64, chunks), dim3(64), in, w, g);
}

// Sweep launch shapes: defaults, plus the variants the tuner may pick for the
// headline scheme (ODR=4, LHLLD).
template <int ODR, int FLUX>
static void sweeps2d(Handle *h, const U2 &in, const W2 &w, const G2 &g)
{
  const int tiles = (g.nx - 5 + 62) / 63;
  const dim3 gx((tiles + 3) / 4, g.ny);
  const int t = (ODR == 4 && FLUX == FLUX_LHLLD) ? h->tune : 0;
  if ((t & 1) == 0) MLAU_LAUNCH(h, K_XSWEEP2D, (k_xsweep2d<ODR, FLUX, 1>), gx, dim3(256), in, w, g);
  else MLAU_LAUNCH(h, K_XSWEEP2D, (k_xsweep2d<ODR, FLUX, 3>), gx, dim3(256), in, w, g);
  switch ((t >> 1) & 7) {
  case 1: launch_ysweep<ODR, FLUX, 16, 1>(h, in, w, g); break;
  case 2: launch_ysweep<ODR, FLUX, 4, 1>(h, in, w, g); break;
  case 3: launch_ysweep<ODR, FLUX, 6, 1>(h, in, w, g); break;
  default: launch_ysweep<ODR, FLUX, 8, 1>(h, in, w, g); break;
  }
}

template <int ODR, int FLUX>
static void stage2d(Handle *h, const U2 &in, const U2 &ut, const U2 &out, const W2 &w, const G2 &g, double ra,
                    double rb)
{
  const int nx = g.nx, ny = g.ny;
  dim3 b256(256);
  dim3 gcell(((nx + 255) / 256) * ny);
  MLAU_LAUNCH(h, K_PRIM2D, (k_prim2d<ODR>), gcell, b256, in, w, g, h->bc);
  if (h->s.ctw) MLAU_LAUNCH(h, K_AUX2D, (k_aux2d<1>), gcell, b256, in, w, g, h->bc);
  else MLAU_LAUNCH(h, K_AUX2D, (k_aux2d<0>), gcell, b256, in, w, g, h->bc);
  sweeps2d<ODR, FLUX>(h, in, w, g);
  MLAU_LAUNCH(h, K_EZ2D, (k_ez2d<ODR>), dim3(((nx - 5 + 255) / 256) * (ny - 5)), b256, w, g);
  const dim3 gu(((nx - 2 * g.xoff + 1 + 255) / 256) * (ny - 2 * g.yoff + 1));
  if (ra == 0.0)
    MLAU_LAUNCH(h, K_UPD2D, (k_update2d<ODR, true>), gu, b256, in, ut, out, w, g, (const StepInfo *)h->info, ra, rb);
  else
    MLAU_LAUNCH(h, K_UPD2D, (k_update2d<ODR, false>), gu, b256, in, ut, out, w, g, (const StepInfo *)h->info, ra, rb);
  const long nframe = (long)ny * 2 * g.xoff + (long)2 * g.yoff * nx;
  MLAU_LAUNCH(h, K_BC2D, k_bc2d, dim3((unsigned)((nframe + 255) / 256), 8), b256, out, g, h->bc);
}

template <int ODR, int FLUX>
static void step2d(Handle *h)
{
  U2 u0, u1;
  W2 w;
  G2 g;
  make_views(h, u0, u1, w, g);
  const int R_K = h->s.rk;
  const double rkf[3][2] = {{0.0, 1.0}, {0.5 + (R_K - 2) * 0.25, 0.5 - (R_K - 2) * 0.25}, {1. / 3., 2. / 3.}};
  for (int rk = 0; rk < R_K; rk++) {
    const U2 &in = (rk == 0) ? u0 : u1;
    const U2 &out = (rk == R_K - 1) ? u0 : u1;
    stage2d<ODR, FLUX>(h, in, u0, out, w, g, rkf[rk][0], rkf[rk][1]);
  }
}

typedef void (*Step2Fn)(Handle *);
template <int ODR>
static Step2Fn pick2(int flux)
{
  switch (flux) {
  case FLUX_MLAU: return step2d<ODR, FLUX_MLAU>;
  case FLUX_ROE: return step2d<ODR, FLUX_ROE>;
  case FLUX_HLLD: return step2d<ODR, FLUX_HLLD>;
  default: return step2d<ODR, FLUX_LHLLD>;
  }
}

int run_cfl_2d(Handle *h, double cfl, double vfloor)
{
  U2 u0, u1;
  W2 w;
  G2 g;
  make_views(h, u0, u1, w, g);
  hipLaunchKernelGGL(k_cfl_reset, dim3(1), dim3(1), 0, h->stream, h->info, vfloor);
  const long cells = (long)(g.nx - 2 * g.xoff) * (g.ny - 2 * g.yoff);
  const unsigned nb = (unsigned)((cells + 255) / 256 < 2048 ? (cells + 255) / 256 : 2048);
  MLAU_LAUNCH(h, K_CFL2D, k_cfl2d, dim3(nb), dim3(256), u0, g, h->info);
  const double dr = (h->g.dx < h->g.dy) ? h->g.dx : h->g.dy;
  hipLaunchKernelGGL(k_finalize_dt, dim3(1), dim3(1), 0, h->stream, h->info, cfl, dr, h->g.dx, h->g.dy, h->g.dz);
  return (int)hipGetLastError();
}

int run_steps_2d(Handle *h, int nsteps, bool cfl_mode, double cfl, double vfloor)
{
  Step2Fn fn;
  switch (h->s.odr) {
  case 1: fn = pick2<1>(h->s.flux); break;
  case 2: fn = pick2<2>(h->s.flux); break;
  case 3: fn = pick2<3>(h->s.flux); break;
  default: fn = pick2<4>(h->s.flux); break;
  }
  for (int n = 0; n < nsteps; n++) {
    fn(h);
    hipLaunchKernelGGL(k_advance_time, dim3(1), dim3(1), 0, h->stream, h->info);
    if (cfl_mode) run_cfl_2d(h, cfl, vfloor);
  }
  return (int)hipGetLastError();
}

}  // namespace mlau
