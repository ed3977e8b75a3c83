// mlau_1d.hip -- 1D ideal-MHD step on gfx950 (FP64).
//
// One RK stage of 1D/SERIAL/shock/mhd_fd4c_1d.c:110-257:
//   k_prim1d    primitives vx,vy,vz,pr                       :121-131
//   k_sweep1d   reconstruction + Riemann solver, one wavefront
//               per 63 faces (left face states passed by shuffle)  :137-207
//   k_openbc1d  characteristic open boundary, stage 0 only    :209-218 (openbc1d_mhd.c:21-175)
//   k_update1d  flux difference + SSP-RK combine              :225-255
//   k_bc1d      periodic ghosts after every stage (1D/SERIAL/wave/prdc_1d.c:18-25)
// Fields in argument order ro,mx,my,mz,en,by,bz; bx is a constant.
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

__global__ __launch_bounds__(256) void k_prim1d(U1d u, W1d w, G1d g)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.nx) return;
  const double bx2 = g.bx * g.bx;
  const double ro = u.f[0][i];
  const double iro = 1.0 / ro;
  const double vx = u.f[1][i] * iro, vy = u.f[2][i] * iro, vz = u.f[3][i] * iro;
  const double by = u.f[5][i], bz = u.f[6][i];
  const double v2 = vx * vx + vy * vy + vz * vz;
  const double b2 = bx2 + by * by + bz * bz;
  w.vx[i] = vx;
  w.vy[i] = vy;
  w.vz[i] = vz;
  w.pr[i] = (g.gamma - 1.0) * (u.f[4][i] - 0.5 * (ro * v2 + b2));
}

template <int ODR, int FLUX, bool MIX>
__global__ __launch_bounds__(256) void k_sweep1d(U1d u, W1d w, G1d g)
{
  const int lane = threadIdx.x & 63;
  const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int nx = g.nx;
  const int c0 = 2 + 63 * tile;
  if (c0 > nx - 4) return;
  const int c = c0 + lane;
  const int cc = (c < nx - 3) ? c : nx - 3;
  Stencil S;
#pragma unroll
  for (int m = 0; m < 5; m++) {
    const int q = cc - 2 + m;
    S.q[0][m] = u.f[0][q];
    S.q[1][m] = w.vx[q];
    S.q[2][m] = w.vy[q];
    S.q[3][m] = w.vz[q];
    S.q[4][m] = u.f[5][q];
    S.q[5][m] = u.f[6][q];
    S.q[6][m] = w.pr[q];
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
  const double dv0 = w.vx[c] - w.vx[c - 1];
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
__device__ void eigen_st_full(const double p[8], double gamma, double al[7], double L[7][7], double R[7][7])
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
__global__ void k_openbc1d(U1d in, U1d dst0, U1d dst1, G1d g, const StepInfo *info)
{
  const int side = threadIdx.x;
  if (side > 1) return;
  const int nx = g.nx, xoff = g.xoff;
  const int sgn = side == 0 ? 1 : -1;
  const int si = side == 0 ? xoff : nx - xoff - 1;
  const int so = si - sgn;
  const double bx = g.bx, bx2 = bx * bx, gamma = g.gamma;
  const double dtdx = info->dtdx;
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

template <int ODR>
__global__ __launch_bounds__(256) void k_update1d(U1d in, U1d ut, U1d out, W1d w, G1d g, const StepInfo *info,
                                                  double ra, double rb)
{
  const int i = 4 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.nx - 4) return;
  const double dtdx = info->dtdx;
#pragma unroll
  for (int n = 0; n < 7; n++) {
    const double *a = w.fx[n] + i;
    const double du = diff4<ODR>(a[-1], a[0], a[1], a[2]);
    out.f[n][i] = ra * ut.f[n][i] + rb * (in.f[n][i] - dtdx * du);
  }
}

// prdc_1d on the seven fields (ghosts copied from the periodic images)
__global__ void k_bc1d(U1d u, G1d g)
{
  const int t = threadIdx.x;  // 7 fields x xoff ghosts
  const int n = t / g.xoff, i = t % g.xoff;
  if (n >= 7) return;
  double *f = u.f[n];
  const int nx = g.nx, xoff = g.xoff;
  f[nx - 1 - i] = f[2 * xoff - 1 - i];
  f[i] = f[nx - 2 * xoff + i];
}

// Driver's one-off dt estimate (1D/SERIAL/shock/main.cpp:114-128): all cells.
__global__ __launch_bounds__(256) void k_cfl1d(U1d u, G1d g, StepInfo *info)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (i < g.nx) {
    const double iro = 1.0 / u.f[0][i];
    const double vx = u.f[1][i] * iro, vy = u.f[2][i] * iro, vz = u.f[3][i] * iro;
    const double v2 = vx * vx + vy * vy + vz * vz;
    const double b2 = g.bx * g.bx + u.f[5][i] * u.f[5][i] + u.f[6][i] * u.f[6][i];
    const double pr = (g.gamma - 1.0) * (u.f[4][i] - 0.5 * (u.f[0][i] * v2 + b2));
    const double cf = sqrt((g.gamma * pr + b2) * iro);
    const double vt = fabs(vx) + cf;
    if (vt > v) v = vt;
  }
  block_max_atomic<256>(v, &info->vmax_bits);
}

static void views1d(Handle *h, U1d &u0, U1d &u1, W1d &w, G1d &g)
{
  for (int n = 0; n < 7; n++) {
    u0.f[n] = h->U0[n];
    u1.f[n] = h->U1[n];
    w.fx[n] = h->fx[n];
  }
  w.vx = h->prim[0]; w.vy = h->prim[1]; w.vz = h->prim[2]; w.pr = h->prim[3];
  g.nx = h->g.nx; g.xoff = h->g.xoff; g.gamma = h->g.gamma; g.bx = h->g.bx;
}

template <int ODR, int FLUX, bool MIX>
static void step1d(Handle *h)
{
  U1d u0, u1;
  W1d w;
  G1d g;
  views1d(h, u0, u1, w, g);
  const int R_K = h->s.rk, nx = g.nx;
  const double rkf[3][2] = {{0.0, 1.0}, {0.5 + (R_K - 2) * 0.25, 0.5 - (R_K - 2) * 0.25}, {1. / 3., 2. / 3.}};
  const int tiles = (nx - 5 + 62) / 63;
  for (int rk = 0; rk < R_K; rk++) {
    const U1d &in = (rk == 0) ? u0 : u1;
    const U1d &out = (rk == R_K - 1) ? u0 : u1;
    MLAU_LAUNCH(h, K_PRIM1D, k_prim1d, dim3((nx + 255) / 256), dim3(256), in, w, g);
    MLAU_LAUNCH(h, K_SWEEP1D, (k_sweep1d<ODR, FLUX, MIX>), dim3((tiles + 3) / 4), dim3(256), in, w, g);
    if (rk == 0 && h->s.bc1d == MLAU_BC1D_OPEN)
      MLAU_LAUNCH(h, K_OPENBC1D, k_openbc1d, dim3(1), dim3(64), in, u0, u1, g, (const StepInfo *)h->info);
    MLAU_LAUNCH(h, K_UPD1D, (k_update1d<ODR>), dim3((nx - 8 + 255) / 256), dim3(256), in, u0, out, w, g,
                (const StepInfo *)h->info, rkf[rk][0], rkf[rk][1]);
    if (h->s.bc1d == MLAU_BC1D_PERIODIC) MLAU_LAUNCH(h, K_BC1D, k_bc1d, dim3(1), dim3(64), out, g);
  }
}

typedef void (*Step1Fn)(Handle *);
template <int ODR, bool MIX>
static Step1Fn pick1f(int flux)
{
  switch (flux) {
  case FLUX_MLAU: return step1d<ODR, FLUX_MLAU, MIX>;
  case FLUX_ROE: return step1d<ODR, FLUX_ROE, MIX>;
  case FLUX_HLLD: return step1d<ODR, FLUX_HLLD, MIX>;
  default: return step1d<ODR, FLUX_LHLLD, MIX>;
  }
}
template <int ODR>
static Step1Fn pick1(int flux, int precision)
{
  return precision == MLAU_MIXED ? pick1f<ODR, true>(flux) : pick1f<ODR, false>(flux);
}

int run_cfl_1d(Handle *h, double cfl, double vfloor)
{
  U1d u0, u1;
  W1d w;
  G1d g;
  views1d(h, u0, u1, w, g);
  hipLaunchKernelGGL(k_cfl_reset, dim3(1), dim3(1), 0, h->stream, h->info, vfloor);
  MLAU_LAUNCH(h, K_CFL1D, k_cfl1d, dim3((g.nx + 255) / 256), dim3(256), u0, g, h->info);
  hipLaunchKernelGGL(k_finalize_dt, dim3(1), dim3(1), 0, h->stream, h->info, cfl, h->g.dx, h->g.dx, h->g.dy,
                     h->g.dz);
  return (int)hipGetLastError();
}

int run_steps_1d(Handle *h, int nsteps, bool cfl_mode, double cfl, double vfloor)
{
  if (h->g.xoff != 4) return (int)hipErrorInvalidValue;
  Step1Fn fn;
  switch (h->s.odr) {
  case 1: fn = pick1<1>(h->s.flux, h->s.precision); break;
  case 2: fn = pick1<2>(h->s.flux, h->s.precision); break;
  case 3: fn = pick1<3>(h->s.flux, h->s.precision); break;
  default: fn = pick1<4>(h->s.flux, h->s.precision); break;
  }
  for (int n = 0; n < nsteps; n++) {
    fn(h);
    hipLaunchKernelGGL(k_advance_time, dim3(1), dim3(1), 0, h->stream, h->info);
    if (cfl_mode) run_cfl_1d(h, cfl, vfloor);
  }
  return (int)hipGetLastError();
}

}  // namespace mlau
