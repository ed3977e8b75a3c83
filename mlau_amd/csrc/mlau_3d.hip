// mlau_3d.hip -- 3D ideal-MHD + CT step on gfx950 (FP64), single rank.
//
// One RK stage of 3D/MPI/OTvortex/mhd_fd_ct_3d.c:178-1176:
//   k_prim3d    face B -> centre (cx,cy,cz) + primitives, halo fills folded in   :186-262
//   k_aux3d_zm  CT weights ct0 (xy), ct1 (yz), ct2 (zx) and shock sensors,
//               a thread per column walking 8 cells along z                    :264-377
//   k_xsweep3d  X reconstruction + Riemann, one wavefront per 63 faces of a row  :379-493
//   k_msweep3d  Y and Z reconstruction + Riemann: one lane per x, marching a
//               chunk along y (or z) with the stencil in registers            :587-701,795-909
//   k_edge3d    edge E-fields ex,ey,ez from the three sweeps                   :494-585,702-793,910-1001
//   k_update3d  cell update, CT Bx/By/Bz, SSP-RK combine                       :1003-1157
//   k_bc3d      ghost cells (single-rank periodic copies / walls)               :1161-1175
// Sweep outputs are stored in the sweep's rotated order
// (ro, m_n, m_t, m_u, B_t upwind part, B_u upwind part, en): x (n,t,u)=(x,y,z),
// y (y,z,x), z (z,x,y); the central induction parts go to fc[0] (t) and fc[1] (u).
#include "mlau_3d_body.h"

namespace mlau {

// Cell kernels: 1D grids of (plane pieces) x (planes or plane chunks), remapped
// so that each XCD works on a contiguous band and neighbouring rows share its L2.
struct Blk2 {
  unsigned piece, plane;
};
__device__ __forceinline__ Blk2 banded2(unsigned pieces)
{
  const unsigned b = xcd_band(blockIdx.x, gridDim.x);
  return Blk2{b % pieces, b / pieces};
}

template <int ODR>
__global__ __launch_bounds__(256) void k_prim3d(U3 u, double *vx_, double *vy_, double *vz_, double *pr_,
                                                double *m2_, double *cx_, double *cy_, double *cz_, G3 g, BcFlags bc)
{
  const Blk2 b = banded2(plane_pieces(g.nx, g.ny));
  int i, j;
  if (!plane_cell(g, b.piece, threadIdx.x, i, j)) return;
  prim3d_body<ODR>(u, vx_, vy_, vz_, pr_, m2_, cx_, cy_, cz_, g, bc, i, j, b.plane);
}

template <int CTW>
__global__ __launch_bounds__(256) void k_aux3d(U3 u, const double *vx, const double *vy, const double *vz,
                                               double *ct0, double *ct1, double *ct2, double *dvx, double *dvy,
                                               double *dvz, G3 g, BcFlags bc)
{
  const Blk2 b = banded2(plane_pieces(g.nx, g.ny));
  int i, j;
  if (!plane_cell(g, b.piece, threadIdx.x, i, j)) return;
  aux3d_body<CTW>(u, vx, vy, vz, ct0, ct1, ct2, dvx, dvy, dvz, g, bc, i, j, b.plane);
}

template <int CTW, int CZ>
__global__ __launch_bounds__(256) void k_aux3d_zm(U3 u, const double *vx, const double *vy, const double *vz,
                                                  double *ct0, double *ct1, double *ct2, double *dvx, double *dvy,
                                                  double *dvz, G3 g, BcFlags bc)
{
  const Blk2 b = banded2(plane_pieces(g.nx, g.ny));
  int i, j;
  if (!plane_cell(g, b.piece, threadIdx.x, i, j)) return;
  aux3d_zm_body<CTW, CZ>(u, vx, vy, vz, ct0, ct1, ct2, dvx, dvy, dvz, g, bc, i, j, b.plane);
}

// x sweep: row-flattened wavefronts, 4 per block
template <int ODR, int FLUX, int MIX>
__global__ __launch_bounds__(256) void k_xsweep3d(Sw3 w, G3 g)
{
  const long W = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (W >= xsweep_waves((long)g.ny * g.nz, g.nx)) return;  // whole wavefront
  xsweep3d_body<ODR, FLUX, MIX>(w, g, W, threadIdx.x & 63);
}

// y/z sweeps: lanes over the flattened (i, third coordinate) columns; blocks
// are (column piece, chunk along the sweep axis)
template <int ODR, int FLUX, int CH, int AX, int MIX>
__global__ __launch_bounds__(64) MLAU_TWO_WAVES void k_msweep3d(Sw3 w, G3 g)
{
  const long nc = (AX == 1) ? g.nz : g.ny;
  const unsigned pieces = (unsigned)((g.nx * nc + 63) / 64);
  const long p = (long)(blockIdx.x % pieces) * 64 + threadIdx.x;
  if (p >= g.nx * nc) return;
  msweep3d_body<ODR, FLUX, CH, AX, MIX>(w, g, (int)(p % g.nx), blockIdx.x / pieces, (int)(p / g.nx));
}

template <int ODR>
__global__ __launch_bounds__(256) void k_edge3d(Edge3 e, G3 g)
{
  edge3d_body<ODR>(e, g, blockIdx.x, blockIdx.y, blockIdx.z, threadIdx.x);
}

template <int ODR, int CZ>
__global__ __launch_bounds__(256) void k_edge3d_zm(Edge3 e, G3 g)
{
  const Blk2 b = banded2(plane_pieces(g.nx, g.ny));
  int i, j;
  if (!plane_cell(g, b.piece, threadIdx.x, i, j)) return;
  edge3d_zm_body<ODR, CZ>(e, g, i, j, b.plane);
}

template <int ODR, int CZ>
__global__ __launch_bounds__(256) void k_update3d_zm(U3 in, U3 ut, U3 out, Fl3 fl, const double *ex,
                                                     const double *ey, const double *ez, G3 g, const StepInfo *info,
                                                     double ra, double rb)
{
  const Blk2 b = banded2(plane_pieces(g.nx - 2 * g.xoff + 1, g.ny - 2 * g.yoff + 1));
  int i, j;
  if (!plane_face_cell(g, b.piece, threadIdx.x, i, j)) return;
  update3d_zm_body<ODR, CZ>(in, ut, out, fl, ex, ey, ez, g, info->dtdx, info->dtdy, info->dtdz, ra, rb, i, j,
                            b.plane);
}

template <int ODR>
__global__ __launch_bounds__(256) void k_update3d(U3 in, U3 ut, U3 out, Fl3 fl, const double *ex, const double *ey,
                                                  const double *ez, G3 g, const StepInfo *info, double ra, double rb)
{
  update3d_body<ODR>(in, ut, out, fl, ex, ey, ez, g, info->dtdx, info->dtdy, info->dtdz, ra, rb, blockIdx.x,
                     blockIdx.y, blockIdx.z, threadIdx.x);
}

// blockIdx.y = field
__global__ __launch_bounds__(256) void k_bc3d(U3 u, G3 g, BcFlags bc)
{
  bc3d_body(u, g, bc, blockIdx.y, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ __launch_bounds__(256) void k_cfl3d(U3 u, G3 g, StepInfo *info)
{
  const long total = (long)(g.nx - 2 * g.xoff) * (g.ny - 2 * g.yoff) * (g.nz - 2 * g.zoff);
  double v = 0.0;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const double vt = cfl3d_cell(u, g, t);
    if (vt > v) v = vt;
  }
  block_max_atomic<256>(v, &info->vmax_bits);
}

// ------------------------------------------------------------------- driver
constexpr int kMChunk = 16;
constexpr int kZChunk = 8;  // planes per thread of the z-marching edge/update kernels

static G3 geom3(Handle *h)
{
  G3 g;
  g.nx = h->g.nx; g.ny = h->g.ny; g.nz = h->g.nz;
  g.xoff = h->g.xoff; g.yoff = h->g.yoff; g.zoff = h->g.zoff;
  g.gamma = h->g.gamma;
  return g;
}

static Sw3 sweep_view(Handle *h, const U3 &in, int ax)
{
  double *const *P = h->prim;  // vx,vy,vz,pr,m2,cx,cy,cz
  const double *v[3] = {P[0], P[1], P[2]};
  const double *c[3] = {P[5], P[6], P[7]};
  const double *dv[3] = {h->aux[3], h->aux[4], h->aux[5]};
  const int n = ax, t = (ax + 1) % 3, u = (ax + 2) % 3;
  Sw3 w;
  w.ro = in.f[0];
  w.vn = v[n]; w.vt = v[t]; w.vu = v[u];
  w.bt = c[t]; w.bu = c[u];
  w.pr = P[3]; w.m2 = P[4];
  w.bnc = c[n];
  w.bface = in.f[5 + n];
  w.dvt = dv[t]; w.dvu = dv[u];
  double *const *F = (ax == 0) ? h->fx : (ax == 1) ? h->fy : h->fz;
  double *const *FC = (ax == 0) ? h->fcx : (ax == 1) ? h->fcy : h->fcz;
  for (int q = 0; q < 7; q++) w.out[q] = F[q];
  w.fc[0] = FC[0];
  w.fc[1] = FC[1];
  return w;
}

template <int ODR, int FLUX, int MIX>
static void stage3d(Handle *h, const U3 &in, const U3 &ut, const U3 &out, double ra, double rb)
{
  const G3 g = geom3(h);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  double *const *P = h->prim;
  dim3 b256(256);
  dim3 gcell((nx + 255) / 256, ny, nz);  // plane-parallel edge kernel (3D grid)
  const dim3 gflat(plane_pieces(nx, ny) * nz);
  MLAU_LAUNCH(h, K_PRIM3D, (k_prim3d<ODR>), gflat, b256, in, P[0], P[1], P[2], P[3], P[4], P[5], P[6], P[7], g,
              h->bc);
  const bool zm = !(h->tune & 8);  // tune bit 3: plane-parallel aux/edge/update kernels instead of z-marching
  const dim3 gz(plane_pieces(nx, ny) * ((nz + kZChunk - 1) / kZChunk));
  if (zm && h->s.ctw)
    MLAU_LAUNCH(h, K_AUX3D, (k_aux3d_zm<1, kZChunk>), gz, b256, in, (const double *)P[0], (const double *)P[1],
                (const double *)P[2], h->aux[0], h->aux[1], h->aux[2], h->aux[3], h->aux[4], h->aux[5], g, h->bc);
  else if (zm)
    MLAU_LAUNCH(h, K_AUX3D, (k_aux3d_zm<0, kZChunk>), gz, b256, in, (const double *)P[0], (const double *)P[1],
                (const double *)P[2], h->aux[0], h->aux[1], h->aux[2], h->aux[3], h->aux[4], h->aux[5], g, h->bc);
  else if (h->s.ctw)
    MLAU_LAUNCH(h, K_AUX3D, (k_aux3d<1>), gflat, b256, in, (const double *)P[0], (const double *)P[1],
                (const double *)P[2], h->aux[0], h->aux[1], h->aux[2], h->aux[3], h->aux[4], h->aux[5], g, h->bc);
  else
    MLAU_LAUNCH(h, K_AUX3D, (k_aux3d<0>), gflat, b256, in, (const double *)P[0], (const double *)P[1],
                (const double *)P[2], h->aux[0], h->aux[1], h->aux[2], h->aux[3], h->aux[4], h->aux[5], g, h->bc);
  Side sd(h);  // the y and z sweeps beside the x sweep
  sd.fork(2);
  MLAU_LAUNCH(h, K_XSWEEP3D, (k_xsweep3d<ODR, FLUX, MIX>), dim3((unsigned)((xsweep_waves((long)ny * nz, nx) + 3) / 4)),
              b256, sweep_view(h, in, 0), g);
  sd.enter(0);
  MLAU_LAUNCH(h, K_YSWEEP3D, (k_msweep3d<ODR, FLUX, kMChunk, 1, MIX>),
              dim3((unsigned)((((long)nx * nz + 63) / 64) * ((ny - 5 + kMChunk - 1) / kMChunk))), dim3(64),
              sweep_view(h, in, 1), g);
  sd.enter(1);
  MLAU_LAUNCH(h, K_ZSWEEP3D, (k_msweep3d<ODR, FLUX, kMChunk, 2, MIX>),
              dim3((unsigned)((((long)nx * ny + 63) / 64) * ((nz - 5 + kMChunk - 1) / kMChunk))), dim3(64),
              sweep_view(h, in, 2), g);
  sd.leave();
  sd.join();
  Edge3 e;
  e.fxt = h->fx[4]; e.fxu = h->fx[5]; e.fcx0 = h->fcx[0]; e.fcx1 = h->fcx[1];
  e.fyt = h->fy[4]; e.fyu = h->fy[5]; e.fcy0 = h->fcy[0]; e.fcy1 = h->fcy[1];
  e.fzt = h->fz[4]; e.fzu = h->fz[5]; e.fcz0 = h->fcz[0]; e.fcz1 = h->fcz[1];
  e.ct0 = h->aux[0]; e.ct1 = h->aux[1]; e.ct2 = h->aux[2];
  e.ex = h->e[0]; e.ey = h->e[1]; e.ez = h->e[2];
  if (zm)
    MLAU_LAUNCH(h, K_EDGE3D, (k_edge3d_zm<ODR, kZChunk>), dim3(plane_pieces(nx, ny) * ((nz + kZChunk - 1) / kZChunk)),
                b256, e, g);
  else
    MLAU_LAUNCH(h, K_EDGE3D, (k_edge3d<ODR>), gcell, b256, e, g);
  Fl3 fl;
  for (int q = 0; q < 7; q++) {
    fl.x[q] = h->fx[q];
    fl.y[q] = h->fy[q];
    fl.z[q] = h->fz[q];
  }
  if (zm)
    MLAU_LAUNCH(h, K_UPD3D, (k_update3d_zm<ODR, kZChunk>),
                dim3(plane_pieces(nx - 2 * g.xoff + 1, ny - 2 * g.yoff + 1) *
                     ((nz - 2 * g.zoff + 1 + kZChunk - 1) / kZChunk)),
                b256, in, ut, out, fl, (const double *)h->e[0], (const double *)h->e[1], (const double *)h->e[2], g,
                (const StepInfo *)h->info, ra, rb);
  else
    MLAU_LAUNCH(h, K_UPD3D, (k_update3d<ODR>),
                dim3((nx - 2 * g.xoff + 1 + 255) / 256, ny - 2 * g.yoff + 1, nz - 2 * g.zoff + 1), b256, in, ut, out,
                fl, (const double *)h->e[0], (const double *)h->e[1], (const double *)h->e[2], g,
                (const StepInfo *)h->info, ra, rb);
  const long nslab = (long)2 * g.xoff * ny * nz + (long)nx * 2 * g.yoff * nz + (long)nx * ny * 2 * g.zoff;
  MLAU_LAUNCH(h, K_BC3D, k_bc3d, dim3((unsigned)((nslab + 255) / 256), 8), b256, out, g, h->bc);
}

template <int ODR, int FLUX, int MIX>
static void step3d(Handle *h)
{
  U3 u0, u1;
  for (int n = 0; n < 8; n++) {
    u0.f[n] = h->U0[n];
    u1.f[n] = h->U1[n];
  }
  const int R_K = h->s.rk;
  const double rkf[3][2] = {{0.0, 1.0}, {0.5 + (R_K - 2) * 0.25, 0.5 - (R_K - 2) * 0.25}, {1. / 3., 2. / 3.}};
  for (int rk = 0; rk < R_K; rk++) {
    const U3 &in = (rk == 0) ? u0 : u1;
    const U3 &out = (rk == R_K - 1) ? u0 : u1;
    stage3d<ODR, FLUX, MIX>(h, in, u0, out, rkf[rk][0], rkf[rk][1]);
  }
}

typedef void (*Step3Fn)(Handle *);
template <int ODR, int MIX>
static Step3Fn pick3f(int flux)
{
  switch (flux) {
  case FLUX_MLAU: return step3d<ODR, FLUX_MLAU, MIX>;
  case FLUX_ROE: return step3d<ODR, FLUX_ROE, MIX>;
  case FLUX_HLLD: return step3d<ODR, FLUX_HLLD, MIX>;
  default: return step3d<ODR, FLUX_LHLLD, MIX>;
  }
}
template <int ODR>
static Step3Fn pick3(int flux, int precision)
{
  switch (precision) {
  case MLAU_MIXED: return pick3f<ODR, MLAU_MIXED>(flux);
  case MLAU_MIXED_RECON: return pick3f<ODR, MLAU_MIXED_RECON>(flux);
  default: return pick3f<ODR, MLAU_FP64>(flux);
  }
}

int run_cfl_3d(Handle *h, double cfl, double vfloor)
{
  const G3 g = geom3(h);
  U3 u0;
  for (int n = 0; n < 8; n++) u0.f[n] = h->U0[n];
  hipLaunchKernelGGL(k_cfl_reset, dim3(1), dim3(1), 0, h->stream, h->info, vfloor);
  const long cells = (long)(g.nx - 2 * g.xoff) * (g.ny - 2 * g.yoff) * (g.nz - 2 * g.zoff);
  const unsigned nb = (unsigned)((cells + 255) / 256 < 2048 ? (cells + 255) / 256 : 2048);
  MLAU_LAUNCH(h, K_CFL3D, k_cfl3d, dim3(nb), dim3(256), u0, g, h->info);
  double dr = h->g.dx;
  if (h->g.dy < dr) dr = h->g.dy;
  if (h->g.dz < dr) dr = h->g.dz;
  hipLaunchKernelGGL(k_finalize_dt, dim3(1), dim3(1), 0, h->stream, h->info, cfl, dr, h->g.dx, h->g.dy, h->g.dz);
  return (int)hipGetLastError();
}

int run_steps_3d(Handle *h, int nsteps, bool cfl_mode, double cfl, double vfloor)
{
  Step3Fn fn;
  switch (h->s.odr) {
  case 1: fn = pick3<1>(h->s.flux, h->s.precision); break;
  case 2: fn = pick3<2>(h->s.flux, h->s.precision); break;
  case 3: fn = pick3<3>(h->s.flux, h->s.precision); break;
  default: fn = pick3<4>(h->s.flux, h->s.precision); break;
  }
  for (int n = 0; n < nsteps; n++) {
    fn(h);
    hipLaunchKernelGGL(k_advance_time, dim3(1), dim3(1), 0, h->stream, h->info);
    if (cfl_mode) run_cfl_3d(h, cfl, vfloor);
  }
  return (int)hipGetLastError();
}

}  // namespace mlau
