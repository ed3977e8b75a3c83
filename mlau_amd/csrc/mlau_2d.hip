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
// followed by k_bc2d (ghost cells, :636-638); on periodic grids the update
// writes the ghost images itself (k_update2d_per) and no boundary pass follows.  The reference's halo fills of
// the intermediate fields (bc2d on cx,cy,dvx,dvy) are folded into the
// kernels that produce them: a ghost value is computed at its source cell and
// multiplied by the boundary factor.
#include "mlau_2d_fused.h"

namespace mlau {

// Single-grid kernels: the cell kernels tile their region in flattened
// 256-cell pieces, remapped so that each XCD works on a contiguous band of
// rows (xcd_band).
#define REGION_CELL(W, H, I0, J0)                                                                   \
  int i, j;                                                                                        \
  if (!region_cell((W), (H), (I0), (J0), xcd_band(blockIdx.x, gridDim.x), threadIdx.x, i, j)) return;

template <int ODR>
__global__ __launch_bounds__(256) void k_prim2d(U2 u, W2 w, G2 g, BcFlags bc)
{
  REGION_CELL(g.nx, g.ny, 0, 0)
  prim2d_body<ODR>(u, w, g, bc, i, j);
}

template <int CTW>
__global__ __launch_bounds__(256) void k_aux2d(U2 u, W2 w, G2 g, BcFlags bc)
{
  REGION_CELL(g.nx, g.ny, 0, 0)
  aux2d_body<CTW>(u, w, g, bc, i, j);
}

template <int ODR, int CTW>
__global__ __launch_bounds__(256) void k_primaux2d(U2 u, W2 w, G2 g, BcFlags bc)
{
  REGION_CELL(g.nx, g.ny, 0, 0)
  primaux2d_body<ODR, CTW>(u, w, g, bc, i, j);
}

// x sweep: row-flattened wavefronts, 4 per block
template <int ODR, int FLUX, int MIX>
__global__ __launch_bounds__(256) void k_xsweep2d(U2 u, W2 w, G2 g)
{
  const long W = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (W >= xsweep_waves(g.ny, g.nx)) return;  // whole wavefront
  xsweep2d_body<ODR, FLUX, MIX>(u, w, g, W, threadIdx.x & 63);
}

template <int ODR, int FLUX, int CH, int MIX>
__global__ __launch_bounds__(64) void k_ysweep2d(U2 u, W2 w, G2 g)
{
  ysweep2d_body<ODR, FLUX, CH, MIX>(u, w, g, blockIdx.x, blockIdx.y, threadIdx.x);
}

template <int ODR>
__global__ __launch_bounds__(256) void k_ez2d(W2 w, G2 g)
{
  REGION_CELL(g.nx - 5, g.ny - 5, 3, 3)
  ez2d_body<ODR>(w, g, i, j);
}

template <int ODR, bool FIRST>
__global__ __launch_bounds__(256) void k_update2d(U2 in, U2 ut, U2 out, W2 w, G2 g, const StepInfo *info,
                                                  double ra, double rb)
{
  REGION_CELL(g.nx - 2 * g.xoff + 1, g.ny - 2 * g.yoff + 1, g.xoff, g.yoff)
  update2d_body<ODR, FIRST>(in, ut, out, w, g, info->dtdx, info->dtdy, ra, rb, i, j);
}

// periodic grids: interior cells, ghost images written with them (no k_bc2d)
template <int ODR, bool FIRST>
__global__ __launch_bounds__(256) void k_update2d_per(U2 in, U2 ut, U2 out, W2 w, G2 g, const StepInfo *info,
                                                      double ra, double rb)
{
  REGION_CELL(g.nx - 2 * g.xoff, g.ny - 2 * g.yoff, g.xoff, g.yoff)
  update2d_body<ODR, FIRST, true>(in, ut, out, w, g, info->dtdx, info->dtdy, ra, rb, i, j);
}

// Periodic grids: the sweeps and the E-field form primitives, shock sensors
// and CT weights from the state themselves (mlau_2d_fused.h).
template <int ODR, int FLUX, int MIX>
__global__ __launch_bounds__(256) void k_xsweep2d_inl(U2 u, W2 w, G2 g)
{
  const long W = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (W >= xsweep_waves(g.ny, g.nx)) return;  // whole wavefront
  xsweep2d_inl_body<ODR, FLUX, MIX>(u, w, g, W, threadIdx.x & 63);
}

template <int ODR, int FLUX, int CH, int MIX, int CTW>
__global__ __launch_bounds__(64) void k_ysweep2d_inl(U2 u, W2 w, G2 g)
{
  ysweep2d_inl_body<ODR, FLUX, CH, MIX, CTW>(u, w, g, blockIdx.x, blockIdx.y, threadIdx.x);
}

// CT weight 1/2 at corners 3..nx-3 x 3..ny-3 (no upwind weighting)
__global__ __launch_bounds__(256) void k_ctx_half2d(W2 w, G2 g)
{
  REGION_CELL(g.nx - 5, g.ny - 5, 3, 3)
  w.ctx[(long)j * g.nx + i] = 0.5;
}

// blockIdx.y = field
__global__ __launch_bounds__(256) void k_bc2d(U2 u, G2 g, BcFlags bc)
{
  bc2d_body(u, g, bc, blockIdx.y, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// vmax over the interior, grid-stride with one atomic per block.
__global__ __launch_bounds__(256) void k_cfl2d(U2 u, G2 g, StepInfo *info)
{
  const long total = (long)(g.nx - 2 * g.xoff) * (g.ny - 2 * g.yoff);
  double v = 0.0;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const double vt = cfl2d_cell(u, g, t);
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

template <int ODR, int FLUX, int CH, int MIX>
static void launch_ysweep(Handle *h, const U2 &in, const W2 &w, const G2 &g)
{
  const int chunks = (g.ny - 5 + CH - 1) / CH;
  MLAU_LAUNCH(h, K_YSWEEP2D, (k_ysweep2d<ODR, FLUX, CH, MIX>), dim3((g.nx + 63) / 64, chunks), dim3(64), in, w, g);
}

// Sweep launch shapes: defaults, plus the chunk lengths the tuner may pick for
// WCNS4 with LHLLD (ODR=4; the 2048^2 vortex benchmark's scheme).
template <int ODR, int FLUX, int MIX>
static void sweeps2d(Handle *h, const U2 &in, const W2 &w, const G2 &g)
{
  const dim3 gx((unsigned)((xsweep_waves(g.ny, g.nx) + 3) / 4));
  Side sd(h);  // the y sweep beside the x sweep
  sd.fork(1);
  MLAU_LAUNCH(h, K_XSWEEP2D, (k_xsweep2d<ODR, FLUX, MIX>), gx, dim3(256), in, w, g);
  sd.enter(0);
  if constexpr (ODR == 4 && FLUX == FLUX_LHLLD) {  // the chunk variants for ODR=4 with LHLLD only
    switch (h->tune & 3) {
    case 1: launch_ysweep<ODR, FLUX, 16, MIX>(h, in, w, g); break;
    case 2: launch_ysweep<ODR, FLUX, 4, MIX>(h, in, w, g); break;
    default: launch_ysweep<ODR, FLUX, 8, MIX>(h, in, w, g); break;
    }
  } else {
    launch_ysweep<ODR, FLUX, 8, MIX>(h, in, w, g);
  }
  sd.leave();
  sd.join();
}

template <int ODR, int FLUX, int MIX>
static void stage2d(Handle *h, const U2 &in, const U2 &ut, const U2 &out, const W2 &w, const G2 &g, double ra,
                    double rb)
{
  const int nx = g.nx, ny = g.ny;
  dim3 b256(256);
  dim3 gcell(region_pieces(nx, ny));
  if (!(h->tune & 4)) {  // separate prim and aux passes (faster on one large grid); tune bit 2: one fused pass
    MLAU_LAUNCH(h, K_PRIM2D, (k_prim2d<ODR>), gcell, b256, in, w, g, h->bc);
    if (h->s.ctw) MLAU_LAUNCH(h, K_AUX2D, (k_aux2d<1>), gcell, b256, in, w, g, h->bc);
    else MLAU_LAUNCH(h, K_AUX2D, (k_aux2d<0>), gcell, b256, in, w, g, h->bc);
  } else if (h->s.ctw) {
    MLAU_LAUNCH(h, K_PRIMAUX2D, (k_primaux2d<ODR, 1>), gcell, b256, in, w, g, h->bc);
  } else {
    MLAU_LAUNCH(h, K_PRIMAUX2D, (k_primaux2d<ODR, 0>), gcell, b256, in, w, g, h->bc);
  }
  sweeps2d<ODR, FLUX, MIX>(h, in, w, g);
  MLAU_LAUNCH(h, K_EZ2D, (k_ez2d<ODR>), dim3(region_pieces(nx - 5, ny - 5)), b256, w, g);
  if (h->periodic_images && !(h->tune & 16)) {  // tune bit 4: separate boundary pass
    const dim3 gp(region_pieces(nx - 2 * g.xoff, ny - 2 * g.yoff));
    const StepInfo *info = h->info;
    if (ra == 0.0) MLAU_LAUNCH(h, K_UPD2D, (k_update2d_per<ODR, true>), gp, b256, in, ut, out, w, g, info, ra, rb);
    else MLAU_LAUNCH(h, K_UPD2D, (k_update2d_per<ODR, false>), gp, b256, in, ut, out, w, g, info, ra, rb);
    return;
  }
  const dim3 gu(region_pieces(nx - 2 * g.xoff + 1, ny - 2 * g.yoff + 1));
  if (ra == 0.0)
    MLAU_LAUNCH(h, K_UPD2D, (k_update2d<ODR, true>), gu, b256, in, ut, out, w, g, (const StepInfo *)h->info, ra, rb);
  else
    MLAU_LAUNCH(h, K_UPD2D, (k_update2d<ODR, false>), gu, b256, in, ut, out, w, g, (const StepInfo *)h->info, ra, rb);
  const long nframe = (long)ny * 2 * g.xoff + (long)2 * g.yoff * nx;
  MLAU_LAUNCH(h, K_BC2D, k_bc2d, dim3((unsigned)((nframe + 255) / 256), 8), b256, out, g, h->bc);
}

template <int ODR, int FLUX, int CH, int MIX>
static void launch_ysweep_inl(Handle *h, const U2 &in, const W2 &w, const G2 &g)
{
  const dim3 gr((g.nx + 63) / 64, (g.ny - 5 + CH - 1) / CH);
  if (h->s.ctw) MLAU_LAUNCH(h, K_YSW2D_INL, (k_ysweep2d_inl<ODR, FLUX, CH, MIX, 1>), gr, dim3(64), in, w, g);
  else MLAU_LAUNCH(h, K_YSW2D_INL, (k_ysweep2d_inl<ODR, FLUX, CH, MIX, 0>), gr, dim3(64), in, w, g);
}

// Periodic grids (with room for the ghost images written by the update): the
// stage without primitive arrays.  Tune bits 0-1 (ODR=4 with LHLLD only):
// y-sweep chunk of 8 (default), 16 or 4 rows.
template <int ODR, int FLUX, int MIX>
static void stage2d_inl(Handle *h, const U2 &in, const U2 &ut, const U2 &out, const W2 &w, const G2 &g, double ra,
                        double rb)
{
  const dim3 gx((unsigned)((xsweep_waves(g.ny, g.nx) + 3) / 4));
  Side sd(h);  // the y sweep beside the x sweep
  sd.fork(1);
  MLAU_LAUNCH(h, K_XSW2D_INL, (k_xsweep2d_inl<ODR, FLUX, MIX>), gx, dim3(256), in, w, g);
  sd.enter(0);
  if constexpr (ODR == 4 && FLUX == FLUX_LHLLD) {  // the chunk variants for ODR=4 with LHLLD only
    switch (h->tune & 3) {
    case 1: launch_ysweep_inl<ODR, FLUX, 16, MIX>(h, in, w, g); break;
    case 2: launch_ysweep_inl<ODR, FLUX, 4, MIX>(h, in, w, g); break;
    default: launch_ysweep_inl<ODR, FLUX, 8, MIX>(h, in, w, g); break;
    }
  } else {
    launch_ysweep_inl<ODR, FLUX, 8, MIX>(h, in, w, g);
  }
  sd.leave();
  sd.join();
  if (!h->s.ctw) {  // constant weight 1/2 (aux2d_body with CTW 0) over the corners the E-field reads
    const dim3 gc(region_pieces(g.nx - 5, g.ny - 5));
    MLAU_LAUNCH(h, K_AUX2D, k_ctx_half2d, gc, dim3(256), w, g);
  }
  MLAU_LAUNCH(h, K_EZ2D, (k_ez2d<ODR>), dim3(region_pieces(g.nx - 5, g.ny - 5)), dim3(256), w, g);
  const dim3 gp(region_pieces(g.nx - 2 * g.xoff, g.ny - 2 * g.yoff));
  const StepInfo *info = h->info;
  if (ra == 0.0) MLAU_LAUNCH(h, K_UPD2D, (k_update2d_per<ODR, true>), gp, dim3(256), in, ut, out, w, g, info, ra, rb);
  else MLAU_LAUNCH(h, K_UPD2D, (k_update2d_per<ODR, false>), gp, dim3(256), in, ut, out, w, g, info, ra, rb);
}

// The stage without primitive arrays applies to periodic grids whose interior
// spans two ghost widths per axis, unless tune bit 5 asks for the separate
// prim/aux passes (or bit 4 for those with a boundary pass).
static bool inline2d(const Handle *h) { return h->periodic_images && !(h->tune & (16 | 32)); }

template <int ODR, int FLUX, int MIX>
static void step2d(Handle *h)
{
  U2 u0, u1;
  W2 w;
  G2 g;
  make_views(h, u0, u1, w, g);
  const int R_K = h->s.rk;
  const double rkf[3][2] = {{0.0, 1.0}, {0.5 + (R_K - 2) * 0.25, 0.5 - (R_K - 2) * 0.25}, {1. / 3., 2. / 3.}};
  const bool inl = inline2d(h);
  for (int rk = 0; rk < R_K; rk++) {
    const U2 &in = (rk == 0) ? u0 : u1;
    const U2 &out = (rk == R_K - 1) ? u0 : u1;
    if (inl) stage2d_inl<ODR, FLUX, MIX>(h, in, u0, out, w, g, rkf[rk][0], rkf[rk][1]);
    else stage2d<ODR, FLUX, MIX>(h, in, u0, out, w, g, rkf[rk][0], rkf[rk][1]);
  }
}

typedef void (*Step2Fn)(Handle *);
template <int ODR, int MIX>
static Step2Fn pick2f(int flux)
{
  switch (flux) {
  case FLUX_MLAU: return step2d<ODR, FLUX_MLAU, MIX>;
  case FLUX_ROE: return step2d<ODR, FLUX_ROE, MIX>;
  case FLUX_HLLD: return step2d<ODR, FLUX_HLLD, MIX>;
  default: return step2d<ODR, FLUX_LHLLD, MIX>;
  }
}
template <int ODR>
static Step2Fn pick2(int flux, int precision)
{
  switch (precision) {
  case MLAU_MIXED: return pick2f<ODR, MLAU_MIXED>(flux);
  case MLAU_MIXED_RECON: return pick2f<ODR, MLAU_MIXED_RECON>(flux);
  default: return pick2f<ODR, MLAU_FP64>(flux);
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
  case 1: fn = pick2<1>(h->s.flux, h->s.precision); break;
  case 2: fn = pick2<2>(h->s.flux, h->s.precision); break;
  case 3: fn = pick2<3>(h->s.flux, h->s.precision); break;
  default: fn = pick2<4>(h->s.flux, h->s.precision); break;
  }
  for (int n = 0; n < nsteps; n++) {
    fn(h);
    hipLaunchKernelGGL(k_advance_time, dim3(1), dim3(1), 0, h->stream, h->info);
    if (cfl_mode) run_cfl_2d(h, cfl, vfloor);
  }
  return (int)hipGetLastError();
}

}  // namespace mlau
