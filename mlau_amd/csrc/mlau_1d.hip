// mlau_1d.hip -- 1D ideal-MHD step on gfx950 (FP64).
//
// One RK stage of 1D/SERIAL/shock/mhd_fd4c_1d.c:110-257 in two launches:
//   k_sweep1d   primitives (:121-131, formed per stencil point) + reconstruction
//               + Riemann solver, one wavefront per 63 faces (left face
//               states passed by shuffle)                                :137-207
//   k_update1d  flux difference + SSP-RK combine (:225-255); its last block
//               runs the characteristic open boundary on stage 0 (:209-218,
//               openbc1d_mhd.c:21-175), periodic ghosts are written with their
//               source cells (1D/SERIAL/wave/prdc_1d.c:18-25), and the last
//               stage advances the step's time.
// RK1 with the open boundary updates in place, so there the boundary runs
// first in a launch of its own (k_openbc1d).
// At ODR <= 2 with R_K >= 2 a stage is one launch (k_stage1d: flux and update
// per wavefront) writing a third state u0 -> u1 -> u2 -> u0 (u2: the flux
// arrays, which it does not use); with the open boundary, stage 0 keeps the
// two launches (the boundary rewrites ghosts its stencils read) and stage 1
// copies those fixed ghosts into u2.
// Fields in argument order ro,mx,my,mz,en,by,bz; bx is a constant.
#include "mlau_1d_body.h"

namespace mlau {

template <int ODR, int FLUX, int MIX>
__global__ __launch_bounds__(256) void k_sweep1d(U1d u, W1d w, G1d g)
{
  sweep1d_body<ODR, FLUX, MIX>(u, w, g, blockIdx.x, threadIdx.x);
}

// thread 0 = left edge, thread 1 = right edge
__global__ void k_openbc1d(U1d in, U1d dst0, U1d dst1, G1d g, const StepInfo *info)
{
  openbc1d_body(in, dst0, dst1, g, info->dtdx, threadIdx.x);
}

// obc: the last block runs the open boundary on u0 and u1 (stage 0 of RK2/RK3:
// it reads the stage input u0 and writes only ghosts, the update reads only
// interior cells of u0 and writes u1, so the two are independent).
// adv: block 0 advances the step's time (tim, nstep; the blocks read only dtdx).
template <int ODR, bool PER>
__global__ __launch_bounds__(256) void k_update1d(U1d in, U1d ut, U1d out, W1d w, G1d g, StepInfo *info, double ra,
                                                  double rb, U1d u0, U1d u1, int obc, int adv)
{
  if (obc && blockIdx.x == gridDim.x - 1) {
    openbc1d_body(in, u0, u1, g, info->dtdx, threadIdx.x);
    return;
  }
  const double dtdx = info->dtdx;
  update1d_body<ODR, PER>(in, ut, out, w, g, dtdx, ra, rb, blockIdx.x, threadIdx.x);
  if (adv && blockIdx.x == 0 && threadIdx.x == 0) {
    info->tim += info->dt;
    info->nstep += 1;
  }
}

template <int ODR, int FLUX, int MIX, bool PER>
__global__ __launch_bounds__(256) void k_stage1d(U1d in, U1d ut, U1d out, G1d g, StepInfo *info, double ra, double rb,
                                                 int copy_ghosts, int adv)
{
  stage1d_body<ODR, FLUX, MIX, PER>(in, ut, out, g, info->dtdx, ra, rb, blockIdx.x * 4 + (threadIdx.x >> 6),
                                    threadIdx.x & 63, copy_ghosts != 0);
  if (adv && blockIdx.x == 0 && threadIdx.x == 0) {
    info->tim += info->dt;
    info->nstep += 1;
  }
}

__global__ __launch_bounds__(256) void k_cfl1d(U1d u, G1d g, StepInfo *info)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (i < g.nx) {
    const double vt = cfl1d_cell(u, g, i);
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

template <int ODR, int FLUX, int MIX>
static void step1d(Handle *h)
{
  U1d u0, u1;
  W1d w;
  G1d g;
  views1d(h, u0, u1, w, g);
  const int R_K = h->s.rk, nx = g.nx;
  const double rkf[3][2] = {{0.0, 1.0}, {0.5 + (R_K - 2) * 0.25, 0.5 - (R_K - 2) * 0.25}, {1. / 3., 2. / 3.}};
  const int tiles = (nx - 5 + 62) / 63;
  const bool open = h->s.bc1d == MLAU_BC1D_OPEN;
  const unsigned ub = (nx - 8 + 255) / 256;
  if constexpr (ODR <= 2) {
    if (R_K >= 2 && !(h->tune & 8)) {  // tune bit 3: separate flux and update launches throughout
      U1d u2;
      for (int n = 0; n < 7; n++) u2.f[n] = h->fx[n];
      const U1d buf[3] = {u0, u1, u2};
      const unsigned sb = (unsigned)(((nx - 8 + 61) / 62 + 3) / 4);  // 4 wavefronts of 62 cells per block
      for (int rk = 0; rk < R_K; rk++) {
        const int sin = (rk == 0) ? 0 : (rk == 1 ? 1 : 2);
        const int sout = (rk == R_K - 1) ? 0 : (rk == 0 ? 1 : 2);
        const int adv = rk == R_K - 1;
        if (rk == 0 && open) {
          MLAU_LAUNCH(h, K_SWEEP1D, (k_sweep1d<ODR, FLUX, MIX>), dim3((tiles + 3) / 4), dim3(256), u0, w, g);
          MLAU_LAUNCH(h, K_UPD1D, (k_update1d<ODR, false>), dim3(ub + 1), dim3(256), u0, u0, u1, w, g, h->info,
                      rkf[0][0], rkf[0][1], u0, u1, 1, 0);
        } else if (open) {
          MLAU_LAUNCH(h, K_STAGE1D, (k_stage1d<ODR, FLUX, MIX, false>), dim3(sb), dim3(256), buf[sin], u0, buf[sout], g,
                      h->info, rkf[rk][0], rkf[rk][1], (int)(sout == 2), adv);
        } else {
          MLAU_LAUNCH(h, K_STAGE1D, (k_stage1d<ODR, FLUX, MIX, true>), dim3(sb), dim3(256), buf[sin], u0, buf[sout], g,
                      h->info, rkf[rk][0], rkf[rk][1], 0, adv);
        }
      }
      return;
    }
  }
  for (int rk = 0; rk < R_K; rk++) {
    const U1d &in = (rk == 0) ? u0 : u1;
    const U1d &out = (rk == R_K - 1) ? u0 : u1;
    MLAU_LAUNCH(h, K_SWEEP1D, (k_sweep1d<ODR, FLUX, MIX>), dim3((tiles + 3) / 4), dim3(256), in, w, g);
    const bool obc_here = rk == 0 && open;
    const int obc = obc_here && R_K > 1;  // RK1 updates in place: boundary first, separately
    if (obc_here && !obc)
      MLAU_LAUNCH(h, K_OPENBC1D, k_openbc1d, dim3(1), dim3(64), in, u0, u1, g, (const StepInfo *)h->info);
    const int adv = rk == R_K - 1;
    if (open)
      MLAU_LAUNCH(h, K_UPD1D, (k_update1d<ODR, false>), dim3(ub + obc), dim3(256), in, u0, out, w, g, h->info,
                  rkf[rk][0], rkf[rk][1], u0, u1, obc, adv);
    else
      MLAU_LAUNCH(h, K_UPD1D, (k_update1d<ODR, true>), dim3(ub), dim3(256), in, u0, out, w, g, h->info, rkf[rk][0],
                  rkf[rk][1], u0, u1, 0, adv);
  }
}

typedef void (*Step1Fn)(Handle *);
template <int ODR, int MIX>
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
  switch (precision) {
  case MLAU_MIXED: return pick1f<ODR, MLAU_MIXED>(flux);
  case MLAU_MIXED_RECON: return pick1f<ODR, MLAU_MIXED_RECON>(flux);
  default: return pick1f<ODR, MLAU_FP64>(flux);
  }
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
    fn(h);  // its last update advances the time
    if (cfl_mode) run_cfl_1d(h, cfl, vfloor);
  }
  return (int)hipGetLastError();
}

}  // namespace mlau
