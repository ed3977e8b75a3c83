// mlau_ens2d.hip -- batched 2D members of the ensemble engine.
//
// One step of a 2D group is: plan (compact running members into work lists),
// fill (rewrite the lists when the running set changed), CFL speed per member,
// dt per member (clipped at its end time), R_K stages of
//   prim+aux (one pass), x-sweep, y-sweep, ez, update, bc
// and advance (time, step count, status).  The stage kernels are persistent:
// a fixed grid walks the tile list of its kind; each tile is one logical block
// of the single-grid kernel of mlau_2d.hip, run through the same device body
// (mlau_2d_body.h) on that member's arrays, so a member's bits do not depend on
// what else runs.  The Riemann solver is chosen per member (per tile).
// Reference: per member, 2D/SERIAL/OTvortex/mhd_fd_ct_2d.c:91-660 per step and
// cflcheck.hpp:1-32 for the time step.
#include <cstring>

#include "mlau_2d_body.h"
#include "mlau_ens.h"

namespace mlau {

struct EnsRec2D {
  U2 u0, u1;
  W2 w;
  G2 g;
  BcFlags bc;
  int flux, pad;
};

enum { T2_CFL = 0, T2_CELL, T2_XSW, T2_YSW, T2_EZ, T2_UPD, T2_BC, T2_N };

static int ych_of(int tune)
{
  switch (tune & 3) {
  case 1: return 16;
  case 2: return 4;
  default: return 8;
  }
}

static long long tiles2d(const mlau_grid &g, int kind, int tune)
{
  const long long nx = g.nx, ny = g.ny, xo = g.xoff, yo = g.yoff;
  switch (kind) {
  case T2_CFL: return ((nx - 2 * xo) * (ny - 2 * yo) + 255) / 256;
  case T2_CELL: return (nx * ny + 255) / 256;
  case T2_XSW: return (xsweep_waves(ny, (int)nx) + 3) / 4;
  case T2_YSW: {
    const int ch = ych_of(tune);
    return ((nx + 63) / 64) * ((ny - 5 + ch - 1) / ch);
  }
  case T2_EZ: return ((nx - 5) * (ny - 5) + 255) / 256;
  case T2_UPD: return ((nx - 2 * xo + 1) * (ny - 2 * yo + 1) + 255) / 256;
  case T2_BC: return ((ny * 2 * xo + 2 * yo * nx + 255) / 256) * 8;
  default: return 0;
  }
}

static long long tiles2d_member(const mlau_member &m, int kind, int tune) { return tiles2d(m.grid, kind, tune); }

static size_t arrays2d(const mlau_grid &) { return 8 + 8 + 7 + 3 + 8 + 8 + 1; }

static void record2d(void *out, const mlau_member &m, double *base, size_t per)
{
  EnsRec2D r{};
  size_t k = 0;
  auto take = [&]() { return base + (k++) * per; };
  for (int n = 0; n < 8; n++) r.u0.f[n] = take();
  for (int n = 0; n < 8; n++) r.u1.f[n] = take();
  r.w.vx = take(); r.w.vy = take(); r.w.vz = take(); r.w.pr = take(); r.w.m2 = take();
  r.w.cx = take(); r.w.cy = take();
  r.w.ctx = take(); r.w.dvx = take(); r.w.dvy = take();
  for (int n = 0; n < 7; n++) r.w.fx[n] = take();
  r.w.fcx = take();
  for (int n = 0; n < 7; n++) r.w.fy[n] = take();
  r.w.fcy = take();
  r.w.ez = take();
  r.g.nx = m.grid.nx; r.g.ny = m.grid.ny; r.g.xoff = m.grid.xoff; r.g.yoff = m.grid.yoff;
  r.g.gamma = m.grid.gamma;
  for (int n = 0; n < 8; n++) {
    r.bc.dnx[n] = m.scheme.dnxs[n];
    r.bc.dny[n] = m.scheme.dnys[n];
    r.bc.dnz[n] = m.scheme.dnzs[n];
    r.bc.stx[n] = (n == 5);
    r.bc.sty[n] = (n == 6);
    r.bc.stz[n] = (n == 7);
  }
  r.flux = m.scheme.flux;
  memcpy(out, &r, sizeof r);
}

static void state2d(const void *rec, double **u0, int *nf, size_t *cells)
{
  const EnsRec2D *r = (const EnsRec2D *)rec;
  for (int n = 0; n < 8; n++) u0[n] = r->u0.f[n];
  *nf = 8;
  *cells = (size_t)r->g.nx * r->g.ny;
}

// ------------------------------------------------------------------ kernels
// sin/sout select the stage input/output state: 0 = u0, 1 = u1.
#define ENS_LOOP(KIND)                                         \
  for (PRange r_ = pers_range(count[KIND]); r_.t < r_.end; r_.t += r_.stride) { \
    const int2 e_ = tiles[r_.t];                               \
    const EnsRec2D &R = rec[e_.x];                             \
    const unsigned l_ = (unsigned)e_.y;

__global__ __launch_bounds__(256) void k_ens_cfl2d(const EnsRec2D *__restrict__ rec, EnsState *st,
                                                   const int2 *__restrict__ tiles, const int *__restrict__ count)
{
  __shared__ double red[4];
  ENS_LOOP(T2_CFL)
    const long total = (long)(R.g.nx - 2 * R.g.xoff) * (R.g.ny - 2 * R.g.yoff);
    const long t = (long)l_ * 256 + threadIdx.x;
    double v = (t < total) ? cfl2d_cell(R.u0, R.g, t) : 0.0;
    // max over bit patterns: equal to the double max for v >= 0, and a NaN
    // (larger bits than +Inf) is kept, so a broken member reports it
    for (int o = 32; o > 0; o >>= 1) {
      const double q = __shfl_xor(v, o, 64);
      if ((unsigned long long)__double_as_longlong(q) > (unsigned long long)__double_as_longlong(v)) v = q;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long b = (unsigned long long)__double_as_longlong(red[0]);
      for (int k = 1; k < 4; k++) {
        const unsigned long long c = (unsigned long long)__double_as_longlong(red[k]);
        if (c > b) b = c;
      }
      if (b) atomicMax(&st[e_.x].vmax_bits, b);
    }
    __syncthreads();
  }
}

template <int ODR, int CTW>
__global__ __launch_bounds__(256) void k_ens_primaux2d(const EnsRec2D *__restrict__ rec,
                                                       const int2 *__restrict__ tiles, const int *__restrict__ count,
                                                       int sin)
{
  ENS_LOOP(T2_CELL)
    int i, j;
    if (!region_cell(R.g.nx, R.g.ny, 0, 0, l_, threadIdx.x, i, j)) continue;
    primaux2d_body<ODR, CTW>(sin ? R.u1 : R.u0, R.w, R.g, R.bc, i, j);
  }
}

// Sweeps are instantiated per Riemann solver (a group holds one solver), so
// each kernel is allocated the registers of its own solver only.
template <int ODR, int FLUX, int MIX>
__global__ __launch_bounds__(256) void k_ens_xsweep2d(const EnsRec2D *__restrict__ rec,
                                                      const int2 *__restrict__ tiles, const int *__restrict__ count,
                                                      int sin)
{
  ENS_LOOP(T2_XSW)
    const U2 u = sin ? R.u1 : R.u0;
    const long W = (long)l_ * 4 + (threadIdx.x >> 6);
    if (W >= xsweep_waves(R.g.ny, R.g.nx)) continue;  // whole wavefront
    xsweep2d_body<ODR, FLUX, MIX>(u, R.w, R.g, W, threadIdx.x & 63);
  }
}

template <int ODR, int FLUX, int CH, int MIX>
__global__ __launch_bounds__(64) MLAU_TWO_WAVES void k_ens_ysweep2d(const EnsRec2D *__restrict__ rec,
                                                     const int2 *__restrict__ tiles, const int *__restrict__ count,
                                                     int sin)
{
  ENS_LOOP(T2_YSW)
    const U2 u = sin ? R.u1 : R.u0;
    const unsigned gyx = (R.g.nx + 63) / 64;
    ysweep2d_body<ODR, FLUX, CH, MIX>(u, R.w, R.g, l_ % gyx, l_ / gyx, threadIdx.x);
  }
}

template <int ODR>
__global__ __launch_bounds__(256) void k_ens_ez2d(const EnsRec2D *__restrict__ rec, const int2 *__restrict__ tiles,
                                                  const int *__restrict__ count)
{
  ENS_LOOP(T2_EZ)
    int i, j;
    if (!region_cell(R.g.nx - 5, R.g.ny - 5, 3, 3, l_, threadIdx.x, i, j)) continue;
    ez2d_body<ODR>(R.w, R.g, i, j);
  }
}

template <int ODR, bool FIRST>
__global__ __launch_bounds__(256) void k_ens_update2d(const EnsRec2D *__restrict__ rec, const EnsState *st,
                                                      const int2 *__restrict__ tiles, const int *__restrict__ count,
                                                      int sin, int sout, double ra, double rb)
{
  ENS_LOOP(T2_UPD)
    const double dtdx = st[e_.x].dtdx, dtdy = st[e_.x].dtdy;
    int i, j;
    if (!region_cell(R.g.nx - 2 * R.g.xoff + 1, R.g.ny - 2 * R.g.yoff + 1, R.g.xoff, R.g.yoff, l_, threadIdx.x, i, j))
      continue;
    update2d_body<ODR, FIRST>(sin ? R.u1 : R.u0, R.u0, sout ? R.u1 : R.u0, R.w, R.g, dtdx, dtdy, ra, rb, i, j);
  }
}

__global__ __launch_bounds__(256) void k_ens_bc2d(const EnsRec2D *__restrict__ rec, const int2 *__restrict__ tiles,
                                                  const int *__restrict__ count, int sout)
{
  ENS_LOOP(T2_BC)
    const unsigned nbf = (unsigned)((bc2d_points(R.g) + 255) / 256);
    const int fld = l_ / nbf;
    bc2d_body(sout ? R.u1 : R.u0, R.g, R.bc, fld, (long)(l_ % nbf) * 256 + threadIdx.x);
  }
}
#undef ENS_LOOP

// ------------------------------------------------------------------- driver
static unsigned pers_blocks(const Ensemble *e)
{
  const int m = (e->tune >> 4) & 15;
  return (unsigned)((m ? m : 2) * e->cus);
}

template <int ODR, int FLUX, int CH, int MIX>
static void sweep2d_launch(Ensemble *e, int y, unsigned pb, const EnsRec2D *rec, const EnsPlan &p, int sin)
{
  if (y)
    ENS_LAUNCH(e, E_YSWEEP, (k_ens_ysweep2d<ODR, FLUX, CH, MIX>), dim3(pb * 4), dim3(64), rec,
               (const int2 *)p.tiles[T2_YSW], (const int *)p.count, sin);
  else
    ENS_LAUNCH(e, E_XSWEEP, (k_ens_xsweep2d<ODR, FLUX, MIX>), dim3(pb), dim3(256), rec, (const int2 *)p.tiles[T2_XSW],
               (const int *)p.count, sin);
}

template <int ODR, int CH, int MIX>
static void sweep2d(Ensemble *e, int flux, int y, unsigned pb, const EnsRec2D *rec, const EnsPlan &p, int sin)
{
  switch (flux) {
  case FLUX_MLAU: sweep2d_launch<ODR, FLUX_MLAU, CH, MIX>(e, y, pb, rec, p, sin); break;
  case FLUX_ROE: sweep2d_launch<ODR, FLUX_ROE, CH, MIX>(e, y, pb, rec, p, sin); break;
  case FLUX_HLLD: sweep2d_launch<ODR, FLUX_HLLD, CH, MIX>(e, y, pb, rec, p, sin); break;
  default: sweep2d_launch<ODR, FLUX_LHLLD, CH, MIX>(e, y, pb, rec, p, sin); break;
  }
}

template <int ODR, int CTW, int CH, int MIX>
static int step2d_group(Ensemble *e, EnsGroup *G)
{
  const EnsPlan &p = G->p;
  const EnsRec2D *rec = (const EnsRec2D *)G->rec;
  const unsigned pb = pers_blocks(e);
  const dim3 b256(256), slots(kEnsSlots / 256);
  ENS_LAUNCH(e, E_PLAN, k_ens_plan, dim3(1), dim3(kPlanThreads), p, (int)T2_N);
  ENS_LAUNCH(e, E_FILL, k_ens_fill, dim3(kEnsSlots), b256, p, (int)T2_N);
  ENS_LAUNCH(e, E_CFL, k_ens_cfl2d, dim3(pb), b256, rec, p.st, (const int2 *)p.tiles[T2_CFL], (const int *)p.count);
  ENS_LAUNCH(e, E_DT, k_ens_dt, slots, b256, p);
  const int R_K = G->rk;
  const double rkf[3][2] = {{0.0, 1.0}, {0.5 + (R_K - 2) * 0.25, 0.5 - (R_K - 2) * 0.25}, {1. / 3., 2. / 3.}};
  const int *cnt = p.count;
  const bool yfirst = (e->tune >> 2) & 1;
  for (int rk = 0; rk < R_K; rk++) {
    const int sin = (rk == 0) ? 0 : 1;
    const int sout = (rk == R_K - 1) ? 0 : 1;
    const double ra = rkf[rk][0], rb = rkf[rk][1];
    ENS_LAUNCH(e, E_PRIM, (k_ens_primaux2d<ODR, CTW>), dim3(pb), b256, rec, (const int2 *)p.tiles[T2_CELL], cnt,
               sin);
    for (int pass = 0; pass < 2; pass++) sweep2d<ODR, CH, MIX>(e, G->flux, (pass == 0) == yfirst, pb, rec, p, sin);
    ENS_LAUNCH(e, E_EZ, (k_ens_ez2d<ODR>), dim3(pb), b256, rec, (const int2 *)p.tiles[T2_EZ], cnt);
    if (ra == 0.0)
      ENS_LAUNCH(e, E_UPD, (k_ens_update2d<ODR, true>), dim3(pb), b256, rec, (const EnsState *)p.st,
                 (const int2 *)p.tiles[T2_UPD], cnt, sin, sout, ra, rb);
    else
      ENS_LAUNCH(e, E_UPD, (k_ens_update2d<ODR, false>), dim3(pb), b256, rec, (const EnsState *)p.st,
                 (const int2 *)p.tiles[T2_UPD], cnt, sin, sout, ra, rb);
    ENS_LAUNCH(e, E_BC, k_ens_bc2d, dim3(pb), b256, rec, (const int2 *)p.tiles[T2_BC], cnt, sout);
  }
  ENS_LAUNCH(e, E_ADVANCE, k_ens_advance, slots, b256, p);
  return (int)hipGetLastError();
}

typedef int (*EnsStepFn)(Ensemble *, EnsGroup *);

template <int ODR, int CTW, int MIX>
static EnsStepFn pick_ch(int tune)
{
  switch (ych_of(tune)) {
  case 16: return step2d_group<ODR, CTW, 16, MIX>;
  case 4: return step2d_group<ODR, CTW, 4, MIX>;
  default: return step2d_group<ODR, CTW, 8, MIX>;
  }
}
template <int ODR>
static EnsStepFn pick_odr(int ctw, int precision, int tune)
{
  switch (precision) {
  case MLAU_MIXED: return ctw ? pick_ch<ODR, 1, MLAU_MIXED>(tune) : pick_ch<ODR, 0, MLAU_MIXED>(tune);
  case MLAU_MIXED_RECON:
    return ctw ? pick_ch<ODR, 1, MLAU_MIXED_RECON>(tune) : pick_ch<ODR, 0, MLAU_MIXED_RECON>(tune);
  default: return ctw ? pick_ch<ODR, 1, MLAU_FP64>(tune) : pick_ch<ODR, 0, MLAU_FP64>(tune);
  }
}

static int launch2d(Ensemble *e, EnsGroup *G)
{
  EnsStepFn fn;
  switch (G->odr) {
  case 1: fn = pick_odr<1>(G->ctw, G->precision, e->tune); break;
  case 2: fn = pick_odr<2>(G->ctw, G->precision, e->tune); break;
  case 3: fn = pick_odr<3>(G->ctw, G->precision, e->tune); break;
  default: fn = pick_odr<4>(G->ctw, G->precision, e->tune); break;
  }
  return fn(e, G);
}

static const EnsDim kDim2d = {sizeof(EnsRec2D), T2_N, tiles2d_member, arrays2d, record2d, launch2d, state2d};
const EnsDim *ens_dim2d() { return &kDim2d; }

}  // namespace mlau
