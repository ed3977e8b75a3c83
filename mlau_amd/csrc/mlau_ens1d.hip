// mlau_ens1d.hip -- batched 1D members of the ensemble engine.
//
// Same step structure as the 2D group (plan, fill, CFL, dt, R_K stages,
// advance); a 1D stage is a sweep (primitives formed per stencil point) and an
// update that also runs the open boundary on stage 0 and writes periodic
// ghosts with their source cells, as the single-grid path; at odr <= 2 with
// rk >= 2 the later stages are one launch each (k_ens_stage1d, flux and update
// per wavefront) writing a third state.  Following the 1D driver (1D/SERIAL/shock/main.cpp:
// 113-130) the CFL step is taken once from the initial state over all cells
// (|vx| + c_f) and kept; only the last step is clipped at the end time.
// Tiles run the device bodies of mlau_1d_body.h on each member's arrays.
#include <cstring>

#include "mlau_1d_body.h"
#include "mlau_ens.h"

namespace mlau {

struct EnsRec1D {
  U1d u0, u1;
  U1d u2;  // third state of the fused stages (aliases w.fx, which they do not use)
  W1d w;
  G1d g;
  int flux, bc1d;
};

enum { T1_CFL = 0, T1_SWEEP, T1_OBC, T1_UPD, T1_FUSE, T1_N };

// Kernel variant: 1 = stages after the first fused (odr <= 2, rk >= 2).
static int variant1d(const mlau_member &m, int flags)
{
  return !(flags & MLAU_ENS_UNFUSED) && m.scheme.odr <= 2 && m.scheme.rk >= 2 ? 1 : 0;
}

static long long tiles1d(const mlau_grid &g, int kind, int)
{
  const long long nx = g.nx;
  switch (kind) {
  case T1_CFL: return (nx + 255) / 256;
  case T1_SWEEP: return ((nx - 5 + 62) / 63 + 3) / 4;
  case T1_UPD: return (nx - 8 + 255) / 256;
  default: return 0;  // T1_OBC depends on the boundary kind (tiles1d_member)
  }
}

static size_t arrays1d(const mlau_grid &) { return 7 + 7 + 4 + 7; }

static void record1d(void *out, const mlau_member &m, double *base, size_t per)
{
  EnsRec1D r{};
  size_t k = 0;
  auto take = [&]() { return base + (k++) * per; };
  for (int n = 0; n < 7; n++) r.u0.f[n] = take();
  for (int n = 0; n < 7; n++) r.u1.f[n] = take();
  r.w.vx = take(); r.w.vy = take(); r.w.vz = take(); r.w.pr = take();
  for (int n = 0; n < 7; n++) r.w.fx[n] = take();
  for (int n = 0; n < 7; n++) r.u2.f[n] = r.w.fx[n];
  r.g.nx = m.grid.nx; r.g.xoff = m.grid.xoff; r.g.gamma = m.grid.gamma; r.g.bx = m.grid.bx;
  r.flux = m.scheme.flux;
  r.bc1d = m.scheme.bc1d;
  memcpy(out, &r, sizeof r);
}

static void state1d(const void *rec, double **u0, int *nf, size_t *cells)
{
  const EnsRec1D *r = (const EnsRec1D *)rec;
  for (int n = 0; n < 7; n++) u0[n] = r->u0.f[n];
  *nf = 7;
  *cells = (size_t)r->g.nx;
}

// Tile counts that depend on the member's boundary kind.
static long long tiles1d_member(const mlau_member &m, int kind, int tune, int variant)
{
  if (kind == T1_OBC) return m.scheme.bc1d == MLAU_BC1D_OPEN ? 1 : 0;
  if (kind == T1_FUSE) return variant ? ((m.grid.nx - 8 + 61) / 62 + 3) / 4 : 0;  // 4 wavefronts of 62 cells
  return tiles1d(m.grid, kind, tune);
}

// ------------------------------------------------------------------ kernels
#define ENS_LOOP(KIND)                                                           \
  for (PRange r_ = pers_range(count[KIND]); r_.t < r_.end; r_.t += r_.stride) { \
    const int2 e_ = tiles[r_.t];                                                 \
    const EnsRec1D &R = rec[e_.x];                                               \
    const unsigned l_ = (unsigned)e_.y;

// CFL speed over all cells, first step of each member only.
__global__ __launch_bounds__(256) void k_ens_cfl1d(const EnsRec1D *__restrict__ rec, EnsState *st,
                                                   const int2 *__restrict__ tiles, const int *__restrict__ count)
{
  __shared__ double red[4];
  ENS_LOOP(T1_CFL)
    if (st[e_.x].nstep > 0) continue;  // uniform per tile
    const int i = l_ * 256 + threadIdx.x;
    const double v = (i < R.g.nx) ? cfl1d_cell_checked(R.u0, R.g, i) : 0.0;
    ens_vmax_block(v, &st[e_.x].vmax_bits, red);
  }
}

// instantiated per Riemann solver (a group holds one solver)
template <int ODR, int FLUX, int MIX>
__global__ __launch_bounds__(256) void k_ens_sweep1d(const EnsRec1D *__restrict__ rec,
                                                     const int2 *__restrict__ tiles, const int *__restrict__ count,
                                                     int sin)
{
  ENS_LOOP(T1_SWEEP)
    sweep1d_body<ODR, FLUX, MIX>(sin ? R.u1 : R.u0, R.w, R.g, l_, threadIdx.x);
  }
}

// stage 0 only: the stage input is u0; ghosts go to u0 and u1
__global__ __launch_bounds__(64) void k_ens_openbc1d(const EnsRec1D *__restrict__ rec, const EnsState *st,
                                                     const int2 *__restrict__ tiles, const int *__restrict__ count)
{
  ENS_LOOP(T1_OBC)
    if (st[e_.x].halt) continue;  // member failed at this step's CFL check
    (void)l_;
    openbc1d_body(R.u0, R.u0, R.u1, R.g, st[e_.x].dtdx, threadIdx.x);
  }
}

// obc (stage 0 of RK2/RK3): the open-boundary tiles run in the same launch
// (the boundary reads u0 and writes ghosts only; the update reads interior
// cells of u0 and writes u1).
template <int ODR>
__global__ __launch_bounds__(256) void k_ens_update1d(const EnsRec1D *__restrict__ rec, const EnsState *st,
                                                      const int2 *__restrict__ tiles, const int2 *__restrict__ otiles,
                                                      const int *__restrict__ count, int sin, int sout, double ra,
                                                      double rb, int obc)
{
  ENS_LOOP(T1_UPD)
    if (st[e_.x].halt) continue;  // member failed at this step's CFL check
    const U1d in = sin ? R.u1 : R.u0, out = sout ? R.u1 : R.u0;
    if (R.bc1d == MLAU_BC1D_PERIODIC)
      update1d_body<ODR, true>(in, R.u0, out, R.w, R.g, st[e_.x].dtdx, ra, rb, l_, threadIdx.x);
    else
      update1d_body<ODR, false>(in, R.u0, out, R.w, R.g, st[e_.x].dtdx, ra, rb, l_, threadIdx.x);
  }
  if (!obc) return;
  tiles = otiles;
  ENS_LOOP(T1_OBC)
    if (st[e_.x].halt) continue;  // member failed at this step's CFL check
    (void)l_;
    openbc1d_body(R.u0, R.u0, R.u1, R.g, st[e_.x].dtdx, threadIdx.x);
  }
}
// stages after the first: tile = 4 wavefronts of stage1d_body (62 cells each);
// sin/sout: 0 = u0, 1 = u1, 2 = u2
__device__ __forceinline__ U1d ubuf1(const EnsRec1D &R, int s) { return s == 0 ? R.u0 : (s == 1 ? R.u1 : R.u2); }

template <int ODR, int FLUX, int MIX>
__global__ __launch_bounds__(256) void k_ens_stage1d(const EnsRec1D *__restrict__ rec, const EnsState *st,
                                                     const int2 *__restrict__ tiles, const int *__restrict__ count,
                                                     int sin, int sout, double ra, double rb)
{
  ENS_LOOP(T1_FUSE)
    if (st[e_.x].halt) continue;  // member failed at this step's CFL check
    const U1d in = ubuf1(R, sin), out = ubuf1(R, sout);
    const double dtdx = st[e_.x].dtdx;
    const unsigned wave = l_ * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (R.bc1d == MLAU_BC1D_PERIODIC)
      stage1d_body<ODR, FLUX, MIX, true>(in, R.u0, out, R.g, dtdx, ra, rb, wave, lane, false);
    else  // the open boundary's ghosts, fixed over the step, go into u2 with it
      stage1d_body<ODR, FLUX, MIX, false>(in, R.u0, out, R.g, dtdx, ra, rb, wave, lane, sout == 2);
  }
}
#undef ENS_LOOP

// ------------------------------------------------------------------- driver
static unsigned pers_blocks1(const Ensemble *e)
{
  const int m = (e->tune >> 4) & 15;
  return (unsigned)((m ? m : 2) * ens_cus(e));
}

template <int ODR, int MIX>
static void sweep1d(Ensemble *e, int flux, unsigned pb, const EnsRec1D *rec, const EnsPlan &p, int sin)
{
  const int2 *t = (const int2 *)p.tiles[T1_SWEEP];
  const int *cnt = p.count;
  const dim3 b(256);
  switch (flux) {
  case FLUX_MLAU: ENS_LAUNCH(e, E_XSWEEP, (k_ens_sweep1d<ODR, FLUX_MLAU, MIX>), dim3(pb), b, rec, t, cnt, sin); break;
  case FLUX_ROE: ENS_LAUNCH(e, E_XSWEEP, (k_ens_sweep1d<ODR, FLUX_ROE, MIX>), dim3(pb), b, rec, t, cnt, sin); break;
  case FLUX_HLLD: ENS_LAUNCH(e, E_XSWEEP, (k_ens_sweep1d<ODR, FLUX_HLLD, MIX>), dim3(pb), b, rec, t, cnt, sin); break;
  default: ENS_LAUNCH(e, E_XSWEEP, (k_ens_sweep1d<ODR, FLUX_LHLLD, MIX>), dim3(pb), b, rec, t, cnt, sin); break;
  }
}

template <int ODR, int MIX>
static void stage1d(Ensemble *e, int flux, unsigned pb, const EnsRec1D *rec, const EnsPlan &p, int sin, int sout,
                    double ra, double rb)
{
  const int2 *t = (const int2 *)p.tiles[T1_FUSE];
  const int *cnt = p.count;
  const EnsState *st = p.st;
  const dim3 b(256);
  const dim3 g(pb);
  switch (flux) {
  case FLUX_MLAU:
    ENS_LAUNCH(e, E_STAGE, (k_ens_stage1d<ODR, FLUX_MLAU, MIX>), g, b, rec, st, t, cnt, sin, sout, ra, rb);
    break;
  case FLUX_ROE:
    ENS_LAUNCH(e, E_STAGE, (k_ens_stage1d<ODR, FLUX_ROE, MIX>), g, b, rec, st, t, cnt, sin, sout, ra, rb);
    break;
  case FLUX_HLLD:
    ENS_LAUNCH(e, E_STAGE, (k_ens_stage1d<ODR, FLUX_HLLD, MIX>), g, b, rec, st, t, cnt, sin, sout, ra, rb);
    break;
  default:
    ENS_LAUNCH(e, E_STAGE, (k_ens_stage1d<ODR, FLUX_LHLLD, MIX>), g, b, rec, st, t, cnt, sin, sout, ra, rb);
    break;
  }
}

template <int ODR, int MIX>
static int step1d_group(Ensemble *e, EnsGroup *G)
{
  const EnsPlan &p = G->p;
  const EnsRec1D *rec = (const EnsRec1D *)G->rec;
  const unsigned pb = pers_blocks1(e);
  const dim3 b256(256), b64(64), slots(kEnsSlots / 256);
  const int *cnt = p.count;
  ENS_LAUNCH(e, E_PLAN, k_ens_plan, dim3(1), dim3(kPlanThreads), p, (int)T1_N);
  ENS_LAUNCH(e, E_FILL, k_ens_fill, dim3(kEnsSlots), b256, p, (int)T1_N);
  ENS_LAUNCH(e, E_CFL, k_ens_cfl1d, dim3(pb), b256, rec, p.st, (const int2 *)p.tiles[T1_CFL], cnt);
  ENS_LAUNCH(e, E_DT, k_ens_dt, slots, b256, p);
  const int R_K = G->rk;
  const double rkf[3][2] = {{0.0, 1.0}, {0.5 + (R_K - 2) * 0.25, 0.5 - (R_K - 2) * 0.25}, {1. / 3., 2. / 3.}};
  for (int rk = 0; rk < R_K; rk++) {
    if constexpr (ODR <= 2) {
      if (G->variant && rk > 0) {  // u1 -> u2 -> u0 (rk 3) or u1 -> u0 (rk 2)
        stage1d<ODR, MIX>(e, G->flux, pb, rec, p, rk == 1 ? 1 : 2, rk == R_K - 1 ? 0 : 2, rkf[rk][0], rkf[rk][1]);
        continue;
      }
    }
    const int sin = (rk == 0) ? 0 : 1;
    const int sout = (rk == R_K - 1) ? 0 : 1;
    sweep1d<ODR, MIX>(e, G->flux, pb, rec, p, sin);
    const int obc = rk == 0 && R_K > 1;  // RK1 updates in place: boundary first, separately
    if (rk == 0 && !obc)
      ENS_LAUNCH(e, E_BC, k_ens_openbc1d, dim3(e->cus), b64, rec, (const EnsState *)p.st,
                 (const int2 *)p.tiles[T1_OBC], cnt);
    ENS_LAUNCH(e, E_UPD, (k_ens_update1d<ODR>), dim3(pb), b256, rec, (const EnsState *)p.st,
               (const int2 *)p.tiles[T1_UPD], (const int2 *)p.tiles[T1_OBC], cnt, sin, sout, rkf[rk][0], rkf[rk][1],
               obc);
  }
  ENS_LAUNCH(e, E_ADVANCE, k_ens_advance, slots, b256, p);
  return (int)hipGetLastError();
}

typedef int (*EnsStepFn)(Ensemble *, EnsGroup *);

template <int ODR>
static EnsStepFn pick1_odr(int precision)
{
  switch (precision) {
  case MLAU_MIXED: return step1d_group<ODR, MLAU_MIXED>;
  case MLAU_MIXED_RECON: return step1d_group<ODR, MLAU_MIXED_RECON>;
  default: return step1d_group<ODR, MLAU_FP64>;
  }
}

static int launch1d(Ensemble *e, EnsGroup *G)
{
  EnsStepFn fn;
  switch (G->odr) {
  case 1: fn = pick1_odr<1>(G->precision); break;
  case 2: fn = pick1_odr<2>(G->precision); break;
  case 3: fn = pick1_odr<3>(G->precision); break;
  default: fn = pick1_odr<4>(G->precision); break;
  }
  return fn(e, G);
}

static const EnsDim kDim1d = {sizeof(EnsRec1D), T1_N, tiles1d_member, arrays1d, record1d, launch1d, state1d,
                              variant1d};
const EnsDim *ens_dim1d() { return &kDim1d; }

}  // namespace mlau
