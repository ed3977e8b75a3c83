// mlau_ens3d.hip -- batched 3D members of the ensemble engine.
//
// Same step structure as the 2D group; a 3D stage is prim, aux, x-, y- and
// z-sweeps, edge E-fields, update and ghost cells, each a persistent kernel
// over its work list running the device bodies of mlau_3d_body.h on each
// member's arrays.  CFL per step as 3D/MPI/OTvortex/cflcheck.hpp:1-39.
#include <cstring>

#include "mlau_3d_body.h"
#include "mlau_ens.h"

namespace mlau {

constexpr int kEnsMChunk = 16;  // y/z sweep chunk (the single-grid path's)
constexpr int kEnsZChunk = 8;   // planes per thread of the z-marching edge/update kernels

struct EnsRec3D {
  U3 u0, u1;
  double *prim[8];          // vx,vy,vz,pr,m2,cx,cy,cz
  double *aux[6];           // ct0,ct1,ct2,dvx,dvy,dvz
  double *fl[3][7];         // x, y, z sweep outputs
  double *fc[3][2];         // central induction parts
  double *e[3];             // ex, ey, ez
  G3 g;
  BcFlags bc;
  int flux, pad;
};

enum { T3_CFL = 0, T3_CELL, T3_XSW, T3_YSW, T3_ZSW, T3_UPD, T3_BC, T3_EDGE, T3_N };

static long long tiles3d(const mlau_member &m, int kind, int, int)
{
  const mlau_grid &g = m.grid;
  const long long nx = g.nx, ny = g.ny, nz = g.nz, xo = g.xoff, yo = g.yoff, zo = g.zoff;
  const long long ch = kEnsMChunk;
  switch (kind) {
  case T3_CFL: return ((nx - 2 * xo) * (ny - 2 * yo) * (nz - 2 * zo) + 255) / 256;
  case T3_CELL: return ((nx * ny + 255) / 256) * nz;
  case T3_XSW: return (xsweep_waves(ny * nz, (int)nx) + 3) / 4;
  case T3_YSW: return ((nx * nz + 63) / 64) * ((ny - 5 + ch - 1) / ch);
  case T3_ZSW: return ((nx * ny + 63) / 64) * ((nz - 5 + ch - 1) / ch);
  case T3_UPD:
    return (((nx - 2 * xo + 1) * (ny - 2 * yo + 1) + 255) / 256) * ((nz - 2 * zo + 1 + kEnsZChunk - 1) / kEnsZChunk);
  case T3_EDGE: return ((nx * ny + 255) / 256) * ((nz + kEnsZChunk - 1) / kEnsZChunk);
  case T3_BC: return ((2 * xo * ny * nz + nx * 2 * yo * nz + nx * ny * 2 * zo + 255) / 256) * 8;
  default: return 0;
  }
}

static size_t arrays3d(const mlau_grid &) { return 8 + 8 + 8 + 6 + 21 + 6 + 3; }

static void record3d(void *out, const mlau_member &m, double *base, size_t per)
{
  EnsRec3D r{};
  size_t k = 0;
  auto take = [&]() { return base + (k++) * per; };
  for (int n = 0; n < 8; n++) r.u0.f[n] = take();
  for (int n = 0; n < 8; n++) r.u1.f[n] = take();
  for (int n = 0; n < 8; n++) r.prim[n] = take();
  for (int n = 0; n < 6; n++) r.aux[n] = take();
  for (int a = 0; a < 3; a++)
    for (int n = 0; n < 7; n++) r.fl[a][n] = take();
  for (int a = 0; a < 3; a++)
    for (int n = 0; n < 2; n++) r.fc[a][n] = take();
  for (int n = 0; n < 3; n++) r.e[n] = take();
  const mlau_grid &g = m.grid;
  r.g.nx = g.nx; r.g.ny = g.ny; r.g.nz = g.nz;
  r.g.xoff = g.xoff; r.g.yoff = g.yoff; r.g.zoff = g.zoff;
  r.g.gamma = g.gamma;
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

static void state3d(const void *rec, double **u0, int *nf, size_t *cells)
{
  const EnsRec3D *r = (const EnsRec3D *)rec;
  for (int n = 0; n < 8; n++) u0[n] = r->u0.f[n];
  *nf = 8;
  *cells = (size_t)r->g.nx * r->g.ny * r->g.nz;
}

// Rotated view of one sweep direction (same roles as the single-grid path):
// (n, t, u) = (x, y, z), (y, z, x), (z, x, y).
__device__ __forceinline__ Sw3 sweep_view_dev(const EnsRec3D &R, const U3 &in, int ax)
{
  const int n = ax, t = (ax + 1) % 3, u = (ax + 2) % 3;
  Sw3 w;
  w.ro = in.f[0];
  w.vn = R.prim[n];
  w.vt = R.prim[t];
  w.vu = R.prim[u];
  w.bt = R.prim[5 + t];
  w.bu = R.prim[5 + u];
  w.pr = R.prim[3];
  w.m2 = R.prim[4];
  w.bnc = R.prim[5 + n];
  w.bface = in.f[5 + n];
  w.dvt = R.aux[3 + t];
  w.dvu = R.aux[3 + u];
  for (int q = 0; q < 7; q++) w.out[q] = R.fl[ax][q];
  w.fc[0] = R.fc[ax][0];
  w.fc[1] = R.fc[ax][1];
  return w;
}

// ------------------------------------------------------------------ kernels
#define ENS_LOOP(KIND)                                                           \
  for (PRange r_ = pers_range(count[KIND]); r_.t < r_.end; r_.t += r_.stride) { \
    const int2 e_ = tiles[r_.t];                                                 \
    const EnsRec3D &R = rec[e_.x];                                               \
    const unsigned l_ = (unsigned)e_.y;

__global__ __launch_bounds__(256) void k_ens_cfl3d(const EnsRec3D *__restrict__ rec, EnsState *st,
                                                   const int2 *__restrict__ tiles, const int *__restrict__ count)
{
  __shared__ double red[4];
  ENS_LOOP(T3_CFL)
    const long total = (long)(R.g.nx - 2 * R.g.xoff) * (R.g.ny - 2 * R.g.yoff) * (R.g.nz - 2 * R.g.zoff);
    const long t = (long)l_ * 256 + threadIdx.x;
    const double v = (t < total) ? cfl3d_cell_checked(R.u0, R.g, t) : 0.0;
    ens_vmax_block(v, &st[e_.x].vmax_bits, red);
  }
}

// cell tiles: l = plane * pieces + piece (256-cell pieces of the flattened plane)
#define CELL_COORDS                                                 \
  const unsigned np_ = plane_pieces(R.g.nx, R.g.ny);                \
  int i_, j_;                                                       \
  const int k_ = l_ / np_;                                          \
  if (!plane_cell(R.g, l_ % np_, threadIdx.x, i_, j_)) continue;

template <int ODR>
__global__ __launch_bounds__(256) void k_ens_prim3d(const EnsRec3D *__restrict__ rec, const int2 *__restrict__ tiles,
                                                    const int *__restrict__ count, int sin)
{
  ENS_LOOP(T3_CELL)
    CELL_COORDS
    prim3d_body<ODR>(sin ? R.u1 : R.u0, R.prim[0], R.prim[1], R.prim[2], R.prim[3], R.prim[4], R.prim[5], R.prim[6],
                     R.prim[7], R.g, R.bc, i_, j_, k_);
  }
}

// z-marching columns of kEnsZChunk cells: tiles of the EDGE list (same shape)
template <int CTW>
__global__ __launch_bounds__(256) void k_ens_aux3d(const EnsRec3D *__restrict__ rec, const int2 *__restrict__ tiles,
                                                   const int *__restrict__ count, int sin)
{
  ENS_LOOP(T3_EDGE)
    const unsigned np = plane_pieces(R.g.nx, R.g.ny);
    int i, j;
    if (!plane_cell(R.g, l_ % np, threadIdx.x, i, j)) continue;
    aux3d_zm_body<CTW, kEnsZChunk>(sin ? R.u1 : R.u0, R.prim[0], R.prim[1], R.prim[2], R.aux[0], R.aux[1], R.aux[2],
                                   R.aux[3], R.aux[4], R.aux[5], R.g, R.bc, i, j, (int)(l_ / np));
  }
}

// Sweeps are instantiated per Riemann solver (a group holds one solver), so
// each kernel is allocated the registers of its own solver only.
template <int ODR, int FLUX, int MIX>
__global__ __launch_bounds__(256) MLAU_TWO_WAVES void k_ens_xsweep3d(const EnsRec3D *__restrict__ rec,
                                                      const int2 *__restrict__ tiles, const int *__restrict__ count,
                                                      int sin)
{
  ENS_LOOP(T3_XSW)
    const Sw3 w = sweep_view_dev(R, sin ? R.u1 : R.u0, 0);
    const long W = (long)l_ * 4 + (threadIdx.x >> 6);
    if (W >= xsweep_waves((long)R.g.ny * R.g.nz, R.g.nx)) continue;  // whole wavefront
    xsweep3d_body<ODR, FLUX, MIX>(w, R.g, W, threadIdx.x & 63);
  }
}

// AX = 1: tiles (bx, chunk, k); AX = 2: tiles (bx, chunk, j)
template <int ODR, int FLUX, int AX, int MIX>
__global__ __launch_bounds__(64) MLAU_TWO_WAVES void k_ens_msweep3d(const EnsRec3D *__restrict__ rec,
                                                     const int2 *__restrict__ tiles, const int *__restrict__ count,
                                                     int sin)
{
  ENS_LOOP(AX == 1 ? T3_YSW : T3_ZSW)
    const Sw3 w = sweep_view_dev(R, sin ? R.u1 : R.u0, AX);
    const long nc = (AX == 1) ? R.g.nz : R.g.ny;
    const unsigned pieces = (unsigned)((R.g.nx * nc + 63) / 64);
    const long p = (long)(l_ % pieces) * 64 + threadIdx.x;
    if (p >= R.g.nx * nc) continue;
    msweep3d_body<ODR, FLUX, kEnsMChunk, AX, MIX>(w, R.g, (int)(p % R.g.nx), l_ / pieces, (int)(p / R.g.nx));
  }
}

template <int ODR>
__global__ __launch_bounds__(256) void k_ens_edge3d(const EnsRec3D *__restrict__ rec, const int2 *__restrict__ tiles,
                                                    const int *__restrict__ count)
{
  ENS_LOOP(T3_EDGE)
    const unsigned np = plane_pieces(R.g.nx, R.g.ny);
    const int kc = l_ / np;
    int i, j;
    if (!plane_cell(R.g, l_ % np, threadIdx.x, i, j)) continue;
    Edge3 e;
    e.fxt = R.fl[0][4]; e.fxu = R.fl[0][5]; e.fcx0 = R.fc[0][0]; e.fcx1 = R.fc[0][1];
    e.fyt = R.fl[1][4]; e.fyu = R.fl[1][5]; e.fcy0 = R.fc[1][0]; e.fcy1 = R.fc[1][1];
    e.fzt = R.fl[2][4]; e.fzu = R.fl[2][5]; e.fcz0 = R.fc[2][0]; e.fcz1 = R.fc[2][1];
    e.ct0 = R.aux[0]; e.ct1 = R.aux[1]; e.ct2 = R.aux[2];
    e.ex = R.e[0]; e.ey = R.e[1]; e.ez = R.e[2];
    edge3d_zm_body<ODR, kEnsZChunk>(e, R.g, i, j, kc);
  }
}

template <int ODR>
__global__ __launch_bounds__(256) void k_ens_update3d(const EnsRec3D *__restrict__ rec, const EnsState *st,
                                                      const int2 *__restrict__ tiles, const int *__restrict__ count,
                                                      int sin, int sout, double ra, double rb)
{
  ENS_LOOP(T3_UPD)
    if (st[e_.x].halt) continue;  // member failed at this step's CFL check
    const unsigned np = plane_pieces(R.g.nx - 2 * R.g.xoff + 1, R.g.ny - 2 * R.g.yoff + 1);
    const int kc = l_ / np;
    int i, j;
    if (!plane_face_cell(R.g, l_ % np, threadIdx.x, i, j)) continue;
    Fl3 fl;
    for (int q = 0; q < 7; q++) {
      fl.x[q] = R.fl[0][q];
      fl.y[q] = R.fl[1][q];
      fl.z[q] = R.fl[2][q];
    }
    const EnsState &S = st[e_.x];
    update3d_zm_body<ODR, kEnsZChunk>(sin ? R.u1 : R.u0, R.u0, sout ? R.u1 : R.u0, fl, R.e[0], R.e[1], R.e[2], R.g,
                                      S.dtdx, S.dtdy, S.dtdz, ra, rb, i, j, kc);
  }
}

__global__ __launch_bounds__(256) void k_ens_bc3d(const EnsRec3D *__restrict__ rec, const int2 *__restrict__ tiles,
                                                  const int *__restrict__ count, int sout)
{
  ENS_LOOP(T3_BC)
    const unsigned nbf = (unsigned)((bc3d_points(R.g) + 255) / 256);
    const int fld = l_ / nbf;
    bc3d_body(sout ? R.u1 : R.u0, R.g, R.bc, fld, (long)(l_ % nbf) * 256 + threadIdx.x);
  }
}
#undef CELL_COORDS
#undef ENS_LOOP

// ------------------------------------------------------------------- driver
static unsigned pers_blocks3(const Ensemble *e)
{
  const int m = (e->tune >> 4) & 15;
  return (unsigned)((m ? m : 2) * ens_cus(e));
}

template <int ODR, int FLUX, int MIX>
static void sweeps3d_launch(Ensemble *e, unsigned pb, const EnsRec3D *rec, const EnsPlan &p, int sin)
{
  const int *cnt = p.count;
  ENS_LAUNCH(e, E_XSWEEP, (k_ens_xsweep3d<ODR, FLUX, MIX>), dim3(pb), dim3(256), rec, (const int2 *)p.tiles[T3_XSW],
             cnt, sin);
  ENS_LAUNCH(e, E_YSWEEP, (k_ens_msweep3d<ODR, FLUX, 1, MIX>), dim3(pb * 4), dim3(64), rec,
             (const int2 *)p.tiles[T3_YSW], cnt, sin);
  ENS_LAUNCH(e, E_YSWEEP, (k_ens_msweep3d<ODR, FLUX, 2, MIX>), dim3(pb * 4), dim3(64), rec,
             (const int2 *)p.tiles[T3_ZSW], cnt, sin);
}

template <int ODR, int MIX>
static void sweeps3d(Ensemble *e, int flux, unsigned pb, const EnsRec3D *rec, const EnsPlan &p, int sin)
{
  switch (flux) {
  case FLUX_MLAU: sweeps3d_launch<ODR, FLUX_MLAU, MIX>(e, pb, rec, p, sin); break;
  case FLUX_ROE: sweeps3d_launch<ODR, FLUX_ROE, MIX>(e, pb, rec, p, sin); break;
  case FLUX_HLLD: sweeps3d_launch<ODR, FLUX_HLLD, MIX>(e, pb, rec, p, sin); break;
  default: sweeps3d_launch<ODR, FLUX_LHLLD, MIX>(e, pb, rec, p, sin); break;
  }
}

template <int ODR, int CTW, int MIX>
static int step3d_group(Ensemble *e, EnsGroup *G)
{
  const EnsPlan &p = G->p;
  const EnsRec3D *rec = (const EnsRec3D *)G->rec;
  const unsigned pb = pers_blocks3(e);
  const dim3 b256(256), slots(kEnsSlots / 256);
  const int *cnt = p.count;
  ENS_LAUNCH(e, E_PLAN, k_ens_plan, dim3(1), dim3(kPlanThreads), p, (int)T3_N);
  ENS_LAUNCH(e, E_FILL, k_ens_fill, dim3(kEnsSlots), b256, p, (int)T3_N);
  ENS_LAUNCH(e, E_CFL, k_ens_cfl3d, dim3(pb), b256, rec, p.st, (const int2 *)p.tiles[T3_CFL], cnt);
  ENS_LAUNCH(e, E_DT, k_ens_dt, slots, b256, p);
  const int R_K = G->rk;
  const double rkf[3][2] = {{0.0, 1.0}, {0.5 + (R_K - 2) * 0.25, 0.5 - (R_K - 2) * 0.25}, {1. / 3., 2. / 3.}};
  for (int rk = 0; rk < R_K; rk++) {
    const int sin = (rk == 0) ? 0 : 1;
    const int sout = (rk == R_K - 1) ? 0 : 1;
    const int2 *tc = (const int2 *)p.tiles[T3_CELL];
    ENS_LAUNCH(e, E_PRIM, (k_ens_prim3d<ODR>), dim3(pb), b256, rec, tc, cnt, sin);
    ENS_LAUNCH(e, E_AUX, (k_ens_aux3d<CTW>), dim3(pb), b256, rec, (const int2 *)p.tiles[T3_EDGE], cnt, sin);
    sweeps3d<ODR, MIX>(e, G->flux, pb, rec, p, sin);
    ENS_LAUNCH(e, E_EZ, (k_ens_edge3d<ODR>), dim3(pb), b256, rec, (const int2 *)p.tiles[T3_EDGE], cnt);
    ENS_LAUNCH(e, E_UPD, (k_ens_update3d<ODR>), dim3(pb), b256, rec, (const EnsState *)p.st,
               (const int2 *)p.tiles[T3_UPD], cnt, sin, sout, rkf[rk][0], rkf[rk][1]);
    ENS_LAUNCH(e, E_BC, k_ens_bc3d, dim3(pb), b256, rec, (const int2 *)p.tiles[T3_BC], cnt, sout);
  }
  ENS_LAUNCH(e, E_ADVANCE, k_ens_advance, slots, b256, p);
  return (int)hipGetLastError();
}

typedef int (*EnsStepFn)(Ensemble *, EnsGroup *);

template <int ODR>
static EnsStepFn pick3_odr(int ctw, int precision)
{
  switch (precision) {
  case MLAU_MIXED: return ctw ? step3d_group<ODR, 1, MLAU_MIXED> : step3d_group<ODR, 0, MLAU_MIXED>;
  case MLAU_MIXED_RECON: return ctw ? step3d_group<ODR, 1, MLAU_MIXED_RECON> : step3d_group<ODR, 0, MLAU_MIXED_RECON>;
  default: return ctw ? step3d_group<ODR, 1, MLAU_FP64> : step3d_group<ODR, 0, MLAU_FP64>;
  }
}

static int launch3d(Ensemble *e, EnsGroup *G)
{
  EnsStepFn fn;
  switch (G->odr) {
  case 1: fn = pick3_odr<1>(G->ctw, G->precision); break;
  case 2: fn = pick3_odr<2>(G->ctw, G->precision); break;
  case 3: fn = pick3_odr<3>(G->ctw, G->precision); break;
  default: fn = pick3_odr<4>(G->ctw, G->precision); break;
  }
  return fn(e, G);
}

static int variant3d(const mlau_member &, int) { return 0; }
static const EnsDim kDim3d = {sizeof(EnsRec3D), T3_N, tiles3d, arrays3d, record3d, launch3d, state3d, variant3d};
const EnsDim *ens_dim3d() { return &kDim3d; }

}  // namespace mlau
