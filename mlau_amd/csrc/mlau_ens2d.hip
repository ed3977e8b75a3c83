// mlau_ens2d.hip -- batched 2D members of the ensemble engine.
//
// One step of a 2D group is: plan (compact running members into work lists),
// fill (rewrite the lists when the running set changed), CFL speed per member,
// dt per member (clipped at its end time), R_K stages of
//   prim+aux (one pass), x-sweep, y-sweep, ez, update, bc
// (or, for periodic members at odr 2 with rk >= 2, the fused stage kernel,
// which also fills the ghost cells) and advance (time, step count, status).  The stage kernels are persistent:
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
  U2 u2;  // third state of the fused stages (aliases w.fx[0..6], w.fcx, which they do not use)
  W2 w;
  G2 g;
  BcFlags bc;
  int flux, pad;
};

enum { T2_CFL = 0, T2_CELL, T2_XSW, T2_YSW, T2_EZ, T2_UPD, T2_BC, T2_FUSE, T2_N };

// Fused stage tiles: kFT x kFT cells of the interior.
constexpr int kFT = 28;

// Kernel variant of a member: 1 = fused stage kernel.  It needs periodic
// boundaries (the ghost cells then hold exact copies of interior cells, which
// the kernel writes itself, the face B at nx-xoff and ny-yoff included), odr 2
// (the stencil reach the tile halo is sized for), rk >= 2 (a stage never
// updates its own input in place), ghost widths of at least 4 and an interior
// of at least two ghost widths per axis (each interior cell is then the source
// of at most one ghost column and one ghost row).
static int variant2d(const mlau_member &m, int flags)
{
  const mlau_scheme &s = m.scheme;
  const mlau_grid &g = m.grid;
  if ((flags & MLAU_ENS_UNFUSED) || s.odr != 2 || s.rk < 2 || g.xoff < 4 || g.yoff < 4 ||
      g.nx - 2 * g.xoff < 2 * g.xoff || g.ny - 2 * g.yoff < 2 * g.yoff)
    return 0;
  for (int n = 0; n < 8; n++)
    if (s.dnxs[n] || s.dnys[n]) return 0;
  return 1;
}

static int ych_of(int tune)
{
  switch (tune & 3) {
  case 1: return 16;
  case 2: return 4;
  default: return 8;
  }
}

static long long tiles2d(const mlau_grid &g, int kind, int tune, int variant)
{
  const long long nx = g.nx, ny = g.ny, xo = g.xoff, yo = g.yoff;
  if (variant && kind != T2_CFL && kind != T2_FUSE) return 0;
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
  case T2_FUSE: return variant ? ((nx - 2 * xo + kFT - 1) / kFT) * ((ny - 2 * yo + kFT - 1) / kFT) : 0;
  default: return 0;
  }
}

static long long tiles2d_member(const mlau_member &m, int kind, int tune, int variant)
{
  return tiles2d(m.grid, kind, tune, variant);
}

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
  for (int n = 0; n < 7; n++) r.u2.f[n] = r.w.fx[n];
  r.u2.f[7] = r.w.fcx;
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
// sin/sout select the stage input/output state: 0 = u0, 1 = u1, 2 = u2.
__device__ __forceinline__ U2 ubuf2(const EnsRec2D &R, int s) { return s == 0 ? R.u0 : (s == 1 ? R.u1 : R.u2); }
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
    const double v = (t < total) ? cfl2d_cell_checked(R.u0, R.g, t) : 0.0;
    ens_vmax_block(v, &st[e_.x].vmax_bits, red);
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
    if (st[e_.x].halt) continue;  // member failed at this step's CFL check
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
    bc2d_body(ubuf2(R, sout), R.g, R.bc, fld, (long)(l_ % nbf) * 256 + threadIdx.x);
  }
}

// ------------------------------------------------------------ fused stage
// One RK stage of a periodic member at odr 2 in one kernel: a tile of up to
// kFT x kFT cells (a..a+tw-1, c..c+th-1) is updated from values formed in
// LDS for the tile and the halo its stencils reach:
//   primitives    cells a-3..a+tw+2 x c-3..c+th+2
//   x faces       a..a+tw on rows c-2..c+th+1 (cells a-1..a+tw reconstructed)
//   y faces       c..c+th on columns a-2..a+tw+1 (cells c-1..c+th)
//   corner E      a..a+tw x c..c+th
// Each value comes from the expressions of the separate stage kernels
// (prim2d/primaux2d, xsweep2d, ysweep2d, ez2d and update2d bodies) on the same
// inputs, so the bits are theirs.  Ghost cells of a periodic grid hold exact
// copies of interior cells, so the centred B and velocity differences the
// separate kernels take through the boundary map equal the ones formed here
// from the ghost cells.  Cells within two ghost widths of an edge also write
// their periodic images into the ghost frame (the face B at nx-xoff and ny-yoff
// is the image of the face at xoff and yoff), as the boundary fill would: no
// tile reads the output state within the stage, and each ghost cell has one
// source cell.
// Per stage the member's state moves through HBM once (input with its halo
// from L2, step-start state, output): no primitive, flux or E-field arrays.
// 512 threads (8 wavefronts, one block per CU: the LDS below is 151 KB of the
// CU's 160 KB; two waves per SIMD).  The sweeps hold one row (x) or column (y)
// of up to kFT + 2 cells per 32-lane half wavefront; with kFT = 28 the kFT + 4
// rows (columns) make 16 halves, two passes of the 8 wavefronts, so no lane
// group idles at a phase's end.  Row strides are odd numbers of doubles, so
// the y-sweep lanes (one per row) spread over the LDS banks.
constexpr int kFThreads = 512;
constexpr int kFWaves = kFThreads / 64;
constexpr int kFGL = 32;                                           // lanes per row/column group
constexpr int kFGW = 64 / kFGL;                                    // groups per wavefront
constexpr int kFCells = (kFT * kFT + kFThreads - 1) / kFThreads;  // tile cells per thread
constexpr int kFPS = kFT + 7;  // primitive row stride
constexpr int kFPW = kFT + 6;  // primitive region edge
constexpr int kFXS = kFT + 1;  // x faces per row, corners per row, flux row stride
constexpr int kFXR = kFT + 4;  // rows of x faces
constexpr int kFYS = kFT + 5;  // y-face row stride (kFT + 4 columns)
static_assert(kFT + 2 <= kFGL && kFXS % 2 == 1 && kFPS % 2 == 1 && kFYS % 2 == 1, "fused tile shape");
enum { FP_RO = 0, FP_VX, FP_VY, FP_VZ, FP_BZ, FP_PR, FP_CX, FP_CY, FP_N };

struct FusedLds {
  double p[FP_N][kFPW * kFPS];   // cell (a-3+li, c-3+lj) at lj*kFPS + li
  double f6[6][kFXS * kFXS];     // ro,mx,my,mz,bz,en flux at face (a+f, c+r): r*kFXS + f
                                 // (x pass: r < th, f <= tw; y pass: r <= th, f < tw)
  double x2[2][kFXR * kFXS];     // By upwind part, central part at x face (a+f, c-2+r): r*kFXS + f
  double y2[2][kFXS * kFYS];     // Bx upwind part, central part at y face (a-2+f, c+r): r*kFYS + f
  double ez[kFXS * kFXS];        // corner (a+f, c+r): r*kFXS + f
  // m2 >= 1 per cell.  The sweeps' test max(m2[-1], m2[0], m2[+1]) >= 1 holds
  // exactly when one of the three does (m2 is NaN only in a failing member).
  unsigned char sup[kFPW * kFPS];
};

__device__ __forceinline__ double fz_dvy(const FusedLds &L, int q)
{
  const double *vy = L.p[FP_VY] + q;
  return dmin((vy[kFPS] - vy[0]), (vy[0] - vy[-kFPS]));
}
__device__ __forceinline__ double fz_dvx(const FusedLds &L, int q)
{
  const double *vx = L.p[FP_VX] + q;
  return dmin((vx[1] - vx[0]), (vx[0] - vx[-1]));
}

template <int FLUX, int MIX>
__global__ __launch_bounds__(kFThreads) void k_ens_stage2d(const EnsRec2D *__restrict__ rec,
                                                                    const EnsState *st,
                                                                    const int2 *__restrict__ tiles,
                                                                    const int *__restrict__ count, int sin, int sout,
                                                                    double ra, double rb, int ctw)
{
  constexpr int ODR = 2;
  __shared__ FusedLds L;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const bool first = ra == 0.0;
  ENS_LOOP(T2_FUSE)
    if (st[e_.x].halt) continue;  // member failed at this step's CFL check
    const U2 in = ubuf2(R, sin), out = ubuf2(R, sout), ut = R.u0;
    const int nx = R.g.nx, ny = R.g.ny, xo = R.g.xoff, yo = R.g.yoff;
    const double gamma = R.g.gamma;
    const int ntx = (nx - 2 * xo + kFT - 1) / kFT;
    const int a = xo + kFT * (int)(l_ % ntx), c = yo + kFT * (int)(l_ / ntx);
    const int tw = min(kFT, nx - xo - a), th = min(kFT, ny - yo - c);
    const double dtdx = st[e_.x].dtdx, dtdy = st[e_.x].dtdy;
    // primitives (prim2d_body) on the fixed region edge (constant divisor)
    {
      const int pw = tw + 6, ph = th + 6;
      for (int t = tid; t < kFPW * kFPW; t += kFThreads) {
        const int li = t % kFPW, lj = t / kFPW;
        if (li >= pw || lj >= ph) continue;
        const long s = (long)(c - 3 + lj) * nx + (a - 3 + li);
        const double cx = bc_centre<ODR>(0.0, in.f[5][s], in.f[5][s + 1], 0.0);
        const double cy = bc_centre<ODR>(0.0, in.f[6][s], in.f[6][s + nx], 0.0);
        const double ro = in.f[0][s];
        const double iro = 1.0 / ro;
        const double vx = in.f[1][s] * iro;
        const double vy = in.f[2][s] * iro;
        const double vz = in.f[3][s] * iro;
        const double bz = in.f[7][s];
        const double v2 = (vx * vx + vy * vy + vz * vz);
        const double b2 = (cx * cx + cy * cy + bz * bz);
        const double pr = (gamma - 1.0) * (in.f[4][s] - 0.5 * (ro * v2 + b2));
        const int q = lj * kFPS + li;
        L.p[FP_RO][q] = ro;
        L.p[FP_VX][q] = vx;
        L.p[FP_VY][q] = vy;
        L.p[FP_VZ][q] = vz;
        L.p[FP_BZ][q] = bz;
        L.p[FP_PR][q] = pr;
        L.sup[q] = v2 / (gamma * pr * iro) >= 1.0;
        L.p[FP_CX][q] = cx;
        L.p[FP_CY][q] = cy;
      }
    }
    __syncthreads();
    // x sweep (xsweep2d_body): lanes kFGL*k+m hold cell a-1+m of row kFGW*g+k;
    // the left state of face i comes from the lane of cell i-1
    {
      const int rows = th + 4, cells = tw + 2, groups = (rows + kFGW - 1) / kFGW;
      for (int gi = wave; gi < groups; gi += kFWaves) {
        const int rr = lane / kFGL;
        int cc = lane - kFGL * rr, row = kFGW * gi + rr;
        const bool live = row < rows && cc < cells;
        row = min(row, rows - 1);
        cc = min(cc, cells - 1);
        const int q = (row + 1) * kFPS + (cc + 2);
        Stencil S;
#pragma unroll
        for (int m = 0; m < 5; m++) {
          S.q[0][m] = L.p[FP_RO][q - 2 + m];
          S.q[1][m] = L.p[FP_VX][q - 2 + m];
          S.q[2][m] = L.p[FP_VY][q - 2 + m];
          S.q[3][m] = L.p[FP_VZ][q - 2 + m];
          S.q[4][m] = L.p[FP_CY][q - 2 + m];
          S.q[5][m] = L.p[FP_BZ][q - 2 + m];
          S.q[6][m] = L.p[FP_PR][q - 2 + m];
        }
        const bool sup = L.sup[q - 1] | L.sup[q] | L.sup[q + 1];
        double vl[7], vr[7];
        recon_p<ODR, MIX>(S, sup, L.p[FP_CX][q], gamma, vl, vr);
        double pp[5];
#pragma unroll
        for (int m = 0; m < 5; m++) pp[m] = L.p[FP_CY][q - 2 + m] * S.q[1][m];
        const double a0 = lcal_l<ODR>(pp[0], pp[1], pp[2], pp[3], pp[4]);
        const double b0 = lcal_r<ODR>(pp[0], pp[1], pp[2], pp[3], pp[4]);
        const double a1 = lcal_l<ODR>(S.q[2][0], S.q[2][1], S.q[2][2], S.q[2][3], S.q[2][4]);
        const double b1 = lcal_r<ODR>(S.q[2][0], S.q[2][1], S.q[2][2], S.q[2][3], S.q[2][4]);
        const long s = (long)(c - 2 + row) * nx + (a - 1 + cc);
        const double bxr = in.f[5][s];
        const double ql = a0 - in.f[5][s + 1] * a1;  // face i+1
        const double qr = b0 - bxr * b1;             // face i
        double Lf[7];
#pragma unroll
        for (int k = 0; k < 7; k++) Lf[k] = shfl_up1(vl[k]);
        const double qlL = shfl_up1(ql);
        if (live && cc >= 1) {
          const double bn = 0.5 * (bxr + bxr);
          const double dv0 = L.p[FP_VX][q] - L.p[FP_VX][q - 1];
          const double dv1 = dmin(fz_dvy(L, q - 1), fz_dvy(L, q));
          double f[7];
          riemann_p<FLUX, MIX>(Lf, vr, bn, gamma, dv0, dv1, f);
          const double fc = 0.5 * (qlL + qr);
          const int fi = cc - 1;
          L.x2[0][row * kFXS + fi] = f[4] - fc;
          L.x2[1][row * kFXS + fi] = fc;
          const int jr = row - 2;
          if (jr >= 0 && jr < th) {
            double *d = &L.f6[0][jr * kFXS + fi];
            d[0 * kFXS * kFXS] = f[0];
            d[1 * kFXS * kFXS] = f[1];
            d[2 * kFXS * kFXS] = f[2];
            d[3 * kFXS * kFXS] = f[3];
            d[4 * kFXS * kFXS] = f[5];
            d[5 * kFXS * kFXS] = f[6];
          }
        }
      }
    }
    __syncthreads();
    // x part of the cell update (update2d_body: in - dtdx*d0, then - dtdy*d1);
    // thread tid holds tile cells tid + kFThreads*u
    const int fld[6] = {0, 1, 2, 3, 7, 4};
    double part[kFCells][6];
#pragma unroll
    for (int u = 0; u < kFCells; u++) {
      const int k = tid + kFThreads * u, ci = k % kFT, cj = k / kFT;
      const bool own = k < kFT * kFT && ci < tw && cj < th;
      const long so = (long)(c + cj) * nx + (a + ci);
#pragma unroll
      for (int t = 0; t < 6; t++) {
        const double *fx = &L.f6[t][(own ? cj * kFXS + ci : 0)];
        part[u][t] = own ? in.f[fld[t]][so] - dtdx * diff4<ODR>(0.0, fx[0], fx[1], 0.0) : 0.0;
      }
    }
    __syncthreads();
    // y sweep (ysweep2d_body): lanes kFGL*k+m hold cell c-1+m of column kFGW*g+k
    {
      const int cols = tw + 4, cells = th + 2, groups = (cols + kFGW - 1) / kFGW;
      for (int gi = wave; gi < groups; gi += kFWaves) {
        const int cg = lane / kFGL;
        int rr = lane - kFGL * cg, col = kFGW * gi + cg;
        const bool live = col < cols && rr < cells;
        col = min(col, cols - 1);
        rr = min(rr, cells - 1);
        const int q = (rr + 2) * kFPS + (col + 1);
        Stencil S;
#pragma unroll
        for (int m = 0; m < 5; m++) {
          const int qm = q + (m - 2) * kFPS;
          S.q[0][m] = L.p[FP_RO][qm];
          S.q[1][m] = L.p[FP_VY][qm];
          S.q[2][m] = L.p[FP_VZ][qm];
          S.q[3][m] = L.p[FP_VX][qm];
          S.q[4][m] = L.p[FP_BZ][qm];
          S.q[5][m] = L.p[FP_CX][qm];
          S.q[6][m] = L.p[FP_PR][qm];
        }
        const bool sup = L.sup[q - kFPS] | L.sup[q] | L.sup[q + kFPS];
        double vl[7], vr[7];
        recon_p<ODR, MIX>(S, sup, L.p[FP_CY][q], gamma, vl, vr);
        double pp[5];
#pragma unroll
        for (int m = 0; m < 5; m++) pp[m] = S.q[5][m] * S.q[1][m];
        const double a0 = lcal_l<ODR>(pp[0], pp[1], pp[2], pp[3], pp[4]);
        const double b0 = lcal_r<ODR>(pp[0], pp[1], pp[2], pp[3], pp[4]);
        const double a1 = lcal_l<ODR>(S.q[3][0], S.q[3][1], S.q[3][2], S.q[3][3], S.q[3][4]);
        const double b1 = lcal_r<ODR>(S.q[3][0], S.q[3][1], S.q[3][2], S.q[3][3], S.q[3][4]);
        const long s = (long)(c - 1 + rr) * nx + (a - 2 + col);
        const double byf = in.f[6][s];
        const double ql = a0 - in.f[6][s + nx] * a1;  // face r+1
        const double qr = b0 - byf * b1;              // face r
        double Lf[7];
#pragma unroll
        for (int k = 0; k < 7; k++) Lf[k] = shfl_up1(vl[k]);
        const double qlp = shfl_up1(ql);
        if (live && rr >= 1) {
          const double bn = 0.5 * (byf + byf);
          const double dv0 = L.p[FP_VY][q] - L.p[FP_VY][q - kFPS];
          const double dv1 = dmin(fz_dvx(L, q - kFPS), fz_dvx(L, q));
          double f[7];
          riemann_p<FLUX, MIX>(Lf, vr, bn, gamma, dv0, dv1, f);
          const double fc = 0.5 * (qlp + qr);
          const int fr = rr - 1;
          L.y2[0][fr * kFYS + col] = f[5] - fc;
          L.y2[1][fr * kFYS + col] = fc;
          const int ic = col - 2;
          if (ic >= 0 && ic < tw) {
            double *d = &L.f6[0][fr * kFXS + ic];
            d[0 * kFXS * kFXS] = f[0];
            d[1 * kFXS * kFXS] = f[3];
            d[2 * kFXS * kFXS] = f[1];
            d[3 * kFXS * kFXS] = f[2];
            d[4 * kFXS * kFXS] = f[4];
            d[5 * kFXS * kFXS] = f[6];
          }
        }
      }
    }
    __syncthreads();
    // corner E-field (primaux2d_body's CT weight, ez2d_body)
    for (int t = tid; t < kFXS * kFXS; t += kFThreads) {
      const int ei = t % kFXS, ej = t / kFXS;
      if (ei > tw || ej > th) continue;
      double ctx = 0.5;
      if (ctw) {
        const int q = (ej + 3) * kFPS + (ei + 3), qa = q - kFPS;
        const double *vx = L.p[FP_VX], *vy = L.p[FP_VY], *ro = L.p[FP_RO];
        double vxc = 0.25 * (vx[qa - 1] + vx[qa] + vx[q - 1] + vx[q]);
        double vyc = 0.25 * (vy[qa - 1] + vy[qa] + vy[q - 1] + vy[q]);
        double roc = 0.25 * (ro[qa - 1] + ro[qa] + ro[q - 1] + ro[q]);
        roc = 1.0 / sqrt(roc);
        const long s = (long)(c + ej) * nx + (a + ei);
        const double bxc = 0.5 * (in.f[5][s - nx] + in.f[5][s]);
        const double byc = 0.5 * (in.f[6][s - 1] + in.f[6][s]);
        vxc = fabs(vxc) + fabs(bxc * roc);
        vyc = fabs(vyc) + fabs(byc * roc);
        const double eps = 1e-6;
        ctx = (vxc + 0.5 * eps) / (vxc + vyc + eps);
      }
      const double *u4 = &L.x2[0][ej * kFXS + ei], *uc = &L.x2[1][ej * kFXS + ei];
      double a0 = lcal_l<ODR>(0.0, u4[0], u4[kFXS], u4[2 * kFXS], 0.0);
      double b0 = lcal_r<ODR>(0.0, u4[kFXS], u4[2 * kFXS], u4[3 * kFXS], 0.0);
      double a1 = lcal_l<ODR>(0.0, uc[0], uc[kFXS], uc[2 * kFXS], 0.0);
      double b1 = lcal_r<ODR>(0.0, uc[kFXS], uc[2 * kFXS], uc[3 * kFXS], 0.0);
      double ez = 0.0;
      ez += -0.5 * ((a0 + b0) + (1.0 - ctx) * (a1 + b1));
      const double *v4 = &L.y2[0][ej * kFYS + ei], *vc = &L.y2[1][ej * kFYS + ei];
      a0 = lcal_l<ODR>(0.0, v4[0], v4[1], v4[2], 0.0);
      b0 = lcal_r<ODR>(0.0, v4[1], v4[2], v4[3], 0.0);
      a1 = lcal_l<ODR>(0.0, vc[0], vc[1], vc[2], 0.0);
      b1 = lcal_r<ODR>(0.0, vc[1], vc[2], vc[3], 0.0);
      ez += +0.5 * ((a0 + b0) + ctx * (a1 + b1));
      L.ez[ej * kFXS + ei] = ez;
    }
    __syncthreads();
    // cell update, CT face B, SSP-RK combine (update2d_body), periodic images
    const int NX = nx - 2 * xo, NY = ny - 2 * yo;
#pragma unroll
    for (int u = 0; u < kFCells; u++) {
      const int k = tid + kFThreads * u, ci = k % kFT, cj = k / kFT;
      if (!(k < kFT * kFT && ci < tw && cj < th)) continue;
      const int i = a + ci, j = c + cj;
      const long so = (long)j * nx + i;
      double v[8];
#pragma unroll
      for (int t = 0; t < 6; t++) {
        const double *fy = &L.f6[t][cj * kFXS + ci];
        const double upd = rb * (part[u][t] - dtdy * diff4<ODR>(0.0, fy[0], fy[kFXS], 0.0));
        v[fld[t]] = first ? upd : ra * ut.f[fld[t]][so] + upd;
      }
      const double *ez = &L.ez[cj * kFXS + ci];
      double du = diff4<ODR>(0.0, ez[0], ez[kFXS], 0.0);
      double upd = rb * (in.f[5][so] - dtdy * du);
      v[5] = first ? upd : ra * ut.f[5][so] + upd;
      du = diff4<ODR>(0.0, ez[0], ez[1], 0.0);
      upd = rb * (in.f[6][so] + dtdx * du);
      v[6] = first ? upd : ra * ut.f[6][so] + upd;
      // ghost column / row whose periodic source is (i, j), or -1
      const int gi = (i < 2 * xo) ? i + NX : (i >= nx - 2 * xo ? i - NX : -1);
      const int gj = (j < 2 * yo) ? j + NY : (j >= ny - 2 * yo ? j - NY : -1);
#pragma unroll
      for (int n = 0; n < 8; n++) {
        double *f = out.f[n];
        f[so] = v[n];
        if (gi >= 0) f[(long)j * nx + gi] = v[n];
        if (gj >= 0) f[(long)gj * nx + i] = v[n];
        if (gi >= 0 && gj >= 0) f[(long)gj * nx + gi] = v[n];
      }
    }
  }
}
#undef ENS_LOOP

// ------------------------------------------------------------------- driver
static unsigned pers_blocks(const Ensemble *e)
{
  const int m = (e->tune >> 4) & 15;
  return (unsigned)((m ? m : 2) * ens_cus(e));
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

template <int MIX>
static void stage2d_launch(Ensemble *e, int flux, unsigned pb, const EnsRec2D *rec, const EnsPlan &p, int sin,
                           int sout, double ra, double rb, int ctw)
{
  const int2 *t = (const int2 *)p.tiles[T2_FUSE];
  const EnsState *st = p.st;
  const int *cnt = p.count;
  const dim3 b(kFThreads);
  switch (flux) {
  case FLUX_MLAU:
    ENS_LAUNCH(e, E_STAGE, (k_ens_stage2d<FLUX_MLAU, MIX>), dim3(pb), b, rec, st, t, cnt, sin, sout, ra, rb, ctw);
    break;
  case FLUX_ROE:
    ENS_LAUNCH(e, E_STAGE, (k_ens_stage2d<FLUX_ROE, MIX>), dim3(pb), b, rec, st, t, cnt, sin, sout, ra, rb, ctw);
    break;
  case FLUX_HLLD:
    ENS_LAUNCH(e, E_STAGE, (k_ens_stage2d<FLUX_HLLD, MIX>), dim3(pb), b, rec, st, t, cnt, sin, sout, ra, rb, ctw);
    break;
  default:
    ENS_LAUNCH(e, E_STAGE, (k_ens_stage2d<FLUX_LHLLD, MIX>), dim3(pb), b, rec, st, t, cnt, sin, sout, ra, rb, ctw);
    break;
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
  for (int rk = 0; rk < R_K && G->variant; rk++) {
    // fused stages: u0 -> u1 -> u2 -> u0 (rk 3) or u0 -> u1 -> u0 (rk 2); the
    // step-start state u0 is read only at the cell being written
    const int sin = (rk == 0) ? 0 : (rk == 1 ? 1 : 2);
    const int sout = (rk == R_K - 1) ? 0 : (rk == 0 ? 1 : 2);
    stage2d_launch<MIX>(e, G->flux, ens_cus(e), rec, p, sin, sout, rkf[rk][0], rkf[rk][1], CTW);  // a block per CU
  }
  for (int rk = 0; rk < R_K && !G->variant; rk++) {
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

static const EnsDim kDim2d = {sizeof(EnsRec2D), T2_N, tiles2d_member, arrays2d, record2d, launch2d, state2d,
                              variant2d};
const EnsDim *ens_dim2d() { return &kDim2d; }

}  // namespace mlau
