/*
 * mlau_oracle.h -- CPU restatement of MLAU's per-step ideal-MHD update.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker, not the product: only the
 * test suite, the smoke check and the CPU-baseline leg of bench.py load it.
 * The product path (libmlau_hip) never links, loads or calls anything here.
 *
 * Every routine follows the reference's arithmetic operation by operation
 * (same evaluation order, same constants, no FMA contraction) so that it is
 * bit-identical to the reference compiled with g++ -O2.  Each function cites
 * the reference file:line it restates.  All arrays use the reference layout:
 * SoA, ghost cells included, x fastest (ss = nx*j + i, nx*(ny*k+j) + i).
 */
#ifndef MLAU_ORACLE_H
#define MLAU_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { OC_FLUX_MLAU = 0, OC_FLUX_ROE = 1, OC_FLUX_HLLD = 2, OC_FLUX_LHLLD = 3 };
enum { OC_BC1D_OPEN = 0, OC_BC1D_PERIODIC = 1 };

typedef struct {
  int odr;   /* spatial order 1..4 (ODR macro) */
  int rk;    /* Runge-Kutta stages 1..3 (R_K macro) */
  int flux;  /* OC_FLUX_* (func_flux assignment) */
  int ctw;   /* CT upwind weighting flag (CTW macro), 2D/3D only */
} oc_scheme;

/* ---- point kernels (per-function known-answer checks) ---- */
double oc_minmod(double a, double b);
double oc_minmod3(double a, double b, double c);
void oc_muscl_mm(double f1, double f2, double f3, double *fl, double *fr);
void oc_muscl_mc(double f1, double f2, double f3, double *fl, double *fr);
void oc_muscl_kr(double f1, double f2, double f3, double *fl, double *fr);
void oc_wcns3(double f0, double f1, double f2, double *fl, double *fr);
void oc_wcns4(double f0, double f1, double f2, double f3, double f4, double *fl, double *fr);
/* p = (ro,vx,vy,vz,bx,by,bz,pr); al[7]; L[49] rows; R[49] columns (R[7*k+n] = n-th comp of k-th vector) */
void oc_eigen_st(const double *p, double gamma, double *al, double *L, double *R);
void oc_eigen_ae(const double *p, double gamma, double *L, double *R);
/* st[7*5]: stencil rows (ro,vn,vt,vu,bt,bu,pr) x 5 points; kind 0 = characteristic (st), 1 = alfven (ae) */
void oc_reconst(int kind, int odr, const double *st, double bn, double gamma, double *ul, double *ur);
/* in[14]: left (ro,vn,vt,vu,bt,bu,pr) then right; out[7]: (ro,mn,mt,mu,bt,bu,en) */
void oc_flux(int kind, const double *in, double bn, double gamma, const double *dv, double *out);

/* ---- full per-step updates (in place, one RK step incl. per-stage BCs) ---- */
void oc_step_1d(double *ro, double *mx, double *my, double *mz, double *en, double *by, double *bz,
                double bx, double dt, double dx, int nx, int xoff, double gamma,
                const oc_scheme *sc, int bc1d);
void oc_step_2d(double *ro, double *mx, double *my, double *mz, double *en, double *bx, double *by, double *bz,
                double dt, double dx, double dy, int nx, int ny, int xoff, int yoff, double gamma,
                const oc_scheme *sc, const int *dnxs, const int *dnys, const int *stxs, const int *stys);
void oc_step_3d(double *ro, double *mx, double *my, double *mz, double *en, double *bx, double *by, double *bz,
                double dt, double dx, double dy, double dz, int nx, int ny, int nz, int xoff, int yoff, int zoff,
                double gamma, const oc_scheme *sc, const int *dnxs, const int *dnys, const int *dnzs,
                const int *stxs, const int *stys, const int *stzs);

/* ghost fills */
void oc_bc2d(double *f, int nx, int ny, int xoff, int yoff, int stx, int dnx, int sty, int dny);
void oc_bc3d_state(double **f, int nn, int nx, int ny, int nz, int xoff, int yoff, int zoff,
                   const int *dnxs, const int *dnys, const int *dnzs,
                   const int *stxs, const int *stys, const int *stzs);

/* CFL speeds: return vmax over the interior, starting from vfloor */
double oc_cfl_vmax_2d(const double *ro, const double *mx, const double *my, const double *mz, const double *en,
                      const double *bx, const double *by, const double *bz,
                      int nx, int ny, int xoff, int yoff, double gamma, double vfloor);
double oc_cfl_vmax_3d(const double *ro, const double *mx, const double *my, const double *mz, const double *en,
                      const double *bx, const double *by, const double *bz,
                      int nx, int ny, int nz, int xoff, int yoff, int zoff, double gamma, double vfloor);
double oc_vmax_1d(const double *ro, const double *mx, const double *my, const double *mz, const double *en,
                  const double *by, const double *bz, double bx, int nx, double gamma, double vfloor);

int oc_num_threads(void);
void oc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
