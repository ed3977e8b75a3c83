/*
 * mlau.h -- C ABI of libmlau_hip, the MI355X (gfx950) implementation of
 * MLAU's per-step ideal-MHD update.
 *
 * Two ways in:
 *  (1) a resident, handle-based API: state lives in HBM across steps and the
 *      step loop (including the CFL time step) never returns to the host;
 *  (2) one-step drop-ins with the reference solvers' argument lists
 *      (mlau_fd4c_1d / mlau_fd_ct_2d / mlau_fd_ct_3d) that copy host arrays in,
 *      run one step and copy them back.  The C++-linkage shims with the
 *      reference's exact names live in mlau_dropin.h.
 *
 * Layout everywhere is the reference's: one float64 array per variable, ghost
 * cells included, x fastest (ss = nx*j+i; nx*(ny*k+j)+i).  Field order is the
 * reference argument order: 1D ro,mx,my,mz,en,by,bz; 2D/3D
 * ro,mx,my,mz,en,bx,by,bz.  bx[i] lives on face i-1/2 (likewise by, bz in 3D).
 *
 * Every call returns MLAU_OK (0) or an error code; mlau_last_error() gives the
 * message of the last failure on the calling thread.  Calls on one handle are
 * serialised on that handle's HIP stream (2D/3D handles run the y and z
 * sweeps of a stage on internal side streams that rejoin it within the
 * stage); use a handle from one host thread at a time, distinct handles from
 * distinct threads at once.
 */
#ifndef MLAU_H
#define MLAU_H

#ifdef __cplusplus
extern "C" {
#endif

enum { MLAU_FLUX_MLAU = 0, MLAU_FLUX_ROE = 1, MLAU_FLUX_HLLD = 2, MLAU_FLUX_LHLLD = 3 };
enum { MLAU_BC1D_OPEN = 0, MLAU_BC1D_PERIODIC = 1 };
enum { MLAU_FP64 = 0, MLAU_MIXED = 1, MLAU_MIXED_RECON = 2 };
enum { MLAU_OK = 0, MLAU_ERR_ARG = 1, MLAU_ERR_HIP = 2, MLAU_ERR_NOMEM = 3, MLAU_ERR_STATE = 4 };

typedef struct mlau_grid {
  int dim;              /* 1, 2 or 3 */
  int nx, ny, nz;       /* cells per axis INCLUDING ghosts (the reference's nx) */
  int xoff, yoff, zoff; /* ghost cells per side (the reference uses 4) */
  double dx, dy, dz;
  double gamma;         /* specific heat ratio */
  double bx;            /* 1D only: constant normal field */
} mlau_grid;

typedef struct mlau_scheme {
  int odr;              /* spatial order 1..4 (the solvers' ODR) */
  int rk;               /* Runge-Kutta stages 1..3 (R_K) */
  int flux;             /* MLAU_FLUX_* (the solvers' func_flux) */
  int ctw;              /* CT upwind weighting 0/1 (CTW), 2D/3D */
  int bc1d;             /* MLAU_BC1D_* : openbc1d_mhd or prdc_1d */
  int dnxs[8], dnys[8], dnzs[8]; /* per-field BC flags in field order: 0 periodic, +1 Neumann, -1 Dirichlet */
  int precision;        /* MLAU_FP64 (reference arithmetic); MLAU_MIXED (FP32 reconstruction and
                           Riemann solver with FP64 pressure terms, accumulation and CT);
                           MLAU_MIXED_RECON (FP32 reconstruction, FP64 Riemann solver: for long
                           runs, where FP32 flux rounding accumulates past 1e-6).  The mixed
                           forms have no reference counterpart. */
} mlau_scheme;

typedef struct mlau_handle mlau_handle;

/* Resident API.  Replaces, for a whole run, the reference drivers' loop of
 * setbc + mhd_fd_ct_2d + cflcheck (2D/SERIAL/OTvortex/main.cpp:26-43), the 1D
 * loop (1D/SERIAL/shock/main.cpp:162-187) and the 3D loop
 * (3D/MPI/OTvortex/main.cpp:45-87) on one GPU. */
int mlau_create(const mlau_grid *grid, const mlau_scheme *scheme, mlau_handle **out);
void mlau_destroy(mlau_handle *h);
/* Copy nfields host arrays (7 in 1D, 8 in 2D/3D) in / out. */
int mlau_upload(mlau_handle *h, const double *const *fields);
int mlau_download(mlau_handle *h, double *const *fields);
/* Set the time step used by the following steps. */
int mlau_set_dt(mlau_handle *h, double dt);
/* Advance nsteps full RK steps with the current dt (the reference solver
 * called nsteps times, mhd_fd_ct_2d.c:91-660, mhd_fd4c_1d.c:60-271,
 * mhd_fd_ct_3d.c:99-1198). */
int mlau_step(mlau_handle *h, int nsteps);
/* dt = cfl*min(dx,dy[,dz])/vmax with vmax the reference cflcheck speed
 * starting from vfloor (2D/SERIAL/OTvortex/cflcheck.hpp:1-32,
 * 3D/MPI/OTvortex/cflcheck.hpp:1-39); 1D: the driver's one-off estimate
 * (1D/SERIAL/shock/main.cpp:114-128).  dt is kept when vmax is not finite.
 * Outputs may be NULL (then no host synchronisation happens). */
int mlau_cfl(mlau_handle *h, double cfl, double vfloor, double *dt_out, double *vmax_out);
/* nsteps of: step with current dt, then recompute dt with mlau_cfl rules. */
int mlau_step_cfl(mlau_handle *h, int nsteps, double cfl, double vfloor);
int mlau_get_time(mlau_handle *h, double *tim, double *dt, long long *nstep);
/* Restore the time and step counter (resuming from a backup, bkup.hpp:4-44). */
int mlau_set_time(mlau_handle *h, double tim, long long nstep);
int mlau_sync(mlau_handle *h);
/* Device pointers of the resident state (field order), for zero-copy callers. */
int mlau_device_fields(mlau_handle *h, double **fields);
/* Per-kernel HIP-event timing on the handle's stream (for the roofline). */
int mlau_kernel_timing(mlau_handle *h, int enable);
int mlau_kernel_stats(mlau_handle *h, int kernel, double *total_ms, long long *launches, const char **name);
int mlau_num_kernels(void);
const char *mlau_last_error(void);

/* One-step drop-ins with the reference argument lists plus the scheme
 * (which the reference fixes with #defines and driver globals).  Host arrays
 * are updated in place.  A device workspace is cached per thread. */
int mlau_fd4c_1d(double *ro, double *mx, double *my, double *mz, double *en, double *by, double *bz,
                 double bx, double dt, double dx, int nx, int xoff, double gamma, const mlau_scheme *scheme);
/* replaces void mhd_fd4c_1d(...) of 1D/SERIAL/shock/mhd_fd4c_1d.c:60-63 */
int mlau_fd_ct_2d(double *ro, double *mx, double *my, double *mz, double *en, double *bx, double *by,
                  double *bz, double dt, double dx, double dy, int nx, int ny, int xoff, int yoff, double gamma,
                  const mlau_scheme *scheme);
/* replaces void mhd_fd_ct_2d(...) of 2D/SERIAL/OTvortex/mhd_fd_ct_2d.c:91-94 */
int mlau_fd_ct_3d(double *ro, double *mx, double *my, double *mz, double *en, double *bx, double *by,
                  double *bz, double dt, double dx, double dy, double dz, int nx, int ny, int nz, int xoff,
                  int yoff, int zoff, double gamma, const mlau_scheme *scheme);
/* replaces void mhd_fd_ct_3d(...) of 3D/MPI/OTvortex/mhd_fd_ct_3d.c:99-103 (single rank) */

/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  One step of mlau_fd_ct_2d on caller-owned DEVICE arrays,
 * updated in place, queued on `stream` (a hipStream_t; NULL = the default
 * stream) without waiting for it.  It allocates nothing: the stage state and
 * every per-stage array live in one fixed device array of the library
 * (128 MiB per GPU, held from load to process end), so grids are limited to
 * mlau_noalloc_max_cells() cells including ghosts (larger ones are refused
 * with MLAU_ERR_ARG before any GPU work).  All calls on one GPU share that
 * array: calls that overlap on one GPU (other threads or unordered streams)
 * race; calls on one stream are ordered. */
int mlau_fd_ct_2d_noalloc(double *ro, double *mx, double *my, double *mz, double *en, double *bx, double *by,
                          double *bz, double dt, double dx, double dy, int nx, int ny, int xoff, int yoff,
                          double gamma, const mlau_scheme *scheme, void *stream);
long long mlau_noalloc_max_cells(void);

#ifdef __cplusplus
}
#endif
#endif
