// mlau_api.hip -- C ABI of libmlau_hip: handle lifetime, host<->device copies,
// step loops, CFL time step, per-kernel timing and the one-step drop-ins.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

#include "mlau_internal.h"

namespace mlau {

const char *kKernelNames[K_COUNT] = {
    "prim2d", "aux2d", "xsweep2d", "ysweep2d", "ez2d", "update2d", "bc2d", "cfl2d",
    "sweep1d", "openbc1d", "update1d", "cfl1d",
    "prim3d", "aux3d", "xsweep3d", "ysweep3d", "zsweep3d", "edge3d", "update3d", "bc3d", "cfl3d",
    "finalize", "primaux2d", "stage1d", "xsweep2d_inl", "ysweep2d_inl"};

static thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
  g_err = msg;
  return code;
}
int hip_fail(hipError_t e, const char *where)
{
  return fail(MLAU_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}
#define HIPCHK(call)                              \
  do {                                            \
    hipError_t _e = (call);                       \
    if (_e != hipSuccess) return hip_fail(_e, #call); \
  } while (0)

__global__ void k_finalize_dt(StepInfo *info, double cfl, double dr, double dx, double dy, double dz)
{
  const double vmax = __longlong_as_double((long long)info->vmax_bits);
  info->vmax = vmax;
  if (isfinite(vmax)) {
    const double dt = cfl * dr / vmax;
    info->dt = dt;
    info->dtdx = dt / dx;
    info->dtdy = dt / dy;
    info->dtdz = dt / dz;
  }
}
__global__ void k_cfl_reset(StepInfo *info, double vfloor)
{
  info->vmax_bits = (unsigned long long)__double_as_longlong(vfloor);
}
__global__ void k_advance_time(StepInfo *info)
{
  info->tim += info->dt;
  info->nstep += 1;
}
__global__ void k_set_time(StepInfo *info, double tim, long long nstep)
{
  info->tim = tim;
  info->nstep = nstep;
}

__global__ void k_set_dt(StepInfo *info, double dt, double dx, double dy, double dz)
{
  info->dt = dt;
  info->dtdx = dt / dx;
  info->dtdy = dt / dy;
  info->dtdz = dt / dz;
}

void timing_begin(Handle *h, int k, hipEvent_t *a)
{
  (void)k;
  Timing &t = h->timing;
  if (t.pool.empty()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    t.pool.push_back(e);
  }
  *a = t.pool.back();
  t.pool.pop_back();
  hipEventRecord(*a, h->stream);
}
void timing_end(Handle *h, int k, hipEvent_t a)
{
  Timing &t = h->timing;
  if (!a) return;
  hipEvent_t b;
  if (t.pool.empty()) {
    if (hipEventCreate(&b) != hipSuccess) return;
  } else {
    b = t.pool.back();
    t.pool.pop_back();
  }
  hipEventRecord(b, h->stream);
  t.pending.push_back({k, a, b});
  if (t.pending.size() > 4096) {  // fold finished records to bound the event count
    hipEventSynchronize(b);
    for (auto &r : t.pending) {
      float ms = 0.f;
      hipEventElapsedTime(&ms, r.a, r.b);
      t.total_ms[r.k] += ms;
      t.count[r.k] += 1;
      t.pool.push_back(r.a);
      t.pool.push_back(r.b);
    }
    t.pending.clear();
  }
}
static void timing_flush(Handle *h)
{
  Timing &t = h->timing;
  if (t.pending.empty()) return;
  hipEventSynchronize(t.pending.back().b);
  for (auto &r : t.pending) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, r.a, r.b);
    t.total_ms[r.k] += ms;
    t.count[r.k] += 1;
    t.pool.push_back(r.a);
    t.pool.push_back(r.b);
  }
  t.pending.clear();
}

int mlau_validate(const mlau_grid *g, const mlau_scheme *s)
{
  if (!g || !s) return fail(MLAU_ERR_ARG, "null grid or scheme");
  if (g->dim < 1 || g->dim > 3) return fail(MLAU_ERR_ARG, "dim must be 1, 2 or 3");
  if (s->odr < 1 || s->odr > 4) return fail(MLAU_ERR_ARG, "odr must be 1..4");
  if (s->rk < 1 || s->rk > 3) return fail(MLAU_ERR_ARG, "rk must be 1..3");
  if (s->flux < 0 || s->flux > 3) return fail(MLAU_ERR_ARG, "flux must be 0..3");
  if (s->precision != MLAU_FP64 && s->precision != MLAU_MIXED && s->precision != MLAU_MIXED_RECON)
    return fail(MLAU_ERR_ARG, "precision must be 0, 1 or 2");
  if (!(g->gamma > 1.0)) return fail(MLAU_ERR_ARG, "gamma must exceed 1");
  if (!(g->dx > 0.0) || (g->dim >= 2 && !(g->dy > 0.0)) || (g->dim == 3 && !(g->dz > 0.0)))
    return fail(MLAU_ERR_ARG, "cell spacing must be positive on every axis of the grid");
  if (g->dim == 1 && s->bc1d != MLAU_BC1D_OPEN && s->bc1d != MLAU_BC1D_PERIODIC)
    return fail(MLAU_ERR_ARG, "bc1d must be 0 (open) or 1 (periodic)");
  if (g->dim >= 2 && s->ctw != 0 && s->ctw != 1) return fail(MLAU_ERR_ARG, "ctw must be 0 or 1");
  if (g->xoff != 4) return fail(MLAU_ERR_ARG, "xoff must be 4 (the reference's fixed ghost width)");
  if (g->nx < 2 * g->xoff + 4) return fail(MLAU_ERR_ARG, "nx too small");
  if (g->dim >= 2) {
    if (g->yoff != 4 || g->ny < 2 * g->yoff + 4) return fail(MLAU_ERR_ARG, "ny/yoff invalid");
    const int axes = g->dim;
    const int *dn[3] = {s->dnxs, s->dnys, s->dnzs};
    for (int a = 0; a < axes; a++) {
      bool any0 = false, anyn = false;
      for (int n = 0; n < 8; n++) {
        if (dn[a][n] == 0) any0 = true;
        else if (dn[a][n] == 1 || dn[a][n] == -1) anyn = true;
        else return fail(MLAU_ERR_ARG, "boundary flags must be 0, +1 or -1");
      }
      if (any0 && anyn) return fail(MLAU_ERR_ARG, "an axis must be periodic for all fields or for none");
    }
  }
  if (g->dim == 3 && (g->zoff != 4 || g->nz < 2 * g->zoff + 4)) return fail(MLAU_ERR_ARG, "nz/zoff invalid");
  return MLAU_OK;
}

}  // namespace mlau

using namespace mlau;

struct mlau_handle : public Handle {};

// Side streams and their events (every launch on them has been joined into
// the main stream, so they are idle once the main stream is).
static void release_side(Handle *h)
{
  for (int k = 0; k < 2; k++) {
    if (h->side[k]) {
      hipStreamSynchronize(h->side[k]);
      hipStreamDestroy(h->side[k]);
    }
    if (h->ev_join[k]) hipEventDestroy(h->ev_join[k]);
    h->side[k] = nullptr;
    h->ev_join[k] = nullptr;
  }
  if (h->ev_fork) hipEventDestroy(h->ev_fork);
  h->ev_fork = nullptr;
}

extern "C" {

const char *mlau_last_error(void) { return g_err.c_str(); }
int mlau_num_kernels(void) { return K_COUNT; }

int mlau_create(const mlau_grid *grid, const mlau_scheme *scheme, mlau_handle **out)
{
  if (!out) return fail(MLAU_ERR_ARG, "null output handle");
  *out = nullptr;
  int rc = mlau_validate(grid, scheme);
  if (rc) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(MLAU_ERR_HIP, "no HIP device");
  mlau_handle *h = new mlau_handle();
  h->g = *grid;
  h->s = *scheme;
  if (h->g.dim < 3) {
    h->g.nz = 1;
    h->g.zoff = 0;
    if (h->g.dz == 0.0) h->g.dz = 1.0;
  }
  if (h->g.dim < 2) {
    h->g.ny = 1;
    h->g.yoff = 0;
    if (h->g.dy == 0.0) h->g.dy = 1.0;
  }
  h->nf = (h->g.dim == 1) ? 7 : 8;
  h->ncell = (long long)h->g.nx * h->g.ny * h->g.nz;
  for (int n = 0; n < 8; n++) {
    h->bc.dnx[n] = scheme->dnxs[n];
    h->bc.dny[n] = scheme->dnys[n];
    h->bc.dnz[n] = scheme->dnzs[n];
    h->bc.stx[n] = (n == 5);
    h->bc.sty[n] = (n == 6);
    h->bc.stz[n] = (n == 7);
  }
  if (h->g.dim == 2 && h->g.nx - 2 * h->g.xoff >= 2 * h->g.xoff && h->g.ny - 2 * h->g.yoff >= 2 * h->g.yoff) {
    h->periodic_images = true;
    for (int n = 0; n < 8; n++)
      if (h->bc.dnx[n] || h->bc.dny[n]) h->periodic_images = false;
  }
  hipGetDevice(&h->device);
  if (const char *t = std::getenv("MLAU_TUNE")) h->tune = std::atoi(t);
  if (const char *t = std::getenv("MLAU_NO_GRAPH")) h->graph_off = std::atoi(t) != 0;
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return fail(MLAU_ERR_HIP, "hipStreamCreate failed");
  }
  if (h->g.dim >= 2) {  // side streams of the y and z sweeps
    bool ok = hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming) == hipSuccess;
    for (int k = 0; k < 2 && ok; k++)
      ok = hipStreamCreateWithFlags(&h->side[k], hipStreamNonBlocking) == hipSuccess &&
           hipEventCreateWithFlags(&h->ev_join[k], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
      release_side(h);
      hipStreamDestroy(h->stream);
      delete h;
      return fail(MLAU_ERR_HIP, "side stream or event creation failed");
    }
  }
  // array budget: 2x state, prims, aux, 3x8 fluxes, 3x2 central parts, 3 edges
  const int nfields = 16 + 8 + 6 + 24 + 6 + 3;
  const size_t per = ((size_t)h->ncell * sizeof(double) + 255) & ~(size_t)255;
  h->base_bytes = per * nfields + 4096;
  if (hipMalloc(&h->base, h->base_bytes) != hipSuccess) {
    release_side(h);
    hipStreamDestroy(h->stream);
    delete h;
    return fail(MLAU_ERR_NOMEM, "hipMalloc of the resident workspace failed");
  }
  char *p = (char *)h->base;
  auto take = [&]() {
    double *r = (double *)p;
    p += per;
    return r;
  };
  for (int n = 0; n < 8; n++) h->U0[n] = take();
  for (int n = 0; n < 8; n++) h->U1[n] = take();
  for (int n = 0; n < 8; n++) h->prim[n] = take();
  for (int n = 0; n < 6; n++) h->aux[n] = take();
  for (int n = 0; n < 8; n++) h->fx[n] = take();
  for (int n = 0; n < 8; n++) h->fy[n] = take();
  for (int n = 0; n < 8; n++) h->fz[n] = take();
  for (int n = 0; n < 2; n++) h->fcx[n] = take();
  for (int n = 0; n < 2; n++) h->fcy[n] = take();
  for (int n = 0; n < 2; n++) h->fcz[n] = take();
  for (int n = 0; n < 3; n++) h->e[n] = take();
  h->info = (StepInfo *)p;
  h->lr = h->fy[0];  // 1D face states reuse the y-flux block (14 arrays available)
  if (hipMemsetAsync(h->base, 0, h->base_bytes, h->stream) != hipSuccess) {
    mlau_destroy(h);
    return fail(MLAU_ERR_HIP, "hipMemset of the workspace failed");
  }
  hipStreamSynchronize(h->stream);
  *out = h;
  return MLAU_OK;
}

void mlau_destroy(mlau_handle *h)
{
  if (!h) return;
  if (h->stream) hipStreamSynchronize(h->stream);
  for (auto &r : h->timing.pending) {
    hipEventDestroy(r.a);
    hipEventDestroy(r.b);
  }
  for (auto e : h->timing.pool) hipEventDestroy(e);
  if (h->gexec) hipGraphExecDestroy(h->gexec);
  if (h->base) hipFree(h->base);
  release_side(h);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
}

int mlau_upload(mlau_handle *h, const double *const *fields)
{
  if (!h || !fields) return fail(MLAU_ERR_ARG, "null argument");
  const size_t bytes = (size_t)h->ncell * sizeof(double);
  for (int n = 0; n < h->nf; n++) {
    if (!fields[n]) return fail(MLAU_ERR_ARG, "null field pointer");
    HIPCHK(hipMemcpyAsync(h->U0[n], fields[n], bytes, hipMemcpyHostToDevice, h->stream));
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  return MLAU_OK;
}

int mlau_download(mlau_handle *h, double *const *fields)
{
  if (!h || !fields) return fail(MLAU_ERR_ARG, "null argument");
  const size_t bytes = (size_t)h->ncell * sizeof(double);
  for (int n = 0; n < h->nf; n++) {
    if (!fields[n]) return fail(MLAU_ERR_ARG, "null field pointer");
    HIPCHK(hipMemcpyAsync(fields[n], h->U0[n], bytes, hipMemcpyDeviceToHost, h->stream));
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  return MLAU_OK;
}

int mlau_device_fields(mlau_handle *h, double **fields)
{
  if (!h || !fields) return fail(MLAU_ERR_ARG, "null argument");
  for (int n = 0; n < h->nf; n++) fields[n] = h->U0[n];
  return MLAU_OK;
}

int mlau_set_dt(mlau_handle *h, double dt)
{
  if (!h) return fail(MLAU_ERR_ARG, "null handle");
  hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, h->stream, h->info, dt, h->g.dx, h->g.dy, h->g.dz);
  HIPCHK(hipGetLastError());
  return MLAU_OK;
}

static int dim_steps(mlau_handle *h, int nsteps, bool cfl_mode, double cfl, double vfloor)
{
  switch (h->g.dim) {
  case 1: return run_steps_1d(h, nsteps, cfl_mode, cfl, vfloor);
  case 2: return run_steps_2d(h, nsteps, cfl_mode, cfl, vfloor);
  default: return run_steps_3d(h, nsteps, cfl_mode, cfl, vfloor);
  }
}

// Step loops of kGraphSteps or more steps replay a captured HIP graph of
// kGraphSteps steps (the same kernels with the same arguments; dt and time
// live on the device), then launch the remainder eagerly.  Per-kernel timing
// uses eager launches only.
static const int kGraphSteps = 8;

static int graph_for(mlau_handle *h, bool cfl_mode, double cfl, double vfloor)
{
  if (h->gexec && h->g_cfl_mode == cfl_mode && h->g_cfl == cfl && h->g_vfloor == vfloor) return MLAU_OK;
  if (h->gexec) hipGraphExecDestroy(h->gexec);
  h->gexec = nullptr;
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  const int e = dim_steps(h, kGraphSteps, cfl_mode, cfl, vfloor);
  const hipError_t ec = hipStreamEndCapture(h->stream, &g);
  if (e || ec != hipSuccess) {
    if (g) hipGraphDestroy(g);
    return hip_fail(e ? (hipError_t)e : ec, "step graph capture");
  }
  const hipError_t ei = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (ei != hipSuccess) {
    h->gexec = nullptr;
    return hip_fail(ei, "hipGraphInstantiate");
  }
  h->g_cfl_mode = cfl_mode;
  h->g_cfl = cfl;
  h->g_vfloor = vfloor;
  return MLAU_OK;
}

static int run_steps(mlau_handle *h, int nsteps, bool cfl_mode, double cfl, double vfloor)
{
  int left = nsteps;
  if (!h->timing.enabled && !h->graph_off && left >= kGraphSteps) {
    const int rc = graph_for(h, cfl_mode, cfl, vfloor);
    if (rc) return rc;
    while (left >= kGraphSteps) {
      HIPCHK(hipGraphLaunch(h->gexec, h->stream));
      left -= kGraphSteps;
    }
  }
  if (left > 0) {
    const int e = dim_steps(h, left, cfl_mode, cfl, vfloor);
    if (e) return hip_fail((hipError_t)e, "kernel launch");
  }
  return MLAU_OK;
}

int mlau_step(mlau_handle *h, int nsteps)
{
  if (!h) return fail(MLAU_ERR_ARG, "null handle");
  if (nsteps < 0) return fail(MLAU_ERR_ARG, "negative step count");
  return run_steps(h, nsteps, false, 0.0, 0.0);
}

int mlau_step_cfl(mlau_handle *h, int nsteps, double cfl, double vfloor)
{
  if (!h) return fail(MLAU_ERR_ARG, "null handle");
  if (nsteps < 0) return fail(MLAU_ERR_ARG, "negative step count");
  return run_steps(h, nsteps, true, cfl, vfloor);
}

int mlau_cfl(mlau_handle *h, double cfl, double vfloor, double *dt_out, double *vmax_out)
{
  if (!h) return fail(MLAU_ERR_ARG, "null handle");
  int e;
  switch (h->g.dim) {
  case 1: e = run_cfl_1d(h, cfl, vfloor); break;
  case 2: e = run_cfl_2d(h, cfl, vfloor); break;
  default: e = run_cfl_3d(h, cfl, vfloor); break;
  }
  if (e) return hip_fail((hipError_t)e, "cfl launch");
  if (dt_out || vmax_out) {
    StepInfo si;
    HIPCHK(hipMemcpyAsync(&si, h->info, sizeof(si), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (dt_out) *dt_out = si.dt;
    if (vmax_out) *vmax_out = si.vmax;
  }
  return MLAU_OK;
}

int mlau_get_time(mlau_handle *h, double *tim, double *dt, long long *nstep)
{
  if (!h) return fail(MLAU_ERR_ARG, "null handle");
  StepInfo si;
  HIPCHK(hipMemcpyAsync(&si, h->info, sizeof(si), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  if (tim) *tim = si.tim;
  if (dt) *dt = si.dt;
  if (nstep) *nstep = si.nstep;
  return MLAU_OK;
}

int mlau_set_time(mlau_handle *h, double tim, long long nstep)
{
  if (!h) return fail(MLAU_ERR_ARG, "null handle");
  hipLaunchKernelGGL(k_set_time, dim3(1), dim3(1), 0, h->stream, h->info, tim, nstep);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(h->stream));
  return MLAU_OK;
}

int mlau_sync(mlau_handle *h)
{
  if (!h) return fail(MLAU_ERR_ARG, "null handle");
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipGetLastError());
  return MLAU_OK;
}

int mlau_kernel_timing(mlau_handle *h, int enable)
{
  if (!h) return fail(MLAU_ERR_ARG, "null handle");
  timing_flush(h);
  h->timing.enabled = enable != 0;
  if (enable) {
    for (int k = 0; k < K_COUNT; k++) {
      h->timing.total_ms[k] = 0.0;
      h->timing.count[k] = 0;
    }
  }
  return MLAU_OK;
}

int mlau_kernel_stats(mlau_handle *h, int kernel, double *total_ms, long long *launches, const char **name)
{
  if (!h || kernel < 0 || kernel >= K_COUNT) return fail(MLAU_ERR_ARG, "bad kernel index");
  timing_flush(h);
  if (total_ms) *total_ms = h->timing.total_ms[kernel];
  if (launches) *launches = h->timing.count[kernel];
  if (name) *name = kKernelNames[kernel];
  return MLAU_OK;
}

// ----------------------------------------------------------- one-step drop-ins
// A per-thread cached handle keyed by shape and scheme, so repeated calls
// (the reference drivers call the solver once per step) reuse the workspace.
struct Cached {
  mlau_handle *h = nullptr;
  mlau_grid g{};
  mlau_scheme s{};
};
static thread_local Cached g_cache;

static int cached_handle(const mlau_grid &g, const mlau_scheme &s, mlau_handle **out)
{
  if (g_cache.h && std::memcmp(&g_cache.g, &g, sizeof(g)) == 0 && std::memcmp(&g_cache.s, &s, sizeof(s)) == 0) {
    *out = g_cache.h;
    return MLAU_OK;
  }
  if (g_cache.h) mlau_destroy(g_cache.h);
  g_cache.h = nullptr;
  int rc = mlau_create(&g, &s, &g_cache.h);
  if (rc) return rc;
  g_cache.g = g;
  g_cache.s = s;
  *out = g_cache.h;
  return MLAU_OK;
}

static int one_step(const mlau_grid &g, const mlau_scheme &s, double dt, double **f)
{
  mlau_handle *h;
  int rc = cached_handle(g, s, &h);
  if (rc) return rc;
  if ((rc = mlau_upload(h, f))) return rc;
  if ((rc = mlau_set_dt(h, dt))) return rc;
  if ((rc = mlau_step(h, 1))) return rc;
  return mlau_download(h, f);
}

int mlau_fd4c_1d(double *ro, double *mx, double *my, double *mz, double *en, double *by, double *bz, double bx,
                 double dt, double dx, int nx, int xoff, double gamma, const mlau_scheme *scheme)
{
  if (!scheme) return fail(MLAU_ERR_ARG, "null scheme");
  mlau_grid g{};
  g.dim = 1; g.nx = nx; g.ny = 1; g.nz = 1; g.xoff = xoff; g.dx = dx; g.dy = 1.0; g.dz = 1.0;
  g.gamma = gamma; g.bx = bx;
  double *f[7] = {ro, mx, my, mz, en, by, bz};
  return one_step(g, *scheme, dt, f);
}

int mlau_fd_ct_2d(double *ro, double *mx, double *my, double *mz, double *en, double *bx, double *by, double *bz,
                  double dt, double dx, double dy, int nx, int ny, int xoff, int yoff, double gamma,
                  const mlau_scheme *scheme)
{
  if (!scheme) return fail(MLAU_ERR_ARG, "null scheme");
  mlau_grid g{};
  g.dim = 2; g.nx = nx; g.ny = ny; g.nz = 1; g.xoff = xoff; g.yoff = yoff; g.dx = dx; g.dy = dy; g.dz = 1.0;
  g.gamma = gamma;
  double *f[8] = {ro, mx, my, mz, en, bx, by, bz};
  return one_step(g, *scheme, dt, f);
}

int mlau_fd_ct_3d(double *ro, double *mx, double *my, double *mz, double *en, double *bx, double *by, double *bz,
                  double dt, double dx, double dy, double dz, int nx, int ny, int nz, int xoff, int yoff, int zoff,
                  double gamma, const mlau_scheme *scheme)
{
  if (!scheme) return fail(MLAU_ERR_ARG, "null scheme");
  mlau_grid g{};
  g.dim = 3; g.nx = nx; g.ny = ny; g.nz = nz; g.xoff = xoff; g.yoff = yoff; g.zoff = zoff;
  g.dx = dx; g.dy = dy; g.dz = dz; g.gamma = gamma;
  double *f[8] = {ro, mx, my, mz, en, bx, by, bz};
  return one_step(g, *scheme, dt, f);
}

}  // extern "C"
