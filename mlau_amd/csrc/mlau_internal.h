// mlau_internal.h -- handle layout and launch helpers shared by the HIP
// translation units of libmlau_hip.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "mlau_common.h"

namespace mlau {

// Kernel identifiers for the per-kernel timing table.
enum KernelId {
  K_PRIM2D = 0, K_AUX2D, K_XSWEEP2D, K_YSWEEP2D, K_EZ2D, K_UPD2D, K_BC2D, K_CFL2D,
  K_SWEEP1D, K_OPENBC1D, K_UPD1D, K_CFL1D,
  K_PRIM3D, K_AUX3D, K_XSWEEP3D, K_YSWEEP3D, K_ZSWEEP3D, K_EDGE3D, K_UPD3D, K_BC3D, K_CFL3D,
  K_FINALIZE, K_PRIMAUX2D, K_STAGE1D, K_XSW2D_INL, K_YSW2D_INL, K_COUNT
};
extern const char *kKernelNames[K_COUNT];

struct Timing {
  bool enabled = false;
  std::vector<hipEvent_t> pool;       // free events
  struct Rec { int k; hipEvent_t a, b; };
  std::vector<Rec> pending;
  double total_ms[K_COUNT] = {};
  long long count[K_COUNT] = {};
};

// Boundary flags per field (field order), plus stagger flags.
struct BcFlags {
  int dnx[8], dny[8], dnz[8];
  int stx[8], sty[8], stz[8];
};

struct Handle {
  mlau_grid g;
  mlau_scheme s;
  BcFlags bc;
  int nf;                  // 7 (1D) or 8
  long long ncell;         // nx*ny*nz including ghosts
  hipStream_t stream = nullptr;
  // 2D/3D: side streams for the y and z sweeps of a stage (see Side below)
  hipStream_t side[2] = {};
  hipEvent_t ev_fork = nullptr, ev_join[2] = {};
  int device = 0;
  double *base = nullptr;  // one device allocation carved into the arrays below
  size_t base_bytes = 0;
  double *U0[8] = {};      // state (step start / result)
  double *U1[8] = {};      // stage state
  // scratch (meaning depends on dim; see the per-dim step files)
  double *prim[8] = {};    // vx,vy,vz,pr,m2,cx,cy,cz
  double *aux[6] = {};     // ct0,ct1,ct2,dvx,dvy,dvz
  double *fx[8] = {}, *fy[8] = {}, *fz[8] = {};
  double *fcx[2] = {}, *fcy[2] = {}, *fcz[2] = {};
  double *e[3] = {};       // ex,ey,ez (2D: ez only)
  double *lr = nullptr;    // 1D face states
  StepInfo *info = nullptr;  // device
  int tune = 0;              // launch-shape variant (autotuner choice), see the per-dim drivers
  bool periodic_images = false;  // 2D periodic: the update writes the ghost images (no boundary pass)
  Timing timing;
  // HIP graph of kGraphSteps steps (replayed by mlau_step / mlau_step_cfl)
  hipGraphExec_t gexec = nullptr;
  bool graph_off = false;    // MLAU_NO_GRAPH set at creation
  bool g_cfl_mode = false;
  double g_cfl = 0.0, g_vfloor = 0.0;
};

// error reporting (message kept per host thread for mlau_last_error)
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *where);

// launch bookkeeping
void timing_begin(Handle *h, int k, hipEvent_t *a);
void timing_end(Handle *h, int k, hipEvent_t a);

#define MLAU_LAUNCH(h, kid, kernel, grid, block, ...)                                 \
  do {                                                                                \
    hipEvent_t _ev = nullptr;                                                         \
    if ((h)->timing.enabled) timing_begin((h), (kid), &_ev);                          \
    hipLaunchKernelGGL(kernel, grid, block, 0, (h)->stream, __VA_ARGS__);             \
    if ((h)->timing.enabled) timing_end((h), (kid), _ev);                             \
  } while (0)

// Side streams.  The sweeps of a stage share no output (the x sweep writes
// fx, the y sweep fy, the z sweep fz; all read the stage's input), so the y
// and z sweeps run on side streams beside the x sweep: fork() before the x
// sweep lets side stream k start after everything queued so far; launches
// between enter(k) and leave() go to side stream k; join() makes the main
// stream wait for every side stream used.  Everything stays on the main
// stream when the handle has no side streams (the no-allocation entry point),
// per-kernel timing is on (its events are read in main-stream order), tune
// bit 6 is set (kernel-by-kernel profiles), or an event call fails.
struct Side {
  Handle *h;
  hipStream_t main;
  bool on[2] = {false, false};
  explicit Side(Handle *h_) : h(h_), main(h_->stream) {}
  void fork(int n)
  {
    if (!h->ev_fork || h->timing.enabled || (h->tune & 64) || hipEventRecord(h->ev_fork, main) != hipSuccess)
      return;
    for (int k = 0; k < n && k < 2; k++)
      on[k] = h->side[k] && hipStreamWaitEvent(h->side[k], h->ev_fork, 0) == hipSuccess;
  }
  void enter(int k) { h->stream = on[k] ? h->side[k] : main; }
  void leave() { h->stream = main; }
  void join()
  {
    for (int k = 0; k < 2; k++)
      if (on[k] && (hipEventRecord(h->ev_join[k], h->side[k]) != hipSuccess ||
                    hipStreamWaitEvent(main, h->ev_join[k], 0) != hipSuccess))
        hipStreamSynchronize(h->side[k]);
  }
};

// per-dimension drivers (return hipError_t as int)
int run_steps_1d(Handle *h, int nsteps, bool cfl_mode, double cfl, double vfloor);
int run_steps_2d(Handle *h, int nsteps, bool cfl_mode, double cfl, double vfloor);
int run_steps_3d(Handle *h, int nsteps, bool cfl_mode, double cfl, double vfloor);
int run_cfl_1d(Handle *h, double cfl, double vfloor);
int run_cfl_2d(Handle *h, double cfl, double vfloor);
int run_cfl_3d(Handle *h, double cfl, double vfloor);

// Device-side dt update from the reduced vmax: dt = cfl*dr/vmax when finite.
__global__ void k_finalize_dt(StepInfo *info, double cfl, double dr, double dx, double dy, double dz);
__global__ void k_cfl_reset(StepInfo *info, double vfloor);
__global__ void k_advance_time(StepInfo *info);
__global__ void k_set_dt(StepInfo *info, double dt, double dx, double dy, double dz);

// Argument checks of mlau_create (grid and scheme).
int mlau_validate(const mlau_grid *g, const mlau_scheme *s);

}  // namespace mlau
