/*
 * mlau_ensemble.h -- batched ensemble engine of libmlau_hip.
 *
 * Many independent problems ("members") advance together on one GPU: every
 * kernel of a step is launched once for all running members, each member
 * with its own grid size, Riemann solver, boundary flags, CFL time step and
 * end time.  State stays resident in HBM.  Members can be added and removed
 * between calls (continuous batching); device memory is admitted against a
 * byte budget fixed at creation.
 *
 * Per member the arithmetic is exactly the single-grid path's
 * (mlau_create/mlau_step of mlau.h): a member gives the same bits alone, in
 * any mix and under any launch tuning.
 *
 * Time stepping replaces, per member, the reference driver loop
 * "cflcheck; mhd_fd_ct_2d; t += dt" (2D/SERIAL/OTvortex/main.cpp:26-43 with
 * cflcheck.hpp:1-32) run to an end time.  The reference has no end-time rule
 * (it runs a fixed number of steps); the rule here is:
 *     vmax = max(vfloor, max over interior cells of |v_perp| + c_f)
 *     dt   = cfl * min(dx, dy) / vmax
 *     if (tim + dt >= tend) { dt = tend - tim; last step }
 *     step with dt;  tim = last ? tend : tim + dt
 * A member whose vmax is not finite (NaN or Inf in its state) stops with
 * MLAU_MEMBER_FAILED; other members are unaffected.  The 1D members keep the
 * 1D driver's rule instead: dt from the initial state only
 * (1D/SERIAL/shock/main.cpp:113-128), clipped at tend the same way.
 *
 * Members of dimension 1, 2 and 3 may be mixed; each group of equal
 * (dim, odr, rk, ctw, precision, flux) shares its launches.
 *
 * Overlapping transfers with compute: mlau_ens_launch queues steps and
 * returns at once; until mlau_ens_sync, the caller may add members (their
 * fields are copied while the steps run and they join at the next launch),
 * and may download, query and remove members that were not running at the
 * last sync (no launched step touches them).  Calls on a running member, and
 * the tuning calls, first wait for the launched steps.  Calls on one
 * ensemble must not overlap (one host thread at a time).
 */
#ifndef MLAU_ENSEMBLE_H
#define MLAU_ENSEMBLE_H

#include "mlau.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mlau_ensemble mlau_ensemble;

typedef struct mlau_member {
  mlau_grid grid;       /* dim, sizes with ghosts, spacing, gamma (as for mlau_create) */
  mlau_scheme scheme;   /* odr, rk, ctw, precision and flux group members into shared
                           launches; the dn*s boundary flags are per member */
  double tend;          /* end time */
  double cfl;           /* CFL number */
  double vfloor;        /* vmax starts from this speed (cflcheck's vfast) */
  long long max_steps;  /* stop after this many steps (0: no limit) */
} mlau_member;

enum {
  MLAU_MEMBER_RUNNING = 0,
  MLAU_MEMBER_DONE = 1,    /* reached tend */
  MLAU_MEMBER_FAILED = 2,  /* vmax not finite: NaN/Inf in the state */
  MLAU_MEMBER_CAPPED = 3   /* hit max_steps before tend */
};

enum {
  MLAU_ENS_GRAPH = 1,   /* capture each group's step loop into a HIP graph */
  MLAU_ENS_BALANCE = 2, /* groups run on their own streams side by side anyway; with this flag a
                           group whose steps are cheaper takes proportionally more of them per
                           call (nsteps times the slowest group's step time over its own, at most
                           64 times nsteps), so cheap groups advance while expensive ones run.
                           Results per member do not change. */
  MLAU_ENS_UNFUSED = 4  /* 2D members with periodic boundaries, MUSCL (odr 2), rk >= 2, ghost widths
                           >= 4 and interiors of at least two ghost widths run each RK stage as one
                           fused kernel (primitives, both sweeps, corner E-field, update and ghost
                           images of a 28 x 28 tile in LDS) by default, and 1D members at odr <= 2
                           with rk >= 2 run the stages after the first as one launch each; this flag
                           runs both through the separate stage kernels instead.  FP64 results are
                           bitwise the same;
                           the mixed forms agree within their tolerance (the compiler contracts
                           FP32 multiply-adds differently in the two kernels). */
};

/* Create an engine that may hold up to budget_bytes of member memory. */
int mlau_ens_create(long long budget_bytes, int flags, mlau_ensemble **out);
void mlau_ens_destroy(mlau_ensemble *e);
/* Device bytes a member takes (state, stage state and per-stage scratch). */
long long mlau_ens_member_bytes(const mlau_member *m);
/* Bytes admitted now (sum over members not yet removed). */
long long mlau_ens_used_bytes(const mlau_ensemble *e);
/* Device bytes the engine holds for member memory (its arena chunks), never
 * more than the budget.  Per-group tables (records, states, work lists: about
 * 4 MB per group of members with one scheme) are outside the budget. */
long long mlau_ens_reserved_bytes(const mlau_ensemble *e);
/* Add a member with its initial fields (mlau_upload layout).  Returns
 * MLAU_ERR_NOMEM without side effects when the budget would be exceeded
 * (also when the budget's free memory is too fragmented to hold it), and
 * MLAU_ERR_HIP when the device refuses an allocation within the budget. */
int mlau_ens_add(mlau_ensemble *e, const mlau_member *m, const double *const *fields, int *id);
/* Release a member that is no longer running (its memory is reused). */
int mlau_ens_remove(mlau_ensemble *e, int id);
/* Advance every running member by up to nsteps steps (more under
 * MLAU_ENS_BALANCE); *nrunning (may be NULL) receives the number of members
 * still running afterwards.  With MLAU_ENS_GRAPH, nsteps = 0 only (re)builds
 * the graphs for the current members' groups and tuning. */
int mlau_ens_run(mlau_ensemble *e, int nsteps, int *nrunning);
/* mlau_ens_run in two halves: queue the steps (MLAU_ERR_STATE if steps are
 * already in flight), then wait for them and refresh the statuses. */
int mlau_ens_launch(mlau_ensemble *e, int nsteps);
int mlau_ens_sync(mlau_ensemble *e, int *nrunning);
/* 1 when mlau_ens_sync would not wait (no steps in flight, or they are
 * done), 0 while launched steps still run, -MLAU_ERR_* on error. */
int mlau_ens_ready(mlau_ensemble *e);
/* Call fn(arg) once the launched steps are done: at once when none are in
 * flight, otherwise from a HIP runtime thread (fn must not call HIP or this
 * library; it may signal the caller's thread). */
int mlau_ens_on_ready(mlau_ensemble *e, void (*fn)(void *), void *arg);
/* Member status as of the end of the last mlau_ens_run or mlau_ens_sync. */
int mlau_ens_status(mlau_ensemble *e, int id, int *status, double *tim, double *dt, long long *nstep);
/* Ids of members (not removed) whose status is no longer RUNNING. */
int mlau_ens_finished(mlau_ensemble *e, int *ids, int cap, int *n);
int mlau_ens_download(mlau_ensemble *e, int id, double *const *fields);
/* Launch shape: bits 0-1 y-sweep chunk (0: 8 rows, 1: 16, 2: 4), bit 2 y-sweep
 * before x-sweep, bits 4-7 persistent blocks per CU (0: 2), bit 8 groups'
 * steps one after another instead of concurrently, bits 9-10 CUs left out of
 * the persistent stage kernels (0, 8, 16, 32), bits 11-12 the most steps
 * MLAU_ENS_BALANCE gives a cheap group per step of the costliest (64, 256,
 * 1024).  Results do not depend on it.  The environment variable
 * MLAU_ENS_TUNE sets it when an ensemble is created. */
int mlau_ens_set_tune(mlau_ensemble *e, int tune);
/* Time candidate launch shapes on the current members (state is restored
 * afterwards) and keep the fastest; *tune receives it. */
int mlau_ens_autotune(mlau_ensemble *e, int steps, int *tune);
/* Per-kernel HIP-event timing of eager (non-graph) runs. */
int mlau_ens_kernel_timing(mlau_ensemble *e, int enable);
int mlau_ens_kernel_stats(mlau_ensemble *e, int kernel, double *total_ms, long long *launches, const char **name);
int mlau_ens_num_kernels(void);
/* Zone-updates done so far (interior cells x steps, summed over members). */
long long mlau_ens_zone_updates(mlau_ensemble *e);

#ifdef __cplusplus
}
#endif
#endif
