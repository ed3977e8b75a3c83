/*
 * mlau_server.h -- solve server of libmlau_hip: continuous batching of
 * independent MHD problems on one GPU.
 *
 * Clients submit solve requests (a member description, see mlau_ensemble.h,
 * plus its initial fields) from any thread.  A worker thread owns the GPU: it
 * admits queued requests in arrival order whenever the device-memory budget
 * has room (a request that does not fit lets smaller later ones in; it keeps
 * its place at the head), advances every admitted request a few steps at a
 * time through the batched ensemble engine, and hands finished results back.
 * Requests that arrive while others run join the running batch at the next
 * step boundary.  Each result is bit-identical to solving the request alone.
 *
 * The reference has no counterpart (it solves one problem per process, e.g.
 * 2D/SERIAL/OTvortex/main.cpp:26-43); per request the arithmetic and time
 * stepping are those of the ensemble engine.
 */
#ifndef MLAU_SERVER_H
#define MLAU_SERVER_H

#include "mlau_ensemble.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mlau_server mlau_server;

typedef struct mlau_srv_times {
  double submit_s;    /* seconds since server start */
  double admit_s;     /* placed on the GPU */
  double finish_s;    /* result downloaded */
  long long nstep;    /* time steps taken */
  long long zone_updates;
} mlau_srv_times;

/* Start a server with a device-memory budget; steps_per_round is how many
 * steps run between admission points (the latency/overhead trade-off). */
int mlau_srv_create(long long budget_bytes, int steps_per_round, mlau_server **out);
/* Drain: finish every admitted and queued request, then stop the worker. */
void mlau_srv_destroy(mlau_server *s);
/* Queue a request; the fields are copied, so the caller may reuse them. */
int mlau_srv_submit(mlau_server *s, const mlau_member *m, const double *const *fields, int *ticket);
/* Block until the request is finished (or timeout_s passes: MLAU_ERR_STATE).
 * *status receives MLAU_MEMBER_DONE / FAILED / CAPPED. */
int mlau_srv_wait(mlau_server *s, int ticket, double timeout_s, int *status);
/* Copy a finished request's fields out, give its times, and forget it.  For
 * a failed request the fields (if it ran) and times are still given and the
 * call returns MLAU_ERR_STATE with the reason in mlau_last_error(). */
int mlau_srv_result(mlau_server *s, int ticket, double *const *fields, mlau_srv_times *times);
/* Counters: requests queued, running, finished-not-collected; bytes admitted. */
int mlau_srv_load(mlau_server *s, int *queued, int *running, int *finished, long long *bytes);
/* Seconds since server start (the clock of mlau_srv_times). */
double mlau_srv_clock(mlau_server *s);

#ifdef __cplusplus
}
#endif
#endif
