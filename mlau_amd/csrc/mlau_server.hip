// mlau_server.hip -- solve server: request queue, admission under a device
// memory budget, continuous batching through the ensemble engine, results
// and timestamps.  Host code only; a worker thread owns the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <cstdlib>
#include <sys/mman.h>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mlau_server.h"
#include "mlau_internal.h"

namespace {

// Launch shape of the server's engine: 16 CUs left out of the persistent stage
// kernels, so the short kernels of the 1D group and of the groups between
// stages start at once, and up to 256 steps of a cheap group per step of the
// costliest (the 1D requests take thousands of steps each).  Measured on the
// serving stream against the default shape: +3 % zone-updates/s.
// MLAU_ENS_TUNE, when set, wins.
int server_tune(mlau_ensemble *ens)
{
  if (std::getenv("MLAU_ENS_TUNE")) return MLAU_OK;
  return mlau_ens_set_tune(ens, (1 << 11) | (2 << 9));
}

// Host buffer of a request: 2 MiB aligned and marked for transparent huge
// pages, so the copy at submit faults in 2 MiB pages instead of 4 KiB ones
// (several times faster on first touch).  Not zero-filled.
struct HostBuf {
  void operator()(double *p) const { free(p); }
};
double *host_buf(size_t n)
{
  const size_t align = (size_t)2 << 20;
  const size_t bytes = (n * sizeof(double) + align - 1) / align * align;
  void *p = nullptr;
  if (posix_memalign(&p, align, bytes) != 0) return nullptr;
  madvise(p, bytes, MADV_HUGEPAGE);  // advice only: ignored where huge pages are off
  return (double *)p;
}

struct Req {
  int ticket = 0;
  mlau_member m{};
  int nf = 0;
  size_t cells = 0;
  std::unique_ptr<double[], HostBuf> data;  // nf * cells: the initial fields, then the result
  int member_id = -1;
  int state = 0;             // 0 queued, 1 on the GPU, 2 finished
  int status = MLAU_MEMBER_RUNNING;
  long long bytes = 0;
  mlau_srv_times t{};
  std::string err;
};

double seconds_since(std::chrono::steady_clock::time_point t0)
{
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

using mlau::fail;

struct mlau_server {
  std::mutex mu;
  std::condition_variable cv_work, cv_done, cv_init;
  std::deque<int> queue;          // queued tickets in arrival order
  std::map<int, Req *> reqs;      // every request not yet collected
  std::map<int, int> by_member;   // ensemble member id -> ticket
  std::thread worker;
  bool stop = false;
  int next_ticket = 0;
  unsigned long long arrivals = 0;  // submissions so far (wakes the worker mid-round)
  bool round_done = false;          // the launched round finished (set from a HIP runtime thread)

  static void round_end(void *p)
  {
    mlau_server *s = (mlau_server *)p;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->round_done = true;
    }
    s->cv_work.notify_all();
  }
  int device = 0;
  long long budget = 0;
  int steps = 8;
  double hold_s = 1.0;            // a head request waiting longer than this blocks backfill
  std::chrono::steady_clock::time_point t0;
  mlau_ensemble *ens = nullptr;
  bool init_done = false;
  int init_rc = 0;
  std::string init_err;
  int nrunning = 0;
  std::string broken;             // set when the engine could not be rebuilt: submissions are refused

  void finish(Req *r, int status, const std::string &err)
  {
    r->state = 2;
    r->status = status;
    r->err = err;
    r->t.finish_s = seconds_since(t0);
  }

  // Try to place queued requests on the GPU (worker thread).  Arrival order;
  // a request that waits for memory lets later ones pass unless it has
  // waited longer than hold_s, so large requests are not starved.
  void admit()
  {
    std::vector<Req *> cand;
    {
      std::lock_guard<std::mutex> lk(mu);
      for (int tk : queue) {
        auto rq = reqs.find(tk);
        if (rq != reqs.end()) cand.push_back(rq->second);
      }
    }
    const double now = seconds_since(t0);
    for (size_t k = 0; k < cand.size(); k++) {
      Req *r = cand[k];
      std::vector<const double *> p(r->nf);
      for (int n = 0; n < r->nf; n++) p[n] = r->data.get() + (size_t)n * r->cells;
      int id = -1;
      const int rc = mlau_ens_add(ens, &r->m, p.data(), &id);
      std::lock_guard<std::mutex> lk(mu);
      if (rc == MLAU_OK) {
        r->state = 1;
        r->member_id = id;
        r->t.admit_s = seconds_since(t0);
        by_member[id] = r->ticket;
        nrunning++;
        queue.erase(std::find(queue.begin(), queue.end(), r->ticket));
      } else if (rc == MLAU_ERR_NOMEM && r->bytes <= budget) {
        if (now - r->t.submit_s > hold_s) break;  // stop backfilling past a long-waiting request
      } else {
        finish(r, MLAU_MEMBER_FAILED, mlau_last_error());
        queue.erase(std::find(queue.begin(), queue.end(), r->ticket));
        cv_done.notify_all();
      }
    }
  }

  // Collect finished members (worker thread).
  void harvest()
  {
    int n = 0;
    if (mlau_ens_finished(ens, nullptr, 0, &n) || n == 0) return;
    std::vector<int> ids(n);
    mlau_ens_finished(ens, ids.data(), n, &n);
    for (int id : ids) {
      Req *r = nullptr;
      {
        std::lock_guard<std::mutex> lk(mu);
        auto bm = by_member.find(id);
        if (bm != by_member.end()) {
          auto rq = reqs.find(bm->second);
          if (rq != reqs.end()) r = rq->second;
        }
      }
      if (!r) {  // no request waits for this member any more: release its memory
        mlau_ens_remove(ens, id);
        continue;
      }
      int st = MLAU_MEMBER_FAILED;
      double tim, dt;
      long long ns = 0;
      mlau_ens_status(ens, id, &st, &tim, &dt, &ns);
      std::vector<double *> p(r->nf);
      for (int k = 0; k < r->nf; k++) p[k] = r->data.get() + (size_t)k * r->cells;
      std::string err;
      if (mlau_ens_download(ens, id, p.data())) err = mlau_last_error();
      if (err.empty() && st == MLAU_MEMBER_FAILED)
        err = "state not finite (NaN or Inf) at step " + std::to_string(ns) + "; fields hold that state";
      mlau_ens_remove(ens, id);
      long long interior = (long long)(r->m.grid.nx - 2 * r->m.grid.xoff);
      if (r->m.grid.dim >= 2) interior *= (r->m.grid.ny - 2 * r->m.grid.yoff);
      if (r->m.grid.dim == 3) interior *= (r->m.grid.nz - 2 * r->m.grid.zoff);
      std::lock_guard<std::mutex> lk(mu);
      r->t.nstep = ns;
      r->t.zone_updates = interior * ns;
      by_member.erase(id);
      nrunning--;
      finish(r, err.empty() ? st : MLAU_MEMBER_FAILED, err);
      cv_done.notify_all();
    }
  }

  // A launch or sync failed: the requests on the GPU fail with the error, and
  // the engine is rebuilt so that their members release their memory (they
  // may still count as running and cannot be removed one by one).  Queued
  // requests stay queued for the new engine; if it cannot be made, they fail
  // too and later submissions are refused.
  void fail_running(const std::string &err)
  {
    {
      std::lock_guard<std::mutex> lk(mu);
      for (auto &kv : by_member) {
        auto rq = reqs.find(kv.second);
        if (rq != reqs.end()) finish(rq->second, MLAU_MEMBER_FAILED, err);
      }
      by_member.clear();
      nrunning = 0;
      cv_done.notify_all();
    }
    mlau_ens_destroy(ens);
    ens = nullptr;
    int rc = mlau_ens_create(budget, MLAU_ENS_GRAPH | MLAU_ENS_BALANCE, &ens);
    if (!rc) rc = server_tune(ens);
    if (rc) {
      const std::string why = "server engine lost after: " + err + "; " + mlau_last_error();
      std::lock_guard<std::mutex> lk(mu);
      ens = nullptr;
      broken = why;
      for (int tk : queue) {
        auto rq = reqs.find(tk);
        if (rq != reqs.end()) finish(rq->second, MLAU_MEMBER_FAILED, why);
      }
      queue.clear();
      cv_done.notify_all();
    }
  }

  void run()
  {
    hipSetDevice(device);
    int rc = mlau_ens_create(budget, MLAU_ENS_GRAPH | MLAU_ENS_BALANCE, &ens);
    if (!rc) rc = server_tune(ens);
    {
      std::lock_guard<std::mutex> lk(mu);
      init_rc = rc;
      if (rc) init_err = mlau_last_error();
      init_done = true;
    }
    cv_init.notify_all();
    if (rc) return;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] { return stop || !queue.empty() || nrunning > 0; });
        if ((stop && queue.empty() && nrunning == 0) || !ens) break;
      }
      bool any;
      {
        std::lock_guard<std::mutex> lk(mu);
        any = nrunning > 0;
      }
      if (!any) {
        admit();  // nothing on the GPU to overlap with
        std::lock_guard<std::mutex> lk(mu);
        any = nrunning > 0;
      }
      if (!any) {
        std::unique_lock<std::mutex> lk(mu);  // queued requests wait for memory that never frees: give up on them
        for (int tk : queue) {
          auto rq = reqs.find(tk);
          if (rq != reqs.end()) finish(rq->second, MLAU_MEMBER_FAILED, "request does not fit in the budget");
        }
        queue.clear();
        cv_done.notify_all();
        continue;
      }
      // One round: queue the steps, then, while they run, download and
      // release the members that had finished before them and upload new
      // arrivals (these join at the next round).
      if (mlau_ens_launch(ens, steps)) {
        fail_running(mlau_last_error());
        continue;
      }
      {
        std::lock_guard<std::mutex> lk(mu);
        round_done = false;
      }
      if (mlau_ens_on_ready(ens, &mlau_server::round_end, this)) {
        std::lock_guard<std::mutex> lk(mu);
        round_done = true;  // no notice: fall back to waiting in the sync below
      }
      harvest();
      for (;;) {  // admit arrivals as they come until the round is done
        unsigned long long seen;
        {
          std::lock_guard<std::mutex> lk(mu);
          seen = arrivals;
        }
        admit();
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] { return stop || round_done || arrivals != seen; });
        if (stop || round_done) break;
      }
      int nrun = 0;
      if (mlau_ens_sync(ens, &nrun)) fail_running(mlau_last_error());
    }
    if (ens) mlau_ens_destroy(ens);
    ens = nullptr;
  }
};

extern "C" {

int mlau_srv_create(long long budget_bytes, int steps_per_round, mlau_server **out)
{
  if (!out || budget_bytes <= 0 || steps_per_round < 1) return fail(MLAU_ERR_ARG, "bad argument");
  *out = nullptr;
  mlau_server *s = new mlau_server();
  s->budget = budget_bytes;
  s->steps = steps_per_round;
  s->t0 = std::chrono::steady_clock::now();
  hipGetDevice(&s->device);
  s->worker = std::thread([s] { s->run(); });
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_init.wait(lk, [&] { return s->init_done; });
  if (s->init_rc) {
    const int rc = s->init_rc;
    const std::string err = s->init_err;
    lk.unlock();
    s->worker.join();
    delete s;
    return fail(rc, err);
  }
  *out = s;
  return MLAU_OK;
}

void mlau_srv_destroy(mlau_server *s)
{
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  s->worker.join();
  for (auto &kv : s->reqs) delete kv.second;
  delete s;
}

int mlau_srv_submit(mlau_server *s, const mlau_member *m, const double *const *fields, int *ticket)
{
  if (!s || !m || !fields || !ticket) return fail(MLAU_ERR_ARG, "null argument");
  const double arrival = seconds_since(s->t0);  // before the copy below, which the latency includes
  const mlau_grid &g = m->grid;
  if (g.dim < 1 || g.dim > 3 || g.nx < 1 || (g.dim >= 2 && g.ny < 1) || (g.dim == 3 && g.nz < 1))
    return fail(MLAU_ERR_ARG, "bad grid");
  Req *r = new Req();
  r->m = *m;
  r->nf = (g.dim == 1) ? 7 : 8;
  r->cells = (size_t)g.nx * (g.dim >= 2 ? g.ny : 1) * (g.dim == 3 ? g.nz : 1);
  r->bytes = mlau_ens_member_bytes(m);
  r->data.reset(host_buf((size_t)r->nf * r->cells));
  if (!r->data) {
    delete r;
    return fail(MLAU_ERR_NOMEM, "host buffer for the request");
  }
  for (int n = 0; n < r->nf; n++) {
    if (!fields[n]) {
      delete r;
      return fail(MLAU_ERR_ARG, "null field pointer");
    }
    memcpy(r->data.get() + (size_t)n * r->cells, fields[n], r->cells * sizeof(double));
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (!s->broken.empty()) {
      delete r;
      return fail(MLAU_ERR_STATE, s->broken);
    }
    r->ticket = s->next_ticket++;
    r->t.submit_s = arrival;
    s->reqs[r->ticket] = r;
    s->queue.push_back(r->ticket);
    s->arrivals++;
    *ticket = r->ticket;
  }
  s->cv_work.notify_all();
  return MLAU_OK;
}

int mlau_srv_wait(mlau_server *s, int ticket, double timeout_s, int *status)
{
  if (!s) return fail(MLAU_ERR_ARG, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return fail(MLAU_ERR_ARG, "no such ticket");
  Req *r = it->second;
  const bool ok = s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] { return r->state == 2; });
  if (!ok) return fail(MLAU_ERR_STATE, "request not finished before the timeout");
  if (status) *status = r->status;
  return MLAU_OK;
}

int mlau_srv_result(mlau_server *s, int ticket, double *const *fields, mlau_srv_times *times)
{
  if (!s) return fail(MLAU_ERR_ARG, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return fail(MLAU_ERR_ARG, "no such ticket");
  Req *r = it->second;
  if (r->state != 2) return fail(MLAU_ERR_STATE, "request not finished");
  if (fields)
    for (int n = 0; n < r->nf; n++)
      if (fields[n]) memcpy(fields[n], r->data.get() + (size_t)n * r->cells, r->cells * sizeof(double));
  if (times) *times = r->t;
  const std::string err = r->err;
  delete r;
  s->reqs.erase(it);
  if (!err.empty()) return fail(MLAU_ERR_STATE, err);
  return MLAU_OK;
}

int mlau_srv_load(mlau_server *s, int *queued, int *running, int *finished, long long *bytes)
{
  if (!s) return fail(MLAU_ERR_ARG, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  int fin = 0;
  long long b = 0;
  for (auto &kv : s->reqs) {
    fin += kv.second->state == 2;
    if (kv.second->state == 1) b += kv.second->bytes;
  }
  if (queued) *queued = (int)s->queue.size();
  if (running) *running = s->nrunning;
  if (finished) *finished = fin;
  if (bytes) *bytes = b;
  return MLAU_OK;
}

double mlau_srv_clock(mlau_server *s) { return s ? seconds_since(s->t0) : -1.0; }

}  // extern "C"
