// mlau_ens.hip -- batched ensemble engine: member admission and memory arena,
// device planning kernels, step loop (eager or HIP graph), autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "mlau_ens.h"

namespace mlau {

static const char *kEnsKernelNames[E_COUNT] = {"ens_plan", "ens_fill", "ens_cfl", "ens_dt", "ens_prim",
                                               "ens_aux",  "ens_xsweep", "ens_ysweep", "ens_ez", "ens_update",
                                               "ens_bc",   "ens_advance", "ens_stage"};

// ------------------------------------------------------------ device plans
// Inclusive prefix sum of v over the block (kPlanThreads threads): a shuffle
// scan within each wavefront, then one over the wavefront totals; *total gets
// the block sum.  ws holds kPlanThreads/64 ints and is free again on return.
__device__ __forceinline__ int plan_scan(int v, int *ws, int *total)
{
  constexpr int nw = kPlanThreads / 64;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) ws[w] = x;
  __syncthreads();
  if (w == 0) {
    int s = (lane < nw) ? ws[lane] : 0;
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(s, o, 64);
      if (lane >= o) s += y;
    }
    if (lane < nw) ws[lane] = s;
  }
  __syncthreads();
  const int incl = x + (w > 0 ? ws[w - 1] : 0);
  *total = ws[nw - 1];
  __syncthreads();
  return incl;
}

// One block: running flags, change detection, per-kind exclusive prefix of
// the running members' tile counts (slot order), totals.
__global__ __launch_bounds__(kPlanThreads) void k_ens_plan(EnsPlan p, int ntypes)
{
  constexpr int per = kEnsSlots / kPlanThreads;
  __shared__ int ws[kPlanThreads / 64];
  __shared__ int chg;
  const int t = threadIdx.x;
  if (t == 0) chg = p.count[kEnsTypes + 1];  // rewrite requested by the host
  __syncthreads();
  int run[per];
  int nrun = 0;
  bool mine = false;
#pragma unroll
  for (int q = 0; q < per; q++) {
    const int s = t * per + q;
    run[q] = (p.stat[s] == MLAU_MEMBER_RUNNING) ? 1 : 0;
    if (run[q] != p.act[s]) {
      mine = true;
      p.act[s] = run[q];
    }
    nrun += run[q];
  }
  if (mine) chg = 1;
  for (int k = 0; k <= ntypes; k++) {  // k == ntypes: running-member count
    int c[per], sum = 0;
#pragma unroll
    for (int q = 0; q < per; q++) {
      c[q] = (k < ntypes) ? (run[q] ? p.nb[k * kEnsSlots + t * per + q] : 0) : 0;
      sum += c[q];
    }
    if (k == ntypes) sum = nrun;
    int total;
    const int incl = plan_scan(sum, ws, &total);
    if (k < ntypes) {
      int base = incl - sum;
#pragma unroll
      for (int q = 0; q < per; q++) {
        p.off[k * kEnsSlots + t * per + q] = base;
        base += c[q];
      }
      if (t == 0) p.count[k] = total;
    } else if (t == 0) {
      p.count[kEnsTypes] = total;
    }
  }
  if (t == 0) {
    p.count[kEnsTypes + 1] = 0;
    p.count[kEnsTypes + 2] = chg;
  }
}

// One block per slot: write the running member's tiles into each list.
__global__ __launch_bounds__(256) void k_ens_fill(EnsPlan p, int ntypes)
{
  if (!p.count[kEnsTypes + 2]) return;
  const int s = blockIdx.x;
  if (!p.act[s]) return;
  for (int k = 0; k < ntypes; k++) {
    const int n = p.nb[k * kEnsSlots + s];
    const int o = p.off[k * kEnsSlots + s];
    int2 *tl = p.tiles[k];
    for (int l = threadIdx.x; l < n; l += blockDim.x) tl[o + l] = make_int2(s, l);
  }
}

// dt per running member from its CFL speed, clipped at its end time.
__global__ __launch_bounds__(256) void k_ens_dt(EnsPlan p)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= kEnsSlots || !p.act[s]) return;
  EnsState &S = p.st[s];
  if (!(S.fixed && S.nstep > 0)) {
    const double vmax = __longlong_as_double((long long)S.vmax_bits);
    if (!isfinite(vmax)) {
      p.stat[s] = MLAU_MEMBER_FAILED;
      S.halt = 1;
      S.dt = S.dtdx = S.dtdy = S.dtdz = 0.0;
      return;
    }
    S.dtc = S.cfl * S.dr / vmax;
  }
  double dt = S.dtc;
  if (S.tim + dt >= S.tend) {
    dt = S.tend - S.tim;
    S.last = 1;
  }
  S.dt = dt;
  S.dtdx = dt / S.dx;
  S.dtdy = dt / S.dy;
  S.dtdz = dt / S.dz;
}

__global__ __launch_bounds__(256) void k_ens_advance(EnsPlan p)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= kEnsSlots || !p.act[s] || p.stat[s] != MLAU_MEMBER_RUNNING) return;
  EnsState &S = p.st[s];
  S.tim = S.last ? S.tend : S.tim + S.dt;
  S.nstep += 1;
  if (S.last) p.stat[s] = MLAU_MEMBER_DONE;
  else if (S.max_steps > 0 && S.nstep >= S.max_steps) p.stat[s] = MLAU_MEMBER_CAPPED;
  S.vmax_bits = (unsigned long long)__double_as_longlong(S.vfloor);
}

// ------------------------------------------------------------------ timing
void ens_timing_begin(Ensemble *e, hipEvent_t *a)
{
  Timing &t = e->timing;
  if (t.pool.empty()) {
    hipEvent_t ev;
    if (hipEventCreate(&ev) != hipSuccess) return;
    t.pool.push_back(ev);
  }
  *a = t.pool.back();
  t.pool.pop_back();
  hipEventRecord(*a, e->cur);
}

// The launches of different groups are on different streams, so the last
// pending event says nothing about the others: wait for each pair's end
// before reading it (and before its events go back to the pool).
static void ens_timing_fold(Ensemble *e)
{
  Timing &t = e->timing;
  if (t.pending.empty()) return;
  for (auto &r : t.pending) {
    float ms = 0.f;
    if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      t.total_ms[r.k] += ms;
      t.count[r.k] += 1;
    }
    t.pool.push_back(r.a);
    t.pool.push_back(r.b);
  }
  t.pending.clear();
}

void ens_timing_end(Ensemble *e, int k, hipEvent_t a)
{
  Timing &t = e->timing;
  if (!a) return;
  hipEvent_t b;
  if (t.pool.empty()) {
    if (hipEventCreate(&b) != hipSuccess) return;
  } else {
    b = t.pool.back();
    t.pool.pop_back();
  }
  hipEventRecord(b, e->cur);
  t.pending.push_back({k, a, b});
  if (t.pending.size() > 4096) ens_timing_fold(e);
}

// ------------------------------------------------------------------- arena
// Member memory comes from chunks of device memory (at least kChunk bytes,
// first fit within a chunk, coalescing free lists).  The chunks themselves are
// counted against the budget: a new chunk is only taken while the bytes of all
// chunks stay within it, and chunks that hold no member are released first
// when a new one would not fit, so fragmentation never takes device memory
// beyond the budget.  A released chunk keeps its slot (members address their
// memory by chunk index).
static const size_t kChunk = (size_t)1 << 30;
enum { ARENA_OK = 0, ARENA_FULL = 1, ARENA_HIP = 2 };

static bool chunk_empty(const Ensemble::Chunk &c)
{
  return c.base && c.free.size() == 1 && c.free.begin()->first == 0 && c.free.begin()->second == c.size;
}

static int arena_alloc(Ensemble *e, size_t bytes, size_t *where)
{
  for (size_t c = 0; c < e->chunks.size(); c++) {
    auto &fl = e->chunks[c].free;
    for (auto it = fl.begin(); it != fl.end(); ++it) {
      if (it->second >= bytes) {
        const size_t off = it->first, sz = it->second;
        fl.erase(it);
        if (sz > bytes) fl[off + bytes] = sz - bytes;
        *where = (c << 40) | off;
        return ARENA_OK;
      }
    }
  }
  auto room = [&]() { return (size_t)e->budget > e->reserved ? (size_t)e->budget - e->reserved : (size_t)0; };
  if (room() < bytes) {  // give back chunks that hold no member
    for (auto &c : e->chunks)
      if (chunk_empty(c)) {
        if (hipFree(c.base) != hipSuccess) return ARENA_HIP;
        e->reserved -= c.size;
        c = Ensemble::Chunk{};
      }
    if (room() < bytes) return ARENA_FULL;
  }
  const size_t sz = std::min(std::max(kChunk, bytes), room());
  char *p = nullptr;
  if (hipMalloc(&p, sz) != hipSuccess) {
    hipGetLastError();
    return ARENA_HIP;
  }
  size_t c = 0;
  while (c < e->chunks.size() && e->chunks[c].base) c++;
  if (c == e->chunks.size()) e->chunks.emplace_back();
  Ensemble::Chunk &ch = e->chunks[c];
  ch.base = p;
  ch.size = sz;
  ch.free.clear();
  if (sz > bytes) ch.free[bytes] = sz - bytes;
  e->reserved += sz;
  *where = (c << 40) | 0;
  return ARENA_OK;
}

static void arena_free(Ensemble *e, size_t where, size_t bytes)
{
  auto &fl = e->chunks[where >> 40].free;
  size_t off = where & (((size_t)1 << 40) - 1), sz = bytes;
  auto nx = fl.lower_bound(off);
  if (nx != fl.end() && nx->first == off + sz) {
    sz += nx->second;
    nx = fl.erase(nx);
  }
  if (nx != fl.begin()) {
    auto pv = std::prev(nx);
    if (pv->first + pv->second == off) {
      off = pv->first;
      sz += pv->second;
      fl.erase(pv);
    }
  }
  fl[off] = sz;
}

static double *arena_ptr(Ensemble *e, size_t where)
{
  return (double *)(e->chunks[where >> 40].base + (where & (((size_t)1 << 40) - 1)));
}

// ------------------------------------------------------------------ groups
static const EnsDim *dim_table(int dim)
{
  switch (dim) {
  case 1: return ens_dim1d();
  case 2: return ens_dim2d();
  case 3: return ens_dim3d();
  default: return nullptr;
  }
}

static void group_free(EnsGroup *G)
{
  if (G->gx) hipGraphExecDestroy(G->gx);
  if (G->ev_start) hipEventDestroy(G->ev_start);
  if (G->ev_done) hipEventDestroy(G->ev_done);
  if (G->stream) hipStreamDestroy(G->stream);
  hipFree(G->rec);
  hipFree(G->p.st);
  hipFree(G->p.stat);
  hipFree(G->p.act);
  hipFree(G->p.nb);
  hipFree(G->p.off);
  hipFree(G->p.count);
  for (int k = 0; k < kEnsTypes; k++) hipFree(G->p.tiles[k]);
  void *pinned[] = {G->rec_host, G->st_src, G->stat_src, G->nb_src, G->flag_src, G->stat_host, G->stat_next};
  for (void *h : pinned)
    if (h) hipHostFree(h);
  delete G;
}

static int group_get(Ensemble *e, const mlau_member &m, EnsGroup **out)
{
  const mlau_scheme &s = m.scheme;
  const int variant = dim_table(m.grid.dim)->variant(m, e->flags);
  for (EnsGroup *G : e->groups)
    if (G->dim == m.grid.dim && G->odr == s.odr && G->rk == s.rk && G->ctw == s.ctw && G->precision == s.precision &&
        G->flux == s.flux && G->variant == variant) {
      *out = G;
      return MLAU_OK;
    }
  EnsGroup *G = new EnsGroup();
  G->dim = m.grid.dim;
  G->odr = s.odr;
  G->rk = s.rk;
  G->ctw = s.ctw;
  G->precision = s.precision;
  G->flux = s.flux;
  G->variant = variant;
  G->D = dim_table(G->dim);
  G->slot_id.assign(kEnsSlots, -1);
  auto pin = [](void **p, size_t sz) {
    if (hipHostMalloc(p, sz, hipHostMallocDefault) != hipSuccess) return false;
    memset(*p, 0, sz);
    return true;
  };
  bool ok = hipMalloc(&G->rec, G->D->rec_bytes * kEnsSlots) == hipSuccess &&
            hipMalloc(&G->p.st, sizeof(EnsState) * kEnsSlots) == hipSuccess &&
            hipMalloc(&G->p.stat, sizeof(int) * kEnsSlots) == hipSuccess &&
            hipMalloc(&G->p.act, sizeof(int) * kEnsSlots) == hipSuccess &&
            hipMalloc(&G->p.nb, sizeof(int) * kEnsSlots * kEnsTypes) == hipSuccess &&
            hipMalloc(&G->p.off, sizeof(int) * kEnsSlots * kEnsTypes) == hipSuccess &&
            hipMalloc(&G->p.count, sizeof(int) * (kEnsTypes + 3)) == hipSuccess &&
            pin((void **)&G->rec_host, G->D->rec_bytes * kEnsSlots) &&
            pin((void **)&G->st_src, sizeof(EnsState) * kEnsSlots) &&
            pin((void **)&G->stat_src, sizeof(int) * kEnsSlots) &&
            pin((void **)&G->nb_src, sizeof(int) * kEnsSlots * kEnsTypes) &&
            pin((void **)&G->flag_src, sizeof(int) * (kEnsTypes + 3)) &&
            pin((void **)&G->stat_host, sizeof(int) * kEnsSlots) &&
            pin((void **)&G->stat_next, sizeof(int) * kEnsSlots) &&
            hipStreamCreateWithFlags(&G->stream, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreate(&G->ev_start) == hipSuccess && hipEventCreate(&G->ev_done) == hipSuccess;
  if (!ok) {
    hipGetLastError();
    group_free(G);
    return fail(MLAU_ERR_NOMEM, "ensemble group allocation failed");
  }
  // stat_next too: a group made while steps are in flight gets no status
  // read-back from that launch, and the sync still takes stat_next as its mirror
  for (int s2 = 0; s2 < kEnsSlots; s2++) G->stat_host[s2] = G->stat_src[s2] = G->stat_next[s2] = kStatFree;
  G->flag_src[kEnsTypes + 1] = 1;  // p.count as first written: no lists, rewrite requested
  ok = hipMemsetAsync(G->rec, 0, G->D->rec_bytes * kEnsSlots, e->stream) == hipSuccess &&
       hipMemsetAsync(G->p.st, 0, sizeof(EnsState) * kEnsSlots, e->stream) == hipSuccess &&
       hipMemcpyAsync(G->p.stat, G->stat_src, sizeof(int) * kEnsSlots, hipMemcpyHostToDevice, e->stream) ==
           hipSuccess &&
       hipMemsetAsync(G->p.act, 0, sizeof(int) * kEnsSlots, e->stream) == hipSuccess &&
       hipMemsetAsync(G->p.nb, 0, sizeof(int) * kEnsSlots * kEnsTypes, e->stream) == hipSuccess &&
       hipMemsetAsync(G->p.off, 0, sizeof(int) * kEnsSlots * kEnsTypes, e->stream) == hipSuccess &&
       hipMemcpyAsync(G->p.count, G->flag_src, sizeof(int) * (kEnsTypes + 3), hipMemcpyHostToDevice, e->stream) ==
           hipSuccess &&
       hipStreamSynchronize(e->stream) == hipSuccess;
  G->flag_src[0] = 1;  // from here on the constant that requests a rewrite
  if (!ok) {
    group_free(G);
    return fail(MLAU_ERR_HIP, "ensemble group initialisation failed");
  }
  e->groups.push_back(G);
  *out = G;
  return MLAU_OK;
}

// Grow the work lists of G to its current need (between runs only).
static int group_reserve(Ensemble *e, EnsGroup *G)
{
  for (int k = 0; k < G->D->ntypes; k++) {
    if (G->need[k] <= G->cap[k]) continue;
    long long nc = std::max<long long>({G->need[k], 2 * G->cap[k], 4096});
    if (nc > 0x7fffffffLL) return fail(MLAU_ERR_NOMEM, "ensemble work list exceeds 2^31 tiles");
    int2 *p = nullptr;
    if (hipStreamSynchronize(e->stream) != hipSuccess) return fail(MLAU_ERR_HIP, "stream sync failed");
    if (hipMalloc(&p, sizeof(int2) * nc) != hipSuccess) {
      hipGetLastError();
      return fail(MLAU_ERR_NOMEM, "ensemble work list allocation failed");
    }
    hipFree(G->p.tiles[k]);
    G->p.tiles[k] = p;
    G->cap[k] = nc;
    G->dirty = 1;
    if (G->gx) hipGraphExecDestroy(G->gx);  // it holds the old list
    G->gx = nullptr;
  }
  return MLAU_OK;
}

// Queued on the compute stream: takes effect at the next launch.
static int group_flag_rewrite(Ensemble *e, EnsGroup *G)
{
  if (hipMemcpyAsync(G->p.count + kEnsTypes + 1, G->flag_src, sizeof(int), hipMemcpyHostToDevice, e->stream) !=
      hipSuccess)
    return fail(MLAU_ERR_HIP, "ensemble flag write failed");
  G->dirty = 0;
  return MLAU_OK;
}

// Tile counts of one slot (v < 0: the member's counts for the current tuning;
// otherwise v for every kind), queued on the compute stream.
static int group_write_nb(Ensemble *e, EnsGroup *G, int slot, const mlau_member *m)
{
  for (int k = 0; k < G->D->ntypes; k++) {
    int *src = G->nb_src + k * kEnsSlots + slot;
    *src = m ? (int)G->D->tiles(*m, k, e->tune, G->variant) : 0;
    if (hipMemcpyAsync(G->p.nb + k * kEnsSlots + slot, src, sizeof(int), hipMemcpyHostToDevice, e->stream) !=
        hipSuccess)
      return fail(MLAU_ERR_HIP, "ensemble tile count write failed");
  }
  return MLAU_OK;
}

static long long member_tiles_max(const EnsDim *D, const mlau_member &m, int k, int variant)
{
  long long r = 0;
  for (int t = 0; t < 3; t++) r = std::max(r, D->tiles(m, k, t, variant));  // y-chunk variants
  return r;
}

// ------------------------------------------------------------------- graph
// n steps of group G on its stream.
static int launch_steps(Ensemble *e, EnsGroup *G, int n)
{
  e->cur = G->stream;
  for (int s = 0; s < n; s++) {
    const int rc = G->D->launch_step(e, G);
    if (rc) return rc;
  }
  return 0;
}

static void graph_drop(Ensemble *e)
{
  for (EnsGroup *G : e->groups) {
    if (G->gx) hipGraphExecDestroy(G->gx);
    G->gx = nullptr;
  }
}

// The group's graph of gsteps steps for the current tuning (captured on first use).
static int graph_get(Ensemble *e, EnsGroup *G, hipGraphExec_t *out)
{
  if (G->gx && G->gx_tune == e->tune) {
    *out = G->gx;
    return MLAU_OK;
  }
  if (G->gx) hipGraphExecDestroy(G->gx);
  G->gx = nullptr;
  hipGraph_t g = nullptr;
  if (hipStreamBeginCapture(G->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return fail(MLAU_ERR_HIP, "hipStreamBeginCapture failed");
  const int rc = launch_steps(e, G, e->gsteps);
  const hipError_t ec = hipStreamEndCapture(G->stream, &g);
  if (rc || ec != hipSuccess) {
    if (g) hipGraphDestroy(g);
    return hip_fail(rc ? (hipError_t)rc : ec, "graph capture");
  }
  hipGraphExec_t x = nullptr;
  const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (ei != hipSuccess) return hip_fail(ei, "hipGraphInstantiate");
  G->gx = x;
  G->gx_tune = e->tune;
  *out = x;
  return MLAU_OK;
}

}  // namespace mlau

using namespace mlau;

struct mlau_ensemble : Ensemble {};

#define ECHK(call)                                   \
  do {                                               \
    hipError_t _e = (call);                          \
    if (_e != hipSuccess) return hip_fail(_e, #call); \
  } while (0)

// Wait for launched steps and take their status read-back as the host
// mirror; slots added or removed while the steps ran keep the status the
// host gave them (written to the device after those steps).
static int sync_launch(Ensemble *e)
{
  if (!e->inflight) return MLAU_OK;
  e->inflight = false;
  ECHK(hipStreamSynchronize(e->stream));  // it waited for every group's steps
  if (e->timing.enabled) ens_timing_fold(e);
  for (EnsGroup *G : e->groups) {
    if (G->launched > 0) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, G->ev_start, G->ev_done) == hipSuccess && ms > 0.f) {
        const double per = ms / G->launched;
        G->step_ms = G->step_ms > 0 ? 0.5 * (G->step_ms + per) : per;
      }
      G->launched = 0;
    }
    memcpy(G->stat_host, G->stat_next, sizeof(int) * kEnsSlots);
    for (int s : G->touched) G->stat_host[s] = G->stat_src[s];
    G->touched.clear();
    G->nrun = 0;
    for (int s = 0; s < kEnsSlots; s++) G->nrun += (G->stat_host[s] == MLAU_MEMBER_RUNNING);
  }
  return MLAU_OK;
}

// Stream for copying a member's memory now: the copy stream when the member
// is not running (no launch reads or writes it), otherwise the compute
// stream once launched steps are complete.
static int quiet_stream(Ensemble *e, EnsGroup *G, int slot, hipStream_t *s)
{
  *s = e->xstream;
  if (G->stat_host[slot] != MLAU_MEMBER_RUNNING) return MLAU_OK;
  *s = e->stream;
  return sync_launch(e);
}

static int read_state(Ensemble *e, EnsGroup *G, int slot, EnsState *st)
{
  hipStream_t s;
  const int rc = quiet_stream(e, G, slot, &s);
  if (rc) return rc;
  ECHK(hipMemcpyAsync(st, G->p.st + slot, sizeof *st, hipMemcpyDeviceToHost, s));
  ECHK(hipStreamSynchronize(s));
  return MLAU_OK;
}

extern "C" {

int mlau_ens_create(long long budget_bytes, int flags, mlau_ensemble **out)
{
  if (!out) return fail(MLAU_ERR_ARG, "null output");
  *out = nullptr;
  if (budget_bytes <= 0) return fail(MLAU_ERR_ARG, "budget must be positive");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(MLAU_ERR_HIP, "no HIP device");
  mlau_ensemble *e = new mlau_ensemble();
  e->budget = budget_bytes;
  e->flags = flags;
  int dev = 0;
  hipGetDevice(&dev);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
    e->cus = (prop.multiProcessorCount + kXcds - 1) / kXcds * kXcds;
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&e->xstream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&e->ctl_ev, hipEventDisableTiming) != hipSuccess) {
    if (e->stream) hipStreamDestroy(e->stream);
    if (e->xstream) hipStreamDestroy(e->xstream);
    delete e;
    return fail(MLAU_ERR_HIP, "hipStreamCreate failed");
  }
  e->cur = e->stream;
  // launch shape at creation (mlau_ens_set_tune bits), for callers that cannot set it
  if (const char *t = std::getenv("MLAU_ENS_TUNE")) {
    const int v = std::atoi(t);
    if (v >= 0 && v <= 8191 && (v & 3) != 3) e->tune = v;
  }
  *out = e;
  return MLAU_OK;
}

void mlau_ens_destroy(mlau_ensemble *e)
{
  if (!e) return;
  if (e->stream) hipStreamSynchronize(e->stream);
  if (e->xstream) hipStreamSynchronize(e->xstream);
  for (EnsGroup *G : e->groups) hipStreamSynchronize(G->stream);
  graph_drop(e);
  for (EnsGroup *G : e->groups) group_free(G);
  for (auto &c : e->chunks)
    if (c.base) hipFree(c.base);
  for (auto &r : e->timing.pending) {
    hipEventDestroy(r.a);
    hipEventDestroy(r.b);
  }
  for (auto ev : e->timing.pool) hipEventDestroy(ev);
  if (e->ctl_ev) hipEventDestroy(e->ctl_ev);
  if (e->stream) hipStreamDestroy(e->stream);
  if (e->xstream) hipStreamDestroy(e->xstream);
  delete e;
}

// Grid with the unused axes of a 1D/2D member normalised (as mlau_create does).
static mlau_grid norm_grid(const mlau_grid &in)
{
  mlau_grid g = in;
  if (g.dim < 3) {
    g.nz = 1;
    g.zoff = 0;
    g.dz = 1.0;
  }
  if (g.dim < 2) {
    g.ny = 1;
    g.yoff = 0;
    g.dy = 1.0;
  }
  return g;
}

static size_t member_per(const mlau_grid &g)
{
  const size_t cells = (size_t)g.nx * g.ny * g.nz;
  return (cells * sizeof(double) + 255) & ~(size_t)255;
}

long long mlau_ens_member_bytes(const mlau_member *m)
{
  if (!m) return -1;
  const EnsDim *D = dim_table(m->grid.dim);
  if (!D) return -1;
  const mlau_grid g = norm_grid(m->grid);
  return (long long)(member_per(g) * D->member_arrays(g));
}

long long mlau_ens_used_bytes(const mlau_ensemble *e) { return e ? e->used : -1; }
long long mlau_ens_reserved_bytes(const mlau_ensemble *e) { return e ? (long long)e->reserved : -1; }

static int validate_member(const mlau_member *m)
{
  if (!m) return fail(MLAU_ERR_ARG, "null member");
  const mlau_grid &g = m->grid;
  const mlau_scheme &s = m->scheme;
  if (!dim_table(g.dim)) return fail(MLAU_ERR_ARG, "dim must be 1, 2 or 3");
  if (g.xoff != 4 || (g.dim >= 2 && g.yoff != 4) || (g.dim == 3 && g.zoff != 4))
    return fail(MLAU_ERR_ARG, "ghost width must be 4 on every axis of the member");
  if (g.nx < 2 * g.xoff + 5 || (g.dim >= 2 && g.ny < 2 * g.yoff + 5) || (g.dim == 3 && g.nz < 2 * g.zoff + 5))
    return fail(MLAU_ERR_ARG, "grid too small");
  if (!(g.dx > 0) || (g.dim >= 2 && !(g.dy > 0)) || (g.dim == 3 && !(g.dz > 0)) || !(g.gamma > 1))
    return fail(MLAU_ERR_ARG, "positive spacing and gamma > 1 required");
  if (s.odr < 1 || s.odr > 4) return fail(MLAU_ERR_ARG, "odr must be 1..4");
  if (s.rk < 1 || s.rk > 3) return fail(MLAU_ERR_ARG, "rk must be 1..3");
  if (s.flux < 0 || s.flux > 3) return fail(MLAU_ERR_ARG, "flux must be 0..3");
  if (s.ctw != 0 && s.ctw != 1) return fail(MLAU_ERR_ARG, "ctw must be 0 or 1");
  if (s.precision != MLAU_FP64 && s.precision != MLAU_MIXED && s.precision != MLAU_MIXED_RECON)
    return fail(MLAU_ERR_ARG, "precision must be 0, 1 or 2");
  if (g.dim == 1 && s.bc1d != MLAU_BC1D_OPEN && s.bc1d != MLAU_BC1D_PERIODIC)
    return fail(MLAU_ERR_ARG, "bc1d must be 0 or 1");
  if (g.dim >= 2) {
    const int *dn[3] = {s.dnxs, s.dnys, s.dnzs};
    for (int a = 0; a < g.dim; a++)
      for (int n = 0; n < 8; n++) {
        if (dn[a][n] < -1 || dn[a][n] > 1) return fail(MLAU_ERR_ARG, "boundary flags must be -1, 0 or 1");
        if ((dn[a][n] == 0) != (dn[a][0] == 0))
          return fail(MLAU_ERR_ARG, "an axis is periodic for all fields or for none");
      }
  }
  if (!(m->tend > 0) || !(m->cfl > 0) || !(m->vfloor >= 0) || m->max_steps < 0)
    return fail(MLAU_ERR_ARG, "tend > 0, cfl > 0, vfloor >= 0 and max_steps >= 0 required");
  return MLAU_OK;
}

int mlau_ens_add(mlau_ensemble *e, const mlau_member *m_in, const double *const *fields, int *id)
{
  if (!e || !fields || !id) return fail(MLAU_ERR_ARG, "null argument");
  int rc = validate_member(m_in);
  if (rc) return rc;
  mlau_member mm = *m_in;
  mm.grid = norm_grid(m_in->grid);
  const mlau_member *m = &mm;
  const EnsDim *D = dim_table(m->grid.dim);
  const size_t per = member_per(m->grid);
  const size_t bytes = per * D->member_arrays(m->grid);
  if (e->used + (long long)bytes > e->budget) return fail(MLAU_ERR_NOMEM, "member exceeds the ensemble budget");
  EnsGroup *G;
  if ((rc = group_get(e, *m, &G))) return rc;
  int slot = -1;
  for (int s = 0; s < kEnsSlots; s++)
    if (G->slot_id[s] < 0) {
      slot = s;
      break;
    }
  if (slot < 0) return fail(MLAU_ERR_NOMEM, "ensemble group is full");
  long long need[kEnsTypes];
  for (int k = 0; k < D->ntypes; k++) {
    need[k] = member_tiles_max(D, *m, k, G->variant);
    G->need[k] += need[k];
  }
  auto undo_need = [&]() {
    for (int k = 0; k < D->ntypes; k++) G->need[k] -= need[k];
  };
  if ((rc = group_reserve(e, G))) {
    undo_need();
    return rc;
  }
  size_t where;
  switch (arena_alloc(e, bytes, &where)) {
  case ARENA_OK: break;
  case ARENA_FULL:
    undo_need();
    return fail(MLAU_ERR_NOMEM, "member does not fit in the budget's free device memory (fragmented)");
  default:
    undo_need();
    return fail(MLAU_ERR_HIP, "device allocation for the member failed (hipMalloc within the budget)");
  }
  double *base = arena_ptr(e, where);
  char *rec = G->rec_host + (size_t)slot * D->rec_bytes;
  D->make_record(rec, *m, base, per / sizeof(double));
  double *u0[8];
  int nf;
  size_t cells;
  D->state_ptrs(rec, u0, &nf, &cells);
  EnsState st{};
  const double dz = m->grid.dz;
  memcpy(&st.vmax_bits, &m->vfloor, sizeof(double));
  st.tend = m->tend;
  st.cfl = m->cfl;
  st.vfloor = m->vfloor;
  st.dx = m->grid.dx;
  st.dy = m->grid.dy;
  st.dz = dz;
  st.dr = m->grid.dx;
  if (m->grid.dim >= 2) st.dr = std::min(st.dr, m->grid.dy);
  if (m->grid.dim == 3) st.dr = std::min(st.dr, dz);
  st.fixed = (m->grid.dim == 1);
  st.max_steps = m->max_steps;
  // The member's memory is fresh (or a removed member's, which no launch
  // touches), so its fields go up on the copy stream while launched steps of
  // the running members proceed; the host waits for the copy only, so the
  // caller's buffers are free on return.
  bool ok = hipMemsetAsync(base, 0, bytes, e->xstream) == hipSuccess;
  for (int n = 0; ok && n < nf; n++) {
    if (!fields[n]) ok = false;
    else ok = hipMemcpyAsync(u0[n], fields[n], cells * sizeof(double), hipMemcpyHostToDevice, e->xstream) == hipSuccess;
  }
  ok = hipStreamSynchronize(e->xstream) == hipSuccess && ok;
  // record, state and status join the plan at the next launch (compute stream order)
  G->st_src[slot] = st;
  G->stat_src[slot] = MLAU_MEMBER_RUNNING;
  ok = ok &&
       hipMemcpyAsync((char *)G->rec + (size_t)slot * D->rec_bytes, rec, D->rec_bytes, hipMemcpyHostToDevice,
                      e->stream) == hipSuccess &&
       hipMemcpyAsync(G->p.st + slot, G->st_src + slot, sizeof st, hipMemcpyHostToDevice, e->stream) == hipSuccess &&
       hipMemcpyAsync(G->p.stat + slot, G->stat_src + slot, sizeof(int), hipMemcpyHostToDevice, e->stream) ==
           hipSuccess;
  if (ok) ok = group_write_nb(e, G, slot, m) == MLAU_OK && group_flag_rewrite(e, G) == MLAU_OK;
  if (!ok) {
    hipGetLastError();
    arena_free(e, where, bytes);
    undo_need();
    return fail(MLAU_ERR_HIP, "member upload failed (null field or copy error)");
  }
  G->stat_host[slot] = MLAU_MEMBER_RUNNING;
  if (e->inflight) G->touched.push_back(slot);
  G->nrun++;
  G->slot_id[slot] = e->next_id;
  EnsMember em{};
  em.group = (int)(std::find(e->groups.begin(), e->groups.end(), G) - e->groups.begin());
  em.slot = slot;
  em.off = where;
  em.bytes = bytes;
  em.cells = (long long)(m->grid.nx - 2 * m->grid.xoff) * (m->grid.ny - 2 * m->grid.yoff) *
             (m->grid.nz - 2 * m->grid.zoff);
  em.m = *m;
  e->members[e->next_id] = em;
  e->used += (long long)bytes;
  *id = e->next_id++;
  return MLAU_OK;
}

static int member_of(mlau_ensemble *e, int id, EnsMember **m)
{
  auto it = e->members.find(id);
  if (it == e->members.end()) return fail(MLAU_ERR_ARG, "no such member");
  *m = &it->second;
  return MLAU_OK;
}

int mlau_ens_remove(mlau_ensemble *e, int id)
{
  if (!e) return fail(MLAU_ERR_ARG, "null ensemble");
  EnsMember *m;
  int rc = member_of(e, id, &m);
  if (rc) return rc;
  EnsGroup *G = e->groups[m->group];
  if (G->stat_host[m->slot] == MLAU_MEMBER_RUNNING) return fail(MLAU_ERR_STATE, "member is still running");
  EnsState st;
  if ((rc = read_state(e, G, m->slot, &st))) return rc;
  // the free status and zero tile counts reach the device between launches
  G->stat_src[m->slot] = kStatFree;
  ECHK(hipMemcpyAsync(G->p.stat + m->slot, G->stat_src + m->slot, sizeof(int), hipMemcpyHostToDevice, e->stream));
  if ((rc = group_write_nb(e, G, m->slot, nullptr))) return rc;
  for (int k = 0; k < G->D->ntypes; k++) G->need[k] -= member_tiles_max(G->D, m->m, k, G->variant);
  G->stat_host[m->slot] = kStatFree;
  if (e->inflight) G->touched.push_back(m->slot);
  G->slot_id[m->slot] = -1;
  e->zone_updates_removed += m->cells * st.nstep;
  arena_free(e, m->off, m->bytes);
  e->used -= (long long)m->bytes;
  e->members.erase(id);
  return MLAU_OK;
}

// Steps of group G in this launch: nsteps, or with MLAU_ENS_BALANCE nsteps
// times the ratio of the slowest running group's step time to G's (1..64),
// so that cheap groups advance while the expensive ones run.
static int group_steps(const Ensemble *e, const EnsGroup *G, int nsteps)
{
  if (!(e->flags & MLAU_ENS_BALANCE) || nsteps == 0 || !(G->step_ms > 0)) return nsteps;
  double tmax = 0.0;
  for (const EnsGroup *H : e->groups)
    if (H->nrun > 0) tmax = std::max(tmax, H->step_ms);
  const double r = tmax / G->step_ms;
  static const int caps[4] = {64, 256, 1024, 1024};
  const int cap = caps[(e->tune >> 11) & 3];  // tune bits 11-12
  return nsteps * (r >= cap ? cap : r >= 1.0 ? (int)r : 1);
}

int mlau_ens_launch(mlau_ensemble *e, int nsteps)
{
  if (!e || nsteps < 0) return fail(MLAU_ERR_ARG, "bad argument");
  if (e->inflight) return fail(MLAU_ERR_STATE, "steps already launched: mlau_ens_sync first");
  for (EnsGroup *G : e->groups)
    if (G->dirty) {
      int rc = group_flag_rewrite(e, G);
      if (rc) return rc;
    }
  ECHK(hipEventRecord(e->ctl_ev, e->stream));
  const bool graph = (e->flags & MLAU_ENS_GRAPH) && !e->timing.enabled;
  // tune bit 8: each group's steps start after the previous group's, so no
  // group's short kernels wait for CUs held by another group's stage kernel
  const bool serial = (e->tune >> 8) & 1;
  const EnsGroup *prev = nullptr;
  int rc = MLAU_OK;
  for (EnsGroup *G : e->groups) {
    G->launched = 0;
    if (G->nrun == 0) continue;
    int left = group_steps(e, G, nsteps);
    hipGraphExec_t x = nullptr;
    if (graph && (left >= e->gsteps || left == 0) && (rc = graph_get(e, G, &x))) break;  // zero steps: capture only
    if (left == 0) continue;
    G->launched = left;
    hipError_t he = hipStreamWaitEvent(G->stream, e->ctl_ev, 0);  // after the control writes
    if (he == hipSuccess && serial && prev) he = hipStreamWaitEvent(G->stream, prev->ev_done, 0);
    if (he == hipSuccess) he = hipEventRecord(G->ev_start, G->stream);
    while (he == hipSuccess && x && left >= e->gsteps) {
      he = hipGraphLaunch(x, G->stream);
      left -= e->gsteps;
    }
    if (he == hipSuccess && left > 0) he = (hipError_t)launch_steps(e, G, left);
    if (he == hipSuccess) he = hipEventRecord(G->ev_done, G->stream);
    if (he == hipSuccess) he = hipStreamWaitEvent(e->stream, G->ev_done, 0);  // control waits for the steps
    if (he != hipSuccess) {
      rc = hip_fail(he, "ensemble step launch");
      break;
    }
    prev = G;
  }
  for (EnsGroup *G : e->groups)
    if (!rc && hipMemcpyAsync(G->stat_next, G->p.stat, sizeof(int) * kEnsSlots, hipMemcpyDeviceToHost,
                              e->stream) != hipSuccess)
      rc = fail(MLAU_ERR_HIP, "status read-back failed");
  if (rc) {
    for (EnsGroup *G : e->groups) {
      hipStreamSynchronize(G->stream);
      G->launched = 0;
    }
    hipStreamSynchronize(e->stream);
    return rc;
  }
  e->inflight = true;
  return MLAU_OK;
}

int mlau_ens_sync(mlau_ensemble *e, int *nrunning)
{
  if (!e) return fail(MLAU_ERR_ARG, "null ensemble");
  int rc = sync_launch(e);
  if (rc) return rc;
  int nrun = 0;
  for (EnsGroup *G : e->groups) nrun += G->nrun;
  if (nrunning) *nrunning = nrun;
  return MLAU_OK;
}

int mlau_ens_ready(mlau_ensemble *e)
{
  if (!e) return -fail(MLAU_ERR_ARG, "null ensemble");
  if (!e->inflight) return 1;
  const hipError_t q = hipStreamQuery(e->stream);  // the control stream waits for every group
  if (q == hipSuccess) return 1;
  if (q == hipErrorNotReady) return 0;
  return -hip_fail(q, "hipStreamQuery");
}

int mlau_ens_on_ready(mlau_ensemble *e, void (*fn)(void *), void *arg)
{
  if (!e || !fn) return fail(MLAU_ERR_ARG, "null argument");
  if (!e->inflight) {
    fn(arg);
    return MLAU_OK;
  }
  ECHK(hipLaunchHostFunc(e->stream, fn, arg));  // runs once the control stream has waited for every group
  return MLAU_OK;
}

int mlau_ens_run(mlau_ensemble *e, int nsteps, int *nrunning)
{
  const int rc = mlau_ens_launch(e, nsteps);
  return rc ? rc : mlau_ens_sync(e, nrunning);
}

int mlau_ens_status(mlau_ensemble *e, int id, int *status, double *tim, double *dt, long long *nstep)
{
  if (!e) return fail(MLAU_ERR_ARG, "null ensemble");
  EnsMember *m;
  int rc = member_of(e, id, &m);
  if (rc) return rc;
  EnsGroup *G = e->groups[m->group];
  EnsState st;
  if ((rc = read_state(e, G, m->slot, &st))) return rc;
  if (status) *status = G->stat_host[m->slot];
  if (tim) *tim = st.tim;
  if (dt) *dt = st.dt;
  if (nstep) *nstep = st.nstep;
  return MLAU_OK;
}

int mlau_ens_finished(mlau_ensemble *e, int *ids, int cap, int *n)
{
  if (!e || !n || (cap > 0 && !ids)) return fail(MLAU_ERR_ARG, "bad argument");
  int k = 0;
  for (auto &kv : e->members) {
    const EnsGroup *G = e->groups[kv.second.group];
    if (G->stat_host[kv.second.slot] != MLAU_MEMBER_RUNNING) {
      if (k < cap) ids[k] = kv.first;
      k++;
    }
  }
  *n = k;
  return MLAU_OK;
}

int mlau_ens_download(mlau_ensemble *e, int id, double *const *fields)
{
  if (!e || !fields) return fail(MLAU_ERR_ARG, "null argument");
  EnsMember *m;
  int rc = member_of(e, id, &m);
  if (rc) return rc;
  EnsGroup *G = e->groups[m->group];
  double *u0[8];
  int nf;
  size_t cells;
  G->D->state_ptrs(G->rec_host + (size_t)m->slot * G->D->rec_bytes, u0, &nf, &cells);
  for (int n = 0; n < nf; n++)
    if (!fields[n]) return fail(MLAU_ERR_ARG, "null field pointer");
  hipStream_t s;
  if ((rc = quiet_stream(e, G, m->slot, &s))) return rc;
  for (int n = 0; n < nf; n++) ECHK(hipMemcpyAsync(fields[n], u0[n], cells * sizeof(double), hipMemcpyDeviceToHost, s));
  ECHK(hipStreamSynchronize(s));
  return MLAU_OK;
}

int mlau_ens_set_tune(mlau_ensemble *e, int tune)
{
  if (!e || tune < 0 || tune > 8191 || (tune & 3) == 3) return fail(MLAU_ERR_ARG, "bad tune");
  if (tune == e->tune) return MLAU_OK;
  int src = sync_launch(e);
  if (src) return src;
  e->tune = tune;
  for (EnsGroup *G : e->groups) {
    for (auto &kv : e->members) {
      if (e->groups[kv.second.group] != G) continue;
      int rc = group_write_nb(e, G, kv.second.slot, &kv.second.m);
      if (rc) return rc;
    }
    int rc = group_flag_rewrite(e, G);
    if (rc) return rc;
  }
  return MLAU_OK;
}

int mlau_ens_autotune(mlau_ensemble *e, int steps, int *tune)
{
  if (!e || steps < 1) return fail(MLAU_ERR_ARG, "bad argument");
  if (int src = sync_launch(e)) return src;
  // snapshot: member states, dynamic state and status of every group
  std::vector<double *> snap;
  std::vector<std::pair<EnsMember *, size_t>> sizes;
  struct GSnap { EnsState *st; int *stat; };
  std::vector<GSnap> gs;
  auto release = [&]() {
    for (double *p : snap) hipFree(p);
    for (auto &g : gs) {
      hipFree(g.st);
      hipFree(g.stat);
    }
  };
  auto copy_all = [&](bool save) -> int {
    size_t i = 0;
    for (auto &kv : e->members) {
      EnsGroup *G = e->groups[kv.second.group];
      double *u0[8];
      int nf;
      size_t cells;
      G->D->state_ptrs(G->rec_host + (size_t)kv.second.slot * G->D->rec_bytes, u0, &nf, &cells);
      for (int n = 0; n < nf; n++, i++) {
        if (save) {
          double *p = nullptr;
          if (hipMalloc(&p, cells * sizeof(double)) != hipSuccess) return fail(MLAU_ERR_NOMEM, "autotune snapshot");
          snap.push_back(p);
        }
        double *dst = save ? snap[i] : u0[n];
        const double *src = save ? u0[n] : snap[i];
        ECHK(hipMemcpyAsync(dst, src, cells * sizeof(double), hipMemcpyDeviceToDevice, e->stream));
      }
    }
    for (size_t k = 0; k < e->groups.size(); k++) {
      EnsGroup *G = e->groups[k];
      if (save) {
        GSnap g{};
        if (hipMalloc(&g.st, sizeof(EnsState) * kEnsSlots) != hipSuccess ||
            hipMalloc(&g.stat, sizeof(int) * kEnsSlots) != hipSuccess)
          return fail(MLAU_ERR_NOMEM, "autotune snapshot");
        gs.push_back(g);
      }
      ECHK(hipMemcpyAsync(save ? gs[k].st : G->p.st, save ? G->p.st : gs[k].st, sizeof(EnsState) * kEnsSlots,
                          hipMemcpyDeviceToDevice, e->stream));
      ECHK(hipMemcpyAsync(save ? gs[k].stat : G->p.stat, save ? G->p.stat : gs[k].stat, sizeof(int) * kEnsSlots,
                          hipMemcpyDeviceToDevice, e->stream));
      if (!save) G->dirty = 1;
    }
    ECHK(hipStreamSynchronize(e->stream));
    return MLAU_OK;
  };
  int rc = copy_all(true);
  if (rc) {
    release();
    return rc;
  }
  hipEvent_t a, b;
  ECHK(hipEventCreate(&a));
  ECHK(hipEventCreate(&b));
  auto time_tune = [&](int t, float *ms) -> int {
    int r = mlau_ens_set_tune(e, t);
    if (!r) r = mlau_ens_run(e, 8, nullptr);  // warm-up (and graph build)
    if (!r) r = copy_all(false);
    if (!r) r = mlau_ens_run(e, 0, nullptr);
    if (r) return r;
    hipEventRecord(a, e->stream);
    r = mlau_ens_run(e, steps, nullptr);
    hipEventRecord(b, e->stream);
    hipEventSynchronize(b);
    hipEventElapsedTime(ms, a, b);
    if (!r) r = copy_all(false);
    return r;
  };
  // coordinate search: blocks per CU, then y-chunk, then sweep order, then
  // concurrent or consecutive groups
  int best;
  float best_ms = 1e30f, ms;
  const int mults[4] = {1, 2, 4, 8};
  int bm = 2;
  for (int mi = 0; mi < 4 && !rc; mi++) {
    const int t = (mults[mi] << 4);
    if ((rc = time_tune(t, &ms))) break;
    if (ms < best_ms) { best_ms = ms; bm = mults[mi]; }
  }
  best = bm << 4;
  for (int ch = 1; ch <= 2 && !rc; ch++) {
    const int t = (bm << 4) | ch;
    if ((rc = time_tune(t, &ms))) break;
    if (ms < best_ms) { best_ms = ms; best = t; }
  }
  if (!rc) {
    const int t = best | 4;
    if (!(rc = time_tune(t, &ms)) && ms < best_ms) {
      best_ms = ms;
      best = t;
    }
  }
  if (!rc && e->groups.size() > 1) {  // groups one after another
    const int t = best | 256;
    if (!(rc = time_tune(t, &ms)) && ms < best_ms) {
      best_ms = ms;
      best = t;
    }
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  if (!rc) rc = mlau_ens_set_tune(e, best);
  if (!rc) rc = copy_all(false);
  release();
  if (rc) return rc;
  for (EnsGroup *G : e->groups)
    ECHK(hipMemcpyAsync(G->stat_host, G->p.stat, sizeof(int) * kEnsSlots, hipMemcpyDeviceToHost, e->stream));
  ECHK(hipStreamSynchronize(e->stream));
  for (EnsGroup *G : e->groups) {
    G->nrun = 0;
    for (int s = 0; s < kEnsSlots; s++) G->nrun += (G->stat_host[s] == MLAU_MEMBER_RUNNING);
  }
  if (tune) *tune = best;
  return MLAU_OK;
}

int mlau_ens_kernel_timing(mlau_ensemble *e, int enable)
{
  if (!e) return fail(MLAU_ERR_ARG, "null ensemble");
  e->timing.enabled = enable != 0;
  return MLAU_OK;
}

int mlau_ens_kernel_stats(mlau_ensemble *e, int kernel, double *total_ms, long long *launches, const char **name)
{
  if (!e || kernel < 0 || kernel >= E_COUNT) return fail(MLAU_ERR_ARG, "bad argument");
  if (int rc = sync_launch(e)) return rc;
  ens_timing_fold(e);
  if (total_ms) *total_ms = e->timing.total_ms[kernel];
  if (launches) *launches = e->timing.count[kernel];
  if (name) *name = kEnsKernelNames[kernel];
  return MLAU_OK;
}

int mlau_ens_num_kernels(void) { return E_COUNT; }

long long mlau_ens_zone_updates(mlau_ensemble *e)
{
  if (!e || sync_launch(e)) return -1;
  long long z = e->zone_updates_removed;
  std::vector<EnsState> st(kEnsSlots);
  for (size_t k = 0; k < e->groups.size(); k++) {
    if (hipMemcpy(st.data(), e->groups[k]->p.st, sizeof(EnsState) * kEnsSlots, hipMemcpyDeviceToHost) != hipSuccess)
      return -1;
    for (auto &kv : e->members)
      if (kv.second.group == (int)k) z += kv.second.cells * st[kv.second.slot].nstep;
  }
  return z;
}

}  // extern "C"
