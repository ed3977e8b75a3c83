// mlau_ens.h -- internal layout of the batched ensemble engine.
//
// Members with the same (dim, odr, rk, ctw, precision, flux) and kernel variant
// (2D: fused stage kernel or separate stage kernels) form a group; a group
// owns a device table of member records (array pointers, grid, flux), the
// members' dynamic state (time, dt, status) and one work list of tiles per
// kernel type.  Every step, a one-block planning kernel compacts the running
// members, prefix-sums their tile counts and, when the running set changed,
// rewrites the work lists; the stage kernels are persistent (a fixed grid
// that loops over the device-resident list), so a captured HIP graph stays
// valid whatever the number of running members.
#pragma once
#include <map>
#include <string>
#include <vector>

#include "../../include/mlau_ensemble.h"
#include "mlau_internal.h"

namespace mlau {

constexpr int kEnsSlots = 4096;   // members per group
constexpr int kEnsTypes = 8;      // work-list kinds per group (at most)
constexpr int kPlanThreads = 1024;

// Dynamic per-member state, device resident.
struct EnsState {
  double dt, dtdx, dtdy, dtdz;
  unsigned long long vmax_bits;  // max of the cells' speeds (NaN speeds skipped; +Inf for a non-finite state)
  double tim, tend, cfl, vfloor, dr, dx, dy, dz;
  long long nstep, max_steps;
  double dtc;                    // CFL step of the current (or, when fixed, the first) step
  int last, fixed;               // fixed: dt from the initial state only (the 1D driver's rule)
  int halt;                      // set with MLAU_MEMBER_FAILED: the step's updates leave the state alone
};

// Kernel argument of the planning kernels (plain device pointers).
struct EnsPlan {
  EnsState *st;
  int *stat;                  // [kEnsSlots] status mirror for host polling (MLAU_MEMBER_*, 4 = free slot)
  int *act;                   // [kEnsSlots] running flag of the last plan
  int *nb;                    // [kEnsTypes][kEnsSlots] tiles of each member per kind (host written)
  int *off;                   // [kEnsTypes][kEnsSlots] first tile of each running member per kind
  int *count;                 // [kEnsTypes] tiles per kind in the lists; [kEnsTypes] running members;
                              // [kEnsTypes+1] rewrite flag (host sets it to force a rewrite)
  int2 *tiles[kEnsTypes];     // work lists: (slot, tile within member)
};

// Per-dimension launch sequences (mlau_ens2d.hip, ...).
struct EnsGroup;
struct Ensemble;
struct EnsDim {
  size_t rec_bytes;                                              // member record size
  int ntypes;                                                    // work-list kinds
  long long (*tiles)(const mlau_member &m, int kind, int tune, int variant);  // tiles of a member per kind
  size_t (*member_arrays)(const mlau_grid &g);                   // double arrays per member
  void (*make_record)(void *rec, const mlau_member &m, double *base, size_t per);
  int (*launch_step)(Ensemble *e, EnsGroup *g);                  // one full step of the group
  void (*state_ptrs)(const void *rec, double **u0, int *nf, size_t *cells);  // resident state
  int (*variant)(const mlau_member &m, int flags);               // kernel variant (part of the group key)
};
const EnsDim *ens_dim1d();
const EnsDim *ens_dim2d();
const EnsDim *ens_dim3d();

// Host-to-device writes of records, states, statuses and tile counts are
// queued on the compute stream from pinned host arrays that stay valid, so
// they take effect between launches without the host waiting for a launch in
// flight.
struct EnsGroup {
  int dim, odr, rk, ctw, precision, flux, variant;
  const EnsDim *D;
  EnsPlan p{};
  void *rec = nullptr;                      // device member records [kEnsSlots]
  char *rec_host = nullptr;                 // pinned host copies of the records
  EnsState *st_src = nullptr;               // pinned: initial state of each added member
  int *stat_src = nullptr;                  // pinned: status written at add/remove
  int *nb_src = nullptr;                    // pinned: [kEnsTypes][kEnsSlots] tile counts
  int *flag_src = nullptr;                  // pinned: constant 1 (work-list rewrite request)
  long long cap[kEnsTypes] = {};            // work-list capacities
  long long need[kEnsTypes] = {};           // sum of member tiles over occupied slots
  std::vector<int> slot_id;                 // member id per slot (-1: free)
  int dirty = 1;                            // force a work-list rewrite on the next plan
  int *stat_host = nullptr;                 // mirror of p.stat as of the last sync
  int *stat_next = nullptr;                 // pinned: p.stat copied back by the launch in flight
  std::vector<int> touched;                 // slots added or removed while steps were in flight
  int nrun = 0;                             // running members as of the mirror
  // the group's steps run on its own stream, beside the other groups' steps
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_done = nullptr;
  hipGraphExec_t gx = nullptr;              // graph of Ensemble::gsteps steps (captured on first use)
  int gx_tune = -1;                         // tuning the graph was captured with
  int launched = 0;                         // steps in the launch in flight
  double step_ms = 0.0;                     // measured time per step (0: not yet measured)
};

struct EnsMember {
  int group, slot;
  size_t off, bytes;        // arena placement
  long long cells;          // interior cells
  mlau_member m;            // description (grid normalised)
};

struct Ensemble {
  long long budget = 0, used = 0;
  int flags = 0;
  int tune = 0;
  int cus = 256;
  hipStream_t stream = nullptr;   // control: writes that must fall between launches, status read-back
  hipStream_t xstream = nullptr;  // bulk copies of members that no launch touches (not running)
  hipStream_t cur = nullptr;      // stream the step kernels being issued go to (the group's)
  hipEvent_t ctl_ev = nullptr;    // end of the control writes before a launch
  bool inflight = false;          // steps launched and not yet synchronised
  // arena: chunks of device memory, first-fit free list per chunk
  struct Chunk { char *base = nullptr; size_t size = 0; std::map<size_t, size_t> free; };
  std::vector<Chunk> chunks;      // released chunks keep their slot (base == nullptr)
  size_t reserved = 0;            // bytes of all chunks: never above the budget
  std::vector<EnsGroup *> groups;
  std::map<int, EnsMember> members;
  int next_id = 0;
  long long zone_updates_removed = 0;
  int gsteps = 8;            // steps per captured graph
  Timing timing;
};

// CUs the persistent stage kernels are sized for: all of them, or (tune bits
// 9-10) all but 8, 16 or 32, which then stay free for other groups' short
// kernels (plan, CFL, dt, and a 1D group's launches) while a stage kernel of
// one group holds the rest.
inline int ens_cus(const Ensemble *e)
{
  static const int reserve[4] = {0, 8, 16, 32};
  const int c = e->cus - reserve[(e->tune >> 9) & 3];
  return c < 8 ? 8 : c;
}

// Kernel ids of the ensemble timing table.
enum EnsKernelId {
  E_PLAN = 0, E_FILL, E_CFL, E_DT, E_PRIM, E_AUX, E_XSWEEP, E_YSWEEP, E_EZ, E_UPD, E_BC, E_ADVANCE, E_STAGE,
  E_COUNT
};

void ens_timing_begin(Ensemble *e, hipEvent_t *a);
void ens_timing_end(Ensemble *e, int k, hipEvent_t a);

#define ENS_LAUNCH(e, kid, kernel, grid, block, ...)                                  \
  do {                                                                                \
    hipEvent_t _ev = nullptr;                                                         \
    if ((e)->timing.enabled) ens_timing_begin((e), &_ev);                             \
    hipLaunchKernelGGL(kernel, grid, block, 0, (e)->cur, __VA_ARGS__);                \
    if ((e)->timing.enabled) ens_timing_end((e), (kid), _ev);                         \
  } while (0)

// Shared kernels (mlau_ens.hip).
__global__ void k_ens_plan(EnsPlan p, int ntypes);
__global__ void k_ens_fill(EnsPlan p, int ntypes);
__global__ void k_ens_dt(EnsPlan p);
__global__ void k_ens_advance(EnsPlan p);
constexpr int kStatFree = 4;

// Persistent-loop range: the tile list is cut into kXcds contiguous bands and
// band x is worked by the blocks that the dispatcher places on XCD x
// (blockIdx % kXcds), so neighbouring tiles of a member share an L2.  Each
// tile goes to exactly one block for any grid: a grid of fewer than kXcds
// blocks strides over the whole list, and the blocks past the last full set
// of kXcds take none.
struct PRange {
  int t, end, stride;
};
__device__ __forceinline__ PRange pers_range(int total)
{
  const int per = gridDim.x / kXcds;
  if (per == 0) return PRange{(int)blockIdx.x, total, (int)gridDim.x};
  const int x = blockIdx.x % kXcds, k = blockIdx.x / kXcds;
  if (k >= per) return PRange{0, 0, 1};
  const int q = total / kXcds, r = total % kXcds;
  const int lo = x * q + (x < r ? x : r);
  const int hi = lo + q + (x < r ? 1 : 0);
  return PRange{lo + k, hi, per};
}

// Max of the block's per-lane CFL speeds with the reference's
// "if (vtmp > vmax) vmax=vtmp" (2D/SERIAL/OTvortex/cflcheck.hpp:21-23): a NaN
// speed never wins, +Inf does; then one atomicMax on the member's bits
// (non-negative doubles order like their bit patterns).  256-thread blocks;
// red holds 4 doubles.  Every thread of the block must call it.
__device__ __forceinline__ void ens_vmax_block(double v, unsigned long long *acc, double *red)
{
  for (int o = 32; o > 0; o >>= 1) {
    const double q = __shfl_xor(v, o, 64);
    if (q > v) v = q;
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; k++)
      if (red[k] > v) v = red[k];
    if (v > 0.0) atomicMax(acc, (unsigned long long)__double_as_longlong(v));
  }
  __syncthreads();
}

}  // namespace mlau
