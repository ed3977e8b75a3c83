// mlau_noalloc.hip -- one 2D step on caller-owned device arrays with no
// device allocation: the stage state, primitives, CT weights, shock sensors,
// face fluxes, corner E-field and the step scalars live in one file-scope
// device array of fixed size (kNoAllocBytes per GPU, held from the module's
// load to the end of the process).  For callers that cannot allocate: code
// under HIP graph capture, tight serving loops, harnesses that forbid it.
//
// Every call on a GPU uses that GPU's one array, so calls that overlap on one
// GPU (two host threads, or two streams without an order between them) race;
// calls on one stream are ordered and safe.  The used part of the array is
// cleared by a fill kernel at every call (no hipMemsetAsync, so the call may
// be captured into a graph).
#include <hip/hip_runtime.h>

#include <mutex>

#include "mlau_internal.h"

namespace mlau {

constexpr size_t kNoAllocBytes = (size_t)128 << 20;
constexpr int kNoAllocArrays = 35;  // U1 (8), prims (7), ctx/dvx/dvy (3), fx+fcx (8), fy+fcy (8), ez (1)
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kNoAllocBytes / sizeof(double)];

// Device address of `scratch` on each GPU (each holds its own copy at its own
// address), looked up once per device with that device current.
static double *g_ptr[kMaxDevices];
static std::mutex g_mu;

static int scratch_ptr(int dev, double **out)
{
  if (dev < 0 || dev >= kMaxDevices) return fail(MLAU_ERR_ARG, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_ptr[dev] = (double *)p;
  }
  *out = g_ptr[dev];
  return MLAU_OK;
}

__global__ __launch_bounds__(256) void k_fill_zero(double *p, long n)
{
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0.0;
}

}  // namespace mlau

using namespace mlau;

extern "C" {

long long mlau_noalloc_max_cells(void)
{
  const size_t per_cell = (size_t)kNoAllocArrays * sizeof(double);
  return (long long)((kNoAllocBytes - 4096 - (size_t)kNoAllocArrays * 256) / per_cell);
}

int mlau_fd_ct_2d_noalloc(double *ro, double *mx, double *my, double *mz, double *en, double *bx, double *by,
                          double *bz, double dt, double dx, double dy, int nx, int ny, int xoff, int yoff,
                          double gamma, const mlau_scheme *scheme, void *stream)
{
  if (!scheme) return fail(MLAU_ERR_ARG, "null scheme");
  double *f[8] = {ro, mx, my, mz, en, bx, by, bz};
  for (int n = 0; n < 8; n++)
    if (!f[n]) return fail(MLAU_ERR_ARG, "null field pointer");
  Handle h;  // on the stack: pointers into the caller's arrays and the scratch array only
  h.g = mlau_grid{};
  h.g.dim = 2; h.g.nx = nx; h.g.ny = ny; h.g.nz = 1; h.g.xoff = xoff; h.g.yoff = yoff; h.g.zoff = 0;
  h.g.dx = dx; h.g.dy = dy; h.g.dz = 1.0; h.g.gamma = gamma;
  h.s = *scheme;
  int rc = mlau_validate(&h.g, &h.s);
  if (rc) return rc;
  const long long cells = (long long)nx * ny;
  if (cells > mlau_noalloc_max_cells())
    return fail(MLAU_ERR_ARG, "grid of " + std::to_string(cells) + " cells exceeds the no-allocation scratch (" +
                                  std::to_string(mlau_noalloc_max_cells()) + " cells, " +
                                  std::to_string(kNoAllocBytes >> 20) + " MiB)");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(MLAU_ERR_HIP, "hipGetDevice failed");
  double *ptr = nullptr;
  if ((rc = scratch_ptr(dev, &ptr))) return rc;
  h.device = dev;
  h.stream = (hipStream_t)stream;
  h.nf = 8;
  h.ncell = cells;
  for (int n = 0; n < 8; n++) {
    h.bc.dnx[n] = scheme->dnxs[n];
    h.bc.dny[n] = scheme->dnys[n];
    h.bc.dnz[n] = 0;
    h.bc.stx[n] = (n == 5);
    h.bc.sty[n] = (n == 6);
    h.bc.stz[n] = (n == 7);
  }
  h.periodic_images = nx - 2 * xoff >= 2 * xoff && ny - 2 * yoff >= 2 * yoff;
  for (int n = 0; n < 8; n++)
    if (h.bc.dnx[n] || h.bc.dny[n]) h.periodic_images = false;
  // carve the scratch like mlau_create carves its workspace (2D arrays only)
  const size_t per = ((size_t)cells * sizeof(double) + 255) & ~(size_t)255;
  char *p = (char *)ptr;
  auto take = [&]() {
    double *r = (double *)p;
    p += per;
    return r;
  };
  for (int n = 0; n < 8; n++) h.U0[n] = f[n];
  for (int n = 0; n < 8; n++) h.U1[n] = take();
  for (int n = 0; n < 7; n++) h.prim[n] = take();
  h.aux[0] = take();
  h.aux[3] = take();
  h.aux[4] = take();
  for (int n = 0; n < 7; n++) h.fx[n] = take();
  h.fcx[0] = take();
  for (int n = 0; n < 7; n++) h.fy[n] = take();
  h.fcy[0] = take();
  h.e[2] = take();
  h.info = (StepInfo *)p;
  p += 4096;
  const long used = (long)((p - (char *)ptr) / sizeof(double));
  hipLaunchKernelGGL(k_fill_zero, dim3(1024), dim3(256), 0, h.stream, ptr, used);
  hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, h.stream, h.info, dt, dx, dy, 1.0);
  const int e = run_steps_2d(&h, 1, false, 0.0, 0.0);
  if (e) return hip_fail((hipError_t)e, "mlau_fd_ct_2d_noalloc launch");
  const hipError_t le = hipGetLastError();
  if (le != hipSuccess) return hip_fail(le, "mlau_fd_ct_2d_noalloc launch");
  return MLAU_OK;
}

}  // extern "C"
