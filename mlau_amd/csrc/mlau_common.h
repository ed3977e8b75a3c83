// mlau_common.h -- host/device shared declarations of the HIP path.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/mlau.h"
#include "mlau_dev.h"

namespace mlau {

// Per-step scalars kept on the device so a step loop never waits on the host:
// dt and the ratios the reference forms once per call (dtdx=dt/dx, ...).
struct StepInfo {
  double dt, dtdx, dtdy, dtdz;
  double vmax;                    // last CFL speed (reference cflcheck vmax)
  unsigned long long vmax_bits;   // reduction accumulator (non-negative doubles order like their bits)
  double tim;                     // accumulated time (driver's tim+=dt)
  long long nstep;                // steps taken
};

// Ghost-cell source of one axis under the reference's boundary rules
// (2D/SERIAL/OTvortex/boundary.c:18-74, 3D/MPI/OTvortex/boundary.c:91-309):
// periodic (dn==0) copies the image across the domain; Neumann/Dirichlet
// (dn=+1/-1) mirrors about the boundary (stagger st shifts the mirror and
// leaves one more face on the right to the solver) and multiplies by dn.
// Returns the source index; multiplies *sg by the factor.  Interior indices map
// to themselves.
__host__ __device__ __forceinline__ int bc_src(int i, int n, int off, int st, int dn, double *sg)
{
  if (dn == 0) {
    if (i < off) return n - 2 * off + i;
    if (i >= n - off) return i - (n - 2 * off);
    return i;
  }
  if (i < off) {
    *sg *= (double)dn;
    return 2 * off - 1 + st - i;
  }
  if (i >= n - (off - st)) {
    *sg *= (double)dn;
    return n - 2 * off + st + (n - 1 - i);
  }
  return i;
}

constexpr int kWave = 64;

__device__ __forceinline__ double shfl_up1(double v)
{
  return __shfl_up(v, 1, kWave);
}

}  // namespace mlau
