// mlau_common.h -- host/device shared declarations of the HIP path.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/mlau.h"
#include "mlau_dev.h"

// Register budget of the marching sweeps: at most 256 VGPRs + AGPRs per lane,
// so two waves share each SIMD (one wave alone leaves the SIMD idle while it
// waits on the next row's loads).
#define MLAU_TWO_WAVES __attribute__((amdgpu_waves_per_eu(2)))

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
constexpr unsigned kXcds = 8;  // MI355X: 8 accelerator dies, each with its own L2

// Workgroups are handed to the XCDs round-robin (block b -> XCD b % 8).  For
// stencils that reuse neighbouring rows, remap the linear block index so each
// XCD walks one contiguous band of the grid and finds its neighbours in its own
// L2.  A bijection on [0, nb).
__device__ __forceinline__ unsigned xcd_band(unsigned b, unsigned nb)
{
  const unsigned x = b % kXcds, k = b / kXcds;
  const unsigned q = nb / kXcds, r = nb % kXcds;
  return x * q + (x < r ? x : r) + k;
}

__device__ __forceinline__ double shfl_up1(double v)
{
  return __shfl_up(v, 1, kWave);
}

// Block-wide max of a non-negative speed (NaN never wins: comparisons with >
// as in the reference's "if (vtmp > vmax) vmax=vtmp"), then one atomic per
// block on the bit pattern (non-negative doubles order like their bits).
template <int BLOCK>
__device__ __forceinline__ void block_max_atomic(double v, unsigned long long *acc)
{
  __shared__ double red[BLOCK / 64];
  for (int o = 32; o > 0; o >>= 1) {
    const double t = __shfl_xor(v, o, 64);
    if (t > v) v = t;
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < BLOCK / 64; k++)
      if (red[k] > v) v = red[k];
    if (v > 0.0) atomicMax(acc, (unsigned long long)__double_as_longlong(v));
  }
}

// Wavefronts of the row-flattened x sweeps: cells 2..nx-3 of `rows` rows at a
// stride of 63 cells per 64-lane wavefront.
__host__ __device__ __forceinline__ long xsweep_waves(long rows, int nx) { return (rows * (nx - 4) - 1 + 62) / 63; }

}  // namespace mlau
