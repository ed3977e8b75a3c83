// hbm_streams_probe.hip -- HBM bandwidth against the number of arrays a
// kernel streams at once.  Each launch reads R arrays and writes W arrays of
// 2048 x 2048 doubles, one coalesced load (store) per array and thread, the
// access shape of the 2D cell update with (R, W) = (37, 8) but without its
// stencil reuse.  Prints the achieved rate per (R, W): the practical ceiling a
// memory-bound kernel of that many streams is compared with.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

struct Ptrs {
  const double *in[40];
  double *out[8];
};

template <int R, int W>
__global__ __launch_bounds__(256) void k_streams(Ptrs p, long n)
{
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
#pragma unroll
  for (int r = 0; r < R; r++) s += p.in[r][i];
#pragma unroll
  for (int w = 0; w < W; w++) p.out[w][i] = s + w;
}

template <int R, int W>
static bool run(const Ptrs &p, long n, hipEvent_t a, hipEvent_t b)
{
  const dim3 g((unsigned)((n + 255) / 256));
  for (int k = 0; k < 3; k++) hipLaunchKernelGGL((k_streams<R, W>), g, dim3(256), 0, 0, p, n);
  const int reps = 10;
  if (hipEventRecord(a) != hipSuccess) return false;
  for (int k = 0; k < reps; k++) hipLaunchKernelGGL((k_streams<R, W>), g, dim3(256), 0, 0, p, n);
  if (hipEventRecord(b) != hipSuccess || hipEventSynchronize(b) != hipSuccess) return false;
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return false;
  const double us = ms * 1e3 / reps;
  printf("R=%2d W=%d  %8.1f us  %.2f TB/s\n", R, W, us, 8.0 * (R + W) * n / (us * 1e-6) / 1e12);
  fflush(stdout);
  return true;
}

int main()
{
  const long n = 2048L * 2048;
  std::vector<double *> bufs(48, nullptr);
  for (auto &q : bufs)
    if (hipMalloc(&q, n * sizeof(double)) != hipSuccess || hipMemset(q, 0, n * sizeof(double)) != hipSuccess) return 1;
  Ptrs p;
  for (int k = 0; k < 40; k++) p.in[k] = bufs[k];
  for (int k = 0; k < 8; k++) p.out[k] = bufs[40 + k];
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return 1;
  const bool ok = run<1, 1>(p, n, a, b) && run<8, 1>(p, n, a, b) && run<8, 8>(p, n, a, b) &&
                  run<16, 8>(p, n, a, b) && run<29, 8>(p, n, a, b) && run<37, 8>(p, n, a, b) &&
                  run<37, 1>(p, n, a, b);
  for (auto q : bufs) (void)hipFree(q);
  return ok ? 0 : 1;
}
