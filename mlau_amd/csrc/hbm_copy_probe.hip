// hbm_copy_probe.hip -- calibration kernel for HBM traffic counters.
// Copies N doubles (reads 8N bytes, writes 8N bytes) so that the FETCH_SIZE /
// WRITE_SIZE readings of rocprofv3 for a kernel of known traffic can scale the
// readings taken on the solver kernels.  Also prints the achieved copy rate.
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void hbm_copy(const double *__restrict__ a, double *__restrict__ b, long n)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

int main()
{
  const long n = 1L << 26;  // 512 MiB per array
  double *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * 8) != hipSuccess || hipMalloc(&b, n * 8) != hipSuccess) return 1;
  if (hipMemset(a, 0, n * 8) != hipSuccess) return 1;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
  for (int r = 0; r < 3; r++) hipLaunchKernelGGL(hbm_copy, dim3(n / 256), dim3(256), 0, 0, a, b, n);
  (void)hipEventRecord(e0);
  const int reps = 10;
  for (int r = 0; r < reps; r++) hipLaunchKernelGGL(hbm_copy, dim3(n / 256), dim3(256), 0, 0, a, b, n);
  (void)hipEventRecord(e1);
  (void)hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  std::printf("hbm_copy: %ld doubles (%.1f MiB read + %.1f MiB written) per launch, %.3f ms/launch, %.1f GB/s\n", n,
              n * 8.0 / 1048576.0, n * 8.0 / 1048576.0, ms / reps, 2.0 * n * 8 / (ms / reps * 1e-3) / 1e9);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
