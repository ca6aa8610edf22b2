// calib_copy.hip -- a plain float4 copy of 1 GiB, the reference kernel for
// scaling the FETCH_SIZE / WRITE_SIZE counters to HBM bytes (it moves exactly
// 1 GiB in and 1 GiB out per launch).  Not part of the library.
#include <hip/hip_runtime.h>
#include <cstdio>
__global__ void k_copy(const float4* __restrict__ x, float4* __restrict__ y, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) y[i] = x[i];
}
int main() {
  const size_t bytes = (size_t)1 << 30, n = bytes / sizeof(float4);
  float4 *x, *y;
  if (hipMalloc(&x, bytes) != hipSuccess || hipMalloc(&y, bytes) != hipSuccess) { printf("alloc failed\n"); return 1; }
  (void)hipMemset(x, 0, bytes);
  for (int r = 0; r < 3; r++) hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, 0, x, y, n);
  (void)hipDeviceSynchronize();
  hipEvent_t e0, e1; (void)hipEventCreate(&e0); (void)hipEventCreate(&e1);
  (void)hipEventRecord(e0);
  hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, 0, x, y, n);
  (void)hipEventRecord(e1); (void)hipEventSynchronize(e1);
  float ms = 0; (void)hipEventElapsedTime(&ms, e0, e1);
  printf("copy %zu bytes: %.3f ms, %.1f GB/s (read+write)\n", bytes, ms, 2.0 * bytes / (ms * 1e-3) / 1e9);
  (void)hipFree(x); (void)hipFree(y);
  return 0;
}
