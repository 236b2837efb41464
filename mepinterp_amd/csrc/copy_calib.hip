// copy_calib.hip -- plain device copy of a known size (1 GiB read + 1 GiB written):
// the reference point that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings.
#include <hip/hip_runtime.h>
#include <cstdio>
__global__ void copyk(const double4* __restrict__ a, double4* __restrict__ b, size_t n) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}
int main() {
  const size_t bytes = 1ull << 30;  // 1 GiB read + 1 GiB written
  const size_t n = bytes / sizeof(double4);
  double4 *a, *b;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
  hipMemset(a, 0, bytes);
  hipLaunchKernelGGL(copyk, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, a, b, n);
  hipDeviceSynchronize();
  hipLaunchKernelGGL(copyk, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, a, b, n);
  hipError_t e = hipDeviceSynchronize();
  printf("copy of %zu bytes done: %s\n", bytes, hipGetErrorString(e));
  hipFree(a); hipFree(b);
  return e == hipSuccess ? 0 : 1;
}
