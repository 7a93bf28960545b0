// Plain device-to-device copy of a known size, used to calibrate the
// FETCH_SIZE / WRITE_SIZE counters (KiB per dispatch) that bench profiles read.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void copy_u32x4(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

int main(int argc, char** argv) {
  const size_t bytes = (argc > 1 ? (size_t)atoll(argv[1]) : (size_t)1 << 30);
  const size_t n = bytes / sizeof(uint4);
  uint4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * sizeof(uint4)) != hipSuccess || hipMalloc(&b, n * sizeof(uint4)) != hipSuccess) return 1;
  (void)hipMemset(a, 1, n * sizeof(uint4));
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  for (int it = 0; it < 4; ++it) {
    (void)hipEventRecord(e0);
    hipLaunchKernelGGL(copy_u32x4, dim3(8192), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e1);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    printf("copy %zu bytes: %.3f ms, %.1f GB/s (read+write)\n", n * sizeof(uint4), ms, 2.0 * n * sizeof(uint4) / ms / 1e6);
  }
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
