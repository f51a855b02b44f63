// Copy kernel of known size, used to calibrate rocprofv3's FETCH_SIZE /
// WRITE_SIZE counters (KiB per dispatch) against exact byte counts.
// Usage: calib_copy [MiB]   (default 1024 MiB read + 1024 MiB written)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void copy_kernel(const float4 *__restrict__ src,
                                                   float4 *__restrict__ dst,
                                                   size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

int main(int argc, char **argv) {
  const size_t mib = argc > 1 ? strtoull(argv[1], nullptr, 10) : 1024;
  const size_t bytes = mib << 20, n = bytes / sizeof(float4);
  float4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    fprintf(stderr, "hipMalloc failed\n");
    return 1;
  }
  (void)hipMemset(a, 1, bytes);
  (void)hipMemset(b, 0, bytes);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  const int reps = 5;
  float best = 1e30f;
  for (int r = 0; r < reps; r++) {
    (void)hipEventRecord(e0);
    hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e1);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (ms < best) best = ms;
  }
  printf("copy %zu MiB: read+write %.1f GB/s (best of %d, %.3f ms)\n", mib,
         2.0 * bytes / (best * 1e-3) / 1e9, reps, best);
  printf("bytes_read_per_dispatch: %zu\nbytes_written_per_dispatch: %zu\n", bytes, bytes);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
