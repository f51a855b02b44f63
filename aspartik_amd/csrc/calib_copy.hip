// Measurement tool, not part of the library (Makefile target `calib`).
//
// 1. A copy kernel of known size, used to calibrate rocprofv3's FETCH_SIZE /
//    WRITE_SIZE counters (KiB per dispatch) against exact byte counts.
// 2. Bandwidth ceilings of this device for the access mixes the pruning
//    kernels have: copy (read + write), write-only fill with nontemporal
//    stores (a full FP32 evaluation is almost all writes), and read-only.
// Usage: calib_copy [MiB]   (default 1024 MiB per buffer)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

typedef float f4v __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void copy_kernel(const float4 *__restrict__ src,
                                                   float4 *__restrict__ dst,
                                                   size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// Each thread writes 4 float4 per iteration, 256 threads apart, with
// nontemporal stores (the store form the pruning kernels use).
__global__ __launch_bounds__(256) void fill_kernel(f4v *__restrict__ dst, size_t n, float v) {
  const size_t stride = (size_t)gridDim.x * blockDim.x * 4;
  const f4v x = {v, v + 1.0f, v + 2.0f, v + 3.0f};
  for (size_t i = (size_t)blockIdx.x * blockDim.x * 4 + threadIdx.x; i < n; i += stride) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const size_t j = i + (size_t)k * blockDim.x;
      if (j < n) __builtin_nontemporal_store(x, &dst[j]);
    }
  }
}

// Read-only: 4 float4 loads in flight per thread, reduced to one value per
// thread so no load is dead; the result is stored only when it is impossible
// (all inputs are 1.0f), so there are no write bytes.
__global__ __launch_bounds__(256) void read_kernel(const float4 *__restrict__ src, size_t n,
                                                   float *sink) {
  const size_t stride = (size_t)gridDim.x * blockDim.x * 4;
  float acc = 0.0f;
  for (size_t i = (size_t)blockIdx.x * blockDim.x * 4 + threadIdx.x; i < n; i += stride) {
    float4 v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const size_t j = i + (size_t)k * blockDim.x;
      v[k] = j < n ? src[j] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) acc += v[k].x + v[k].y + v[k].z + v[k].w;
  }
  if (acc < 0.0f) sink[0] = acc;
}

template <typename F>
static float best_ms(F launch, int reps) {
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  float best = 1e30f;
  for (int r = 0; r < reps; r++) {
    (void)hipEventRecord(e0);
    launch();
    (void)hipEventRecord(e1);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (ms < best) best = ms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}

int main(int argc, char **argv) {
  const size_t mib = argc > 1 ? strtoull(argv[1], nullptr, 10) : 1024;
  const size_t bytes = mib << 20, n = bytes / sizeof(float4);
  float4 *a = nullptr, *b = nullptr;
  float *sink = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess ||
      hipMalloc(&sink, 256) != hipSuccess) {
    fprintf(stderr, "hipMalloc failed\n");
    return 1;
  }
  // 0x3f800000 = 1.0f in every float
  (void)hipMemsetD32(reinterpret_cast<hipDeviceptr_t>(a), 0x3f800000u, bytes / 4);
  (void)hipMemset(b, 0, bytes);
  (void)hipDeviceSynchronize();
  const int reps = 5;
  // the counter calibration reads the first dispatches: the copy comes first
  const float copy = best_ms([&] {
    hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, 0, a, b, n);
  }, reps);
  const float fill = best_ms([&] {
    hipLaunchKernelGGL(fill_kernel, dim3(8192), dim3(256), 0, 0,
                       reinterpret_cast<f4v *>(b), n, 2.0f);
  }, reps);
  const float rd = best_ms([&] {
    hipLaunchKernelGGL(read_kernel, dim3(8192), dim3(256), 0, 0, a, n, sink);
  }, reps);
  printf("copy %zu MiB: read+write %.1f GB/s (best of %d, %.3f ms)\n", mib,
         2.0 * bytes / (copy * 1e-3) / 1e9, reps, copy);
  printf("fill %zu MiB (nontemporal stores): %.1f GB/s (%.3f ms)\n", mib,
         bytes / (fill * 1e-3) / 1e9, fill);
  printf("read %zu MiB: %.1f GB/s (%.3f ms)\n", mib, bytes / (rd * 1e-3) / 1e9, rd);
  printf("bytes_read_per_dispatch: %zu\nbytes_written_per_dispatch: %zu\n", bytes, bytes);
  (void)hipFree(a);
  (void)hipFree(b);
  (void)hipFree(sink);
  return 0;
}
