// Deterministic FP64 reductions shared by the pruning kernels.
#pragma once

#include "common.hpp"

namespace b3h {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) v += __shfl_xor(v, s);
  return v;
}

// Each of `count` tiles contributes one partial sum at blocksum[index]; the
// tile that arrives last sums all partials in a fixed order (64 strided lane
// sums, then a fixed butterfly) so the result does not depend on timing, and
// resets the counter.  `flag` is an LDS int.  Must be called by all threads
// of the block.
__device__ __forceinline__ void last_block_total(double block_value,
                                                 double *blocksum,
                                                 uint32_t *counter,
                                                 double *total, int *flag,
                                                 uint32_t index, uint32_t count) {
  if (threadIdx.x == 0) {
    blocksum[index] = block_value;
    __threadfence();
    const uint32_t ticket = atomicAdd(counter, 1u);
    *flag = (ticket == count - 1) ? 1 : 0;
  }
  __syncthreads();
  if (*flag && threadIdx.x < 64) {
    __threadfence();
    double acc = 0.0;
    for (uint32_t j = threadIdx.x; j < count; j += 64)
      acc += __hip_atomic_load(&blocksum[j], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    acc = wave_sum(acc);
    if (threadIdx.x == 0) {
      *total = acc;
      __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

}  // namespace b3h
