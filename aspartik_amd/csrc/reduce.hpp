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
// tile that takes the last ticket on `counter` adds all partials in a fixed
// order (64 strided lane sums, then a fixed butterfly), so the result does
// not depend on timing, and resets the counter.  `flag` is an LDS int.  Must
// be called by all threads of the block.
//
// Ordering without an L2 write-back: the partial is written with an
// agent-scope atomic store, which writes through to the level every XCD
// sees, and the ticket is taken only after that store has completed (a
// workgroup-scope release plus vmcnt(0), which orders but flushes nothing);
// the last tile reads the partials with agent-scope atomic loads, which read
// that level.  An agent-scope __threadfence() here would write back the
// whole L2 (buffer_wbl2) once per tile, which measured slower than the
// pruning itself at thousands of tiles.  rocPRIM's look-back scan state uses
// the same order-only idiom.
__device__ __forceinline__ void last_block_total(double block_value, double *blocksum,
                                                 uint32_t *counter, double *total,
                                                 double *host_total, uint32_t *host_flag,
                                                 uint32_t seq, int *flag, uint32_t index,
                                                 uint32_t count) {
  if (threadIdx.x == 0) {
    __hip_atomic_store(&blocksum[index], block_value, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0 | (0x7 << 4) | (0xf << 8));  // vmcnt(0)
    const uint32_t ticket =
        __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    *flag = (ticket == count - 1) ? 1 : 0;
  }
  __syncthreads();
  if (*flag && threadIdx.x < 64) {
    __builtin_amdgcn_s_waitcnt(0 | (0x7 << 4) | (0xf << 8));
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    double acc = 0.0;
    for (uint32_t j = threadIdx.x; j < count; j += 64)
      acc += __hip_atomic_load(&blocksum[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    acc = wave_sum(acc);
    if (threadIdx.x == 0) {
      *total = acc;
      // fine-grained pinned host memory: the value, then the sequence
      // number with a system-scope release, so a host that sees the number
      // also sees the value, before the kernel has retired
      if (host_total) {
        *host_total = acc;
        __hip_atomic_store(host_flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
      }
      __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

}  // namespace b3h
