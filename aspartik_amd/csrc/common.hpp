// Shared definitions of the b3hip library (host and device).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "entry.hpp"

namespace b3h {

constexpr double LN2 = 0.693147180559945309417232121458176568;

// Cumulative binary exponents of scaled partials are stored as int16.  A
// pattern whose exponent at some node leaves [-32767, 32767] (a site ln L
// below about -22,700 nats) cannot be represented: from that node on its
// exponent is EXP_DEAD, stored as INT16_MIN, and the root reports ln L =
// -inf for it (b3's Mcmc::step rejects such a state, mcmc.rs:339-341)
// instead of a wrapped value.  Full and partial evaluations agree, because
// the rule applies at every node whether the child came from registers, LDS
// or memory.
constexpr int EXP_DEAD = -(1 << 28);

__host__ __device__ __forceinline__ int exp_clamp(int e) {
  return (e < -32767 || e > 32767) ? EXP_DEAD : e;
}
__host__ __device__ __forceinline__ int16_t exp_pack(int e) {
  return (int16_t)(e == EXP_DEAD ? -32768 : e);
}
__host__ __device__ __forceinline__ int exp_unpack(int16_t s) {
  return s == -32768 ? EXP_DEAD : (int)s;
}
// site ln L from the root value's dot product with pi and its exponent
__host__ __device__ __forceinline__ double exp_site(double dot, int e) {
  return e == EXP_DEAD ? -INFINITY : fma((double)e, LN2, log(dot));
}

// Substitution model held on the device (FP64).
struct DevModel {
  int32_t n;
  int32_t special;  // apply t > 20 / t < 1e-10 special cases
  int32_t pad[2];
  double freqs[64];
  double q[64 * 64];
  double evec[64 * 64];
  double ievc[64 * 64];
  double eval[64];
};

}  // namespace b3h
