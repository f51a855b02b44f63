// Shared definitions of the b3hip library (host and device).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "entry.hpp"

namespace b3h {

constexpr double LN2 = 0.693147180559945309417232121458176568;

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
