// Shared definitions of the b3hip library (host and device).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace b3h {

// One update-list entry as the pruning kernels read it (16 bytes, loaded
// wave-uniformly with scalar loads).
//   out   : partial slot the node's result is stored in (slot = internal
//           index * 2 + EditBuf active bit, crates/sk/src/editbuf.rs:36-40)
//   left  : source of the left child (SRC_* tag in bits 31..30)
//   right : source of the right child
//   meta  : bits 7..0 LDS stack slot the result is pushed to (NO_PUSH: none),
//           bits 31..8 the entry's index in the caller's list (which selects
//           its transition matrices and edge times)
struct Entry {
  uint32_t out;
  uint32_t left;
  uint32_t right;
  uint32_t meta;
};

constexpr uint32_t SRC_TIP = 0u << 30;  // low bits: leaf id
constexpr uint32_t SRC_HBM = 1u << 30;  // low bits: partial slot in HBM
constexpr uint32_t SRC_LDS = 2u << 30;  // low bits: LDS stack slot
constexpr uint32_t SRC_TAG = 3u << 30;
constexpr uint32_t SRC_IDX = (1u << 30) - 1;
constexpr uint32_t NO_PUSH = 0xFFu;

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
