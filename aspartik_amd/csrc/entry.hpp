// Update-list entry format shared by the host scheduler and the kernels.
// Plain C++ (no HIP), so the scheduler can also be built for CPU tests.
#pragma once

#include <stdint.h>

namespace b3h {

// One update-list entry as the pruning kernels read it (16 bytes).
//   out   : partial slot the node's result is stored in (slot = internal
//           index * 2 + EditBuf active bit, crates/sk/src/editbuf.rs:36-40)
//   left  : source of the left child (SRC_* tag in bits 31..30; for SRC_HBM
//           bit 29 is SRC_CHERRY)
//   right : source of the right child
//   meta  : bits 7..0 LDS stack slot the result is pushed to (NO_PUSH: none),
//           bits 30..8 the entry's index in the caller's list (which selects
//           its transition matrices and edge times), bit 31 META_KEEP
struct Entry {
  uint32_t out;
  uint32_t left;
  uint32_t right;
  uint32_t meta;
};

constexpr uint32_t SRC_TIP = 0u << 30;   // low bits: leaf id
constexpr uint32_t SRC_HBM = 1u << 30;   // low bits: partial slot in HBM
constexpr uint32_t SRC_LDS = 2u << 30;   // low bits: LDS stack slot
constexpr uint32_t SRC_PREV = 3u << 30;  // the entry computed just before
constexpr uint32_t SRC_TAG = 3u << 30;
constexpr uint32_t SRC_IDX = (1u << 30) - 1;
// With SRC_HBM: the slot holds a cherry (both children leaves), whose partial
// the FP32 path does not store; readers recompute it from the two leaves'
// codes and the two transition matrices recorded for the slot.
constexpr uint32_t SRC_CHERRY = 1u << 29;
constexpr uint32_t SRC_SLOT = SRC_CHERRY - 1;  // slot of an SRC_HBM source
constexpr uint32_t NO_PUSH = 0xFFu;
// meta bit 31: the FP32 path stores this cherry's partial after all, because
// a parent in the same pass reads it back from memory (the slot is then not
// a cherry slot); bits 30..8 are the list index.
constexpr uint32_t META_KEEP = 1u << 31;
inline constexpr uint32_t entry_index(uint32_t meta) { return (meta >> 8) & 0x7FFFFFu; }

}  // namespace b3h
