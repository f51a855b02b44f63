// Update-list scheduling for the pruning kernels (host, plain C++).
//
// Input is what b3's DatedTree::partials_lists() produces (crates/b3/src/
// parameters/tree.rs:398-461): the dirty internal nodes with children before
// parents and the root last, and each node's sorted child pair.  The node set
// is kept; only the order may change, which does not change any node's value.
//
// For the FP32 kernel the list is re-ordered depth first (post-order).  A
// child whose value is produced in the same pass is then consumed without
// touching HBM:
//  - the child emitted immediately before its parent (the second child, or
//    an only child) is still in the lane's registers: SRC_PREV;
//  - the other (first) child was pushed to a per-wavefront LDS stack slot:
//    SRC_LDS.  A subtree evaluated at stack base b leaves a pushed result in
//    slot b and uses slots >= b; the child order at each node is chosen to
//    minimise the number of slots (a Sethi-Ullman labelling), which is at
//    most about log2(leaves).
// Subtrees deeper than the stack capacity fall back to reading the child's
// stored partial from HBM, which every node writes anyway.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "entry.hpp"

namespace b3h {

struct Scheduler {
  uint32_t capacity = 16;  // LDS stack slots per wavefront
  uint32_t slots_used = 0;
  uint32_t hbm_sources = 0;  // children read from HBM
  uint32_t cherry_sources = 0;  // of those, stored cherries (SRC_CHERRY)
  std::vector<Entry> entries;

  std::vector<int32_t> pos;    // node id -> list index, -1 when absent
  std::vector<int32_t> kid;    // [2n] in-pass child list index or -1
  std::vector<int32_t> first;  // [n] side (0/1) evaluated first, -1 if < 2 kids
  std::vector<uint32_t> need;  // slots to evaluate a subtree, result in registers
  std::vector<uint8_t> consumed;
  std::vector<uint8_t> pushed;
  std::vector<uint32_t> base;
  std::vector<uint32_t> outpos;  // list index -> position in entries

  // partial slot of an internal node: two slots per node selected by the
  // EditBuf active bit, or one slot per node when editbuf is null
  static uint32_t slot_of(uint32_t node, uint32_t L, const uint8_t *editbuf) {
    if (!editbuf) return node - L;
    return 2u * (node - L) + (editbuf[node - L] & 1u);
  }

  // slot_cherry (may be null): per partial slot, 1 when it holds a cherry
  // (both children leaves), whose partial the FP32 path does not store.  A
  // stored child in such a slot gets SRC_CHERRY.  A cherry of this pass that
  // its parent reads back from memory (LDS stack overflow) is stored after
  // all: its entry gets META_KEEP and its slot is no longer a cherry slot.
  bool build(uint32_t n, const uint32_t *internals, const uint32_t *children,
             uint32_t L, const uint8_t *editbuf, bool reorder, uint32_t cap,
             std::string *err, uint8_t *slot_cherry = nullptr) {
    outpos.assign(n, 0);
    // k: the child's list index when it is computed in this pass, else -1
    auto hbm_src = [&](uint32_t ch, int32_t k) {
      const uint32_t s = slot_of(ch, L, editbuf);
      if (!slot_cherry || !slot_cherry[s]) return SRC_HBM | s;
      if (k >= 0) {
        slot_cherry[s] = 0;
        entries[outpos[k]].meta |= META_KEEP;
        return SRC_HBM | s;
      }
      cherry_sources++;
      return SRC_HBM | SRC_CHERRY | s;
    };
    const uint32_t N = 2 * L - 1;
    pos.assign(N, -1);
    for (uint32_t i = 0; i < n; i++) {
      if (pos[internals[i]] >= 0) {
        *err = "node " + std::to_string(internals[i]) + " listed twice";
        return false;
      }
      pos[internals[i]] = (int32_t)i;
    }
    kid.assign(2 * (size_t)n, -1);
    consumed.assign(n, 0);
    for (uint32_t i = 0; i < n; i++) {
      for (int s = 0; s < 2; s++) {
        const uint32_t ch = children[2 * i + s];
        if (ch >= L && pos[ch] >= 0) {
          if ((uint32_t)pos[ch] >= i) {
            *err = "child " + std::to_string(ch) + " listed after its parent";
            return false;
          }
          if (consumed[pos[ch]]) {
            *err = "node " + std::to_string(ch) + " has two parents in the list";
            return false;
          }
          consumed[pos[ch]] = 1;
          kid[2 * i + s] = pos[ch];
        }
      }
    }
    if (consumed[n - 1]) {
      *err = "the last listed node must be the root of the update";
      return false;
    }
    entries.resize(n);
    slots_used = 0;
    hbm_sources = 0;
    cherry_sources = 0;
    if (!reorder) {
      for (uint32_t i = 0; i < n; i++) {
        Entry &e = entries[i];
        outpos[i] = i;
        e.out = slot_of(internals[i], L, editbuf);
        e.meta = NO_PUSH | (i << 8);
        for (int s = 0; s < 2; s++) {
          const uint32_t ch = children[2 * i + s];
          const uint32_t src = ch < L ? (SRC_TIP | ch) : hbm_src(ch, kid[2 * i + s]);
          hbm_sources += ch >= L;
          (s ? e.right : e.left) = src;
        }
      }
      return true;
    }
    // Sethi-Ullman labelling over the in-pass forest (list order = post-order)
    need.assign(n, 0);
    first.assign(n, -1);
    for (uint32_t i = 0; i < n; i++) {
      const int32_t a = kid[2 * i], b = kid[2 * i + 1];
      if (a >= 0 && b >= 0) {
        const uint32_t pa = need[a] > 1 ? need[a] : 1, pb = need[b] > 1 ? need[b] : 1;
        const uint32_t o1 = pa > 1 + need[b] ? pa : 1 + need[b];  // a first
        const uint32_t o2 = pb > 1 + need[a] ? pb : 1 + need[a];  // b first
        need[i] = o1 <= o2 ? o1 : o2;
        first[i] = o1 <= o2 ? 0 : 1;
      } else if (a >= 0) {
        need[i] = need[a];
      } else if (b >= 0) {
        need[i] = need[b];
      }
    }
    base.assign(n, 0);
    pushed.assign(n, 0);
    struct Frame {
      int32_t v;
      uint32_t base;
      uint8_t state;
      uint8_t push;
    };
    std::vector<Frame> st;
    st.reserve(64);
    uint32_t out_i = 0;
    for (uint32_t r = 0; r < n; r++) {
      if (consumed[r]) continue;  // not the top of an in-pass subtree
      st.push_back({(int32_t)r, 0, 0, 0});
      while (!st.empty()) {
        const size_t top = st.size() - 1;
        const int32_t v = st[top].v;
        const uint32_t b0 = st[top].base;
        const int32_t a = kid[2 * v], b = kid[2 * v + 1];
        int32_t k1 = -1, k2 = -1;  // evaluated first (pushed), second (in registers)
        if (a >= 0 && b >= 0) {
          k1 = first[v] == 0 ? a : b;
          k2 = first[v] == 0 ? b : a;
        } else {
          k2 = a >= 0 ? a : b;
        }
        if (st[top].state == 0) {
          st[top].state = 1;
          if (k1 >= 0) {
            st.push_back({k1, b0, 0, 1});
            continue;
          }
        }
        if (st[top].state == 1) {
          st[top].state = 2;
          if (k2 >= 0) {
            st.push_back({k2, k1 >= 0 ? b0 + 1 : b0, 0, 0});
            continue;
          }
        }
        base[v] = b0;
        const bool push = st[top].push && b0 < cap;
        pushed[v] = push ? 1 : 0;
        outpos[v] = out_i;
        Entry &e = entries[out_i++];
        e.out = slot_of(internals[v], L, editbuf);
        e.meta = (push ? b0 : NO_PUSH) | ((uint32_t)v << 8);
        for (int s = 0; s < 2; s++) {
          const uint32_t ch = children[2 * v + s];
          const int32_t k = kid[2 * v + s];
          uint32_t src;
          if (ch < L) src = SRC_TIP | ch;
          else if (k >= 0 && k == k2) src = SRC_PREV;
          else if (k >= 0 && pushed[k]) src = SRC_LDS | base[k];
          else src = hbm_src(ch, k);
          hbm_sources += (src & SRC_TAG) == SRC_HBM;
          (s ? e.right : e.left) = src;
        }
        if (push && b0 + 1 > slots_used) slots_used = b0 + 1;
        st.pop_back();
      }
    }
    return out_i == n;
  }
};

}  // namespace b3h
