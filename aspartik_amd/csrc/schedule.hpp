// Update-list scheduling for the pruning kernels.
//
// Input is what b3's DatedTree::partials_lists() produces (crates/b3/src/
// parameters/tree.rs:398-461): the dirty internal nodes with children before
// parents and the root last, and each node's sorted child pair.  The node set
// is kept; only the order may change, which does not change any node's
// value.
//
// For the FP32 kernel the list is re-ordered depth first, visiting the child
// subtree with the larger register need first (Strahler / Sethi-Ullman
// order).  A child whose value is produced in the same pass is then consumed
// from a small per-wavefront LDS stack: stack slot b holds the value of the
// subtree evaluated at depth b, and a subtree needs at most its Strahler
// number of slots (<= log2(leaves) + 1).  Subtrees deeper than the stack
// capacity fall back to reading the child's stored partial from HBM, which
// every node writes anyway.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "common.hpp"

namespace b3h {

struct Scheduler {
  uint32_t capacity = 16;  // LDS stack slots per wavefront
  uint32_t slots_used = 0;
  std::vector<Entry> entries;

  std::vector<int32_t> pos;    // node id -> list index, -1 when absent
  std::vector<int32_t> kid;    // [2n] in-pass child list index or -1
  std::vector<uint32_t> need;  // [n]
  std::vector<uint8_t> consumed;
  std::vector<uint32_t> base;

  static uint32_t slot_of(uint32_t node, uint32_t L, const uint8_t *editbuf) {
    return 2u * (node - L) + (editbuf[node - L] & 1u);
  }

  bool build(uint32_t n, const uint32_t *internals, const uint32_t *children,
             uint32_t L, const uint8_t *editbuf, bool reorder, uint32_t cap,
             std::string *err) {
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
    auto src_of = [&](uint32_t i, int s) -> uint32_t {
      const uint32_t ch = children[2 * i + s];
      if (ch < L) return SRC_TIP | ch;
      const int32_t k = kid[2 * i + s];
      if (reorder && k >= 0 && base[k] < cap) return SRC_LDS | base[k];
      return SRC_HBM | slot_of(ch, L, editbuf);
    };
    if (!reorder) {
      for (uint32_t i = 0; i < n; i++) {
        Entry &e = entries[i];
        e.out = slot_of(internals[i], L, editbuf);
        e.left = src_of(i, 0);
        e.right = src_of(i, 1);
        e.meta = NO_PUSH | (i << 8);
      }
      return true;
    }
    // register need of every in-pass subtree (list order is post-order)
    need.assign(n, 1);
    for (uint32_t i = 0; i < n; i++) {
      const int32_t a = kid[2 * i], b = kid[2 * i + 1];
      if (a >= 0 && b >= 0)
        need[i] = need[a] == need[b] ? need[a] + 1 : (need[a] > need[b] ? need[a] : need[b]);
      else if (a >= 0)
        need[i] = need[a];
      else if (b >= 0)
        need[i] = need[b];
    }
    base.assign(n, 0);
    struct Frame {
      int32_t v;
      uint32_t base;
      uint8_t state;
    };
    std::vector<Frame> st;
    st.reserve(64);
    uint32_t out_i = 0;
    for (uint32_t r = 0; r < n; r++) {
      if (consumed[r]) continue;  // not the top of an in-pass subtree
      st.push_back({(int32_t)r, 0, 0});
      while (!st.empty()) {
        const size_t top = st.size() - 1;
        const int32_t v = st[top].v;
        const uint32_t b0 = st[top].base;
        int32_t first = kid[2 * v], second = kid[2 * v + 1];
        if (first < 0 || (second >= 0 && need[second] > need[first])) {
          const int32_t t = first;
          first = second;
          second = t;
        }
        if (st[top].state == 0) {
          st[top].state = 1;
          if (first >= 0) {
            st.push_back({first, b0, 0});
            continue;
          }
        }
        if (st[top].state == 1) {
          st[top].state = 2;
          if (second >= 0) {
            st.push_back({second, b0 + 1, 0});
            continue;
          }
        }
        base[v] = b0;
        Entry &e = entries[out_i++];
        e.out = slot_of(internals[v], L, editbuf);
        e.left = src_of((uint32_t)v, 0);
        e.right = src_of((uint32_t)v, 1);
        const bool pushed = consumed[v] && b0 < cap;
        e.meta = (pushed ? b0 : NO_PUSH) | ((uint32_t)v << 8);
        if (pushed && b0 + 1 > slots_used) slots_used = b0 + 1;
        st.pop_back();
      }
    }
    return out_i == n;
  }
};

}  // namespace b3h
