// CPU check of the kernel schedule (aspartik_amd/csrc/schedule.hpp).
// Emulates how prune_f32_kernel consumes entries (registers for SRC_PREV, a
// bounded LDS stack for SRC_LDS, the two-slot HBM store for SRC_HBM) with a
// nonlinear node function, over random trees, random dirty sets, EditBuf
// slot flips and stack capacities, and compares the root value with a direct
// recursive evaluation.  Prints "OK <cases>" or the first failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "../aspartik_amd/csrc/schedule.hpp"
extern "C" {
#include "../oracle/b3_oracle.h"
}

using namespace b3h;

int main(int argc, char **argv) {
  const int trials = argc > 1 ? atoi(argv[1]) : 300;
  std::mt19937_64 rng(12345);
  int cases = 0;
  for (int trial = 0; trial < trials; trial++) {
    const uint32_t L = 2 + rng() % 400;
    const uint32_t N = 2 * L - 1;
    std::vector<uint32_t> parents(N, OR_TOMBSTONE), children(2 * (L - 1));
    std::vector<double> heights(N, 0.0);
    std::vector<uint32_t> lin(L);
    for (uint32_t i = 0; i < L; i++) lin[i] = i;
    double t = 0;
    uint32_t nxt = L;
    const bool caterpillar = trial % 7 == 0;
    while (lin.size() > 1) {
      t += 0.1;
      size_t i = caterpillar ? 0 : rng() % lin.size();
      uint32_t a = lin[i];
      lin.erase(lin.begin() + i);
      size_t j = caterpillar ? lin.size() - 1 : rng() % lin.size();
      uint32_t b = lin[j];
      lin.erase(lin.begin() + j);
      parents[a] = parents[b] = nxt;
      children[2 * (nxt - L)] = a;
      children[2 * (nxt - L) + 1] = b;
      heights[nxt] = t;
      lin.push_back(nxt++);
    }
    or_tree *tr = or_tree_new(L, parents.data(), children.data(), heights.data());
    std::vector<double> leafv(L);
    for (auto &v : leafv) v = (double)(rng() % 1000) / 1000.0;
    std::vector<double> hbm(2 * (L - 1), NAN);
    std::vector<uint8_t> eb(L - 1, 0);
    auto f = [](double a, double b, uint32_t node) {
      return std::sin(a) * 1.3 + std::cos(b) + 0.01 * node;
    };
    auto truth = [&](auto &&self, uint32_t n) -> double {
      if (n < L) return leafv[n];
      uint32_t a = children[2 * (n - L)], b = children[2 * (n - L) + 1];
      if (a > b) std::swap(a, b);
      return f(self(self, a), self(self, b), n);
    };
    // the FP32 path's cherry rule: a cherry slot's value is never stored;
    // readers must carry SRC_CHERRY exactly when the slot holds a cherry and
    // recompute it from its recorded leaves
    std::vector<uint8_t> slot_cherry(2 * (L - 1), 0);
    struct Rec {
      uint32_t a, b, node;
    };
    std::vector<Rec> crec(2 * (L - 1), Rec{0, 0, 0});
    Scheduler s;
    s.capacity = 1 + rng() % 6;
    for (int step = 0; step < 8; step++) {
      if (step == 0) {
        or_tree_mark_all_edges_updated(tr);
      } else {
        const int m = 1 + rng() % 3;
        for (int k = 0; k < m; k++) or_tree_mark_edge_updated(tr, rng() % (N - 1));
      }
      std::vector<uint32_t> ints(N), ch(2 * N);
      const uint32_t n = or_tree_partials_lists(tr, ints.data(), ch.data());
      std::vector<uint8_t> e2 = eb;
      for (uint32_t i = 0; i < n; i++) or_editbuf_set_edited(e2.data(), ints[i] - L);
      for (uint32_t i = 0; i < n; i++)
        slot_cherry[Scheduler::slot_of(ints[i], L, e2.data())] = ch[2 * i] < L && ch[2 * i + 1] < L;
      std::string err;
      if (!s.build(n, ints.data(), ch.data(), L, e2.data(), true, s.capacity, &err,
                   slot_cherry.data())) {
        printf("FAIL build: %s\n", err.c_str());
        return 1;
      }
      eb = e2;
      if (s.slots_used > s.capacity) {
        printf("FAIL slots_used %u > capacity %u\n", s.slots_used, s.capacity);
        return 1;
      }
      std::vector<double> stk(64, NAN);
      double x = NAN;
      std::vector<uint8_t> seen(n, 0);
      for (uint32_t i = 0; i < n; i++) {
        const Entry &e = s.entries[i];
        const uint32_t orig = entry_index(e.meta);
        if (orig >= n || seen[orig]) {
          printf("FAIL entry order\n");
          return 1;
        }
        seen[orig] = 1;
        auto get = [&](uint32_t src, uint32_t chid) -> double {
          const uint32_t tag = src & SRC_TAG, idx = src & SRC_IDX;
          if (tag == SRC_TIP) {
            if (idx != chid) { printf("FAIL tip id\n"); exit(1); }
            return leafv[idx];
          }
          if (tag == SRC_PREV) return x;
          if (tag == SRC_LDS) {
            if (idx >= s.slots_used) { printf("FAIL stack read out of range\n"); exit(1); }
            return stk[idx];
          }
          const uint32_t slot = src & SRC_SLOT;
          if (slot != 2 * (chid - L) + (eb[chid - L] & 1u)) { printf("FAIL hbm slot\n"); exit(1); }
          if ((src & SRC_CHERRY) != 0 != (slot_cherry[slot] != 0)) {
            printf("FAIL cherry tag\n");
            exit(1);
          }
          if (src & SRC_CHERRY) return f(leafv[crec[slot].a], leafv[crec[slot].b], crec[slot].node);
          return hbm[slot];
        };
        const double a = get(e.left, ch[2 * orig]);
        const double b = get(e.right, ch[2 * orig + 1]);
        x = f(a, b, ints[orig]);
        if (e.out != 2 * (ints[orig] - L) + (eb[ints[orig] - L] & 1u)) {
          printf("FAIL out slot\n");
          return 1;
        }
        if (ch[2 * orig] < L && ch[2 * orig + 1] < L && !(e.meta & META_KEEP)) {
          hbm[e.out] = NAN;  // not stored
          crec[e.out] = Rec{ch[2 * orig], ch[2 * orig + 1], ints[orig]};
        } else {
          hbm[e.out] = x;
        }
        const uint32_t push = e.meta & 0xFFu;
        if (push != NO_PUSH) {
          if (push >= s.capacity) { printf("FAIL push beyond capacity\n"); return 1; }
          stk[push] = x;
        }
      }
      const double tv = truth(truth, or_tree_root(tr));
      cases++;
      if (!(std::fabs(tv - x) <= 1e-12)) {
        printf("FAIL trial %d step %d L=%u n=%u cap=%u: %.17g vs %.17g\n", trial, step, L, n,
               s.capacity, tv, x);
        return 1;
      }
      for (auto &m : eb) m &= 1u;  // accept
      or_tree_accept(tr);
    }
    or_tree_free(tr);
  }
  printf("OK %d\n", cases);
  return 0;
}
