// Sequence simulation along a tree (host, workload generation).
//
// For each site: a rate category uniform over K, a root state drawn from the
// stationary frequencies, and each child's state drawn from row (parent
// state) of the edge's transition matrix for that category.  Leaf states are
// written [L x sites].  Sites are split over threads; every site has its own
// counter-based random stream (splitmix64 of seed and site), so the output
// does not depend on the thread count.
#include <cstdint>
#include <thread>
#include <vector>

namespace {

inline uint64_t splitmix(uint64_t &x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline double unif(uint64_t &x) { return (splitmix(x) >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

extern "C" {

// order[N_int]: internal nodes, parents before children; root = order[0].
// children[2 * (node - L)]; cum[(edge * K + c) * S * S + from * S + to]:
// cumulative rows of P(t_edge * r_c) for edge = child id; freqs_cum[S].
// out: [L x sites] state indices.
int b3sim_states(uint32_t L, uint32_t S, uint32_t K, const uint32_t *order,
                 const uint32_t *children, const double *cum, const double *freqs_cum,
                 uint64_t sites, uint64_t seed, uint32_t threads, uint8_t *out) {
  if (L < 2 || S < 2 || S > 255 || K < 1 || !order || !children || !cum || !freqs_cum || !out)
    return 1;
  const uint32_t N = 2 * L - 1;
  auto work = [&](uint64_t begin, uint64_t end) {
    std::vector<uint8_t> st(N);
    for (uint64_t site = begin; site < end; site++) {
      uint64_t x = seed ^ (site * 0xD1B54A32D192ED03ull);
      splitmix(x);
      const uint32_t c = (uint32_t)(splitmix(x) % K);
      double u = unif(x) * freqs_cum[S - 1];
      uint32_t s = 0;
      while (s + 1 < S && u >= freqs_cum[s]) s++;
      st[order[0]] = (uint8_t)s;
      for (uint32_t k = 0; k < L - 1; k++) {
        const uint32_t node = order[k];
        const uint32_t from = st[node];
        for (int side = 0; side < 2; side++) {
          const uint32_t ch = children[2 * (node - L) + side];
          const double *row = cum + ((size_t)ch * K + c) * S * S + (size_t)from * S;
          const double v = unif(x) * row[S - 1];
          uint32_t t = 0;
          while (t + 1 < S && v >= row[t]) t++;
          st[ch] = (uint8_t)t;
        }
      }
      for (uint32_t l = 0; l < L; l++) out[(size_t)l * sites + site] = st[l];
    }
  };
  if (threads < 1) threads = 1;
  std::vector<std::thread> pool;
  const uint64_t chunk = (sites + threads - 1) / threads;
  for (uint32_t t = 0; t < threads; t++) {
    const uint64_t b = t * chunk, e = b + chunk < sites ? b + chunk : sites;
    if (b >= e) break;
    pool.emplace_back(work, b, e);
  }
  for (auto &th : pool) th.join();
  return 0;
}

}  // extern "C"
