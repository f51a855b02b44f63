// MCMC driver over the GPU likelihood (host C++), the native counterpart of
// b3's Mcmc::step / finalize_step (crates/b3/src/mcmc.rs:314-393) for a dated
// tree under GTR+Gamma, plus delayed-acceptance MH with the MLP surrogate.
//
// State: DatedTree (tree.rs bookkeeping: dirty edges, partials lists,
// accept/reject), GTR exchangeabilities (kept summing to 1), fixed base
// frequencies, Gamma shape alpha.  Priors: Yule on the node heights,
// (n - 1) ln(lambda) - lambda * tree length (the pure-birth density of the
// ranked tree up to a constant); flat Dirichlet on the exchangeabilities;
// Exponential on alpha.  Operators as b3's Python ones (see b3mcmc.h).
//
// Delayed acceptance (Christen & Fox 2005): the first stage accepts with
// min(1, exp(df + dprior + h)), df = beta x the surrogate's change of ln L;
// only then is the exact likelihood evaluated and the move accepted with
// min(1, exp(dlnL - df)).  Topology moves go straight to the exact MH test,
// or, with screen_topology, pass the first stage too: the surrogate's
// features (log edge lengths by node label, log rates, log alpha) are a
// function of the whole labelled state, which is all the two-stage test
// needs.  Every stage is an MH test, so the chain keeps the exact posterior.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/b3hip.h"
#include "../../include/b3mcmc.h"

namespace b3h {
b3h_status host_fail(b3h_status s, const std::string &msg);
b3h_status mlp_predict(b3m_mlp *m, const float *x, uint32_t B, float *y);
}  // namespace b3h

namespace {

constexpr uint32_t kTomb = 0xFFFFFFFFu;

// xoshiro256** seeded by splitmix64 (Blackman & Vigna)
struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    for (auto &w : s) {
      seed += 0x9E3779B97F4A7C15ull;
      uint64_t z = seed;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      w = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0, 1)
  uint32_t below(uint32_t n) { return (uint32_t)(uniform() * n); }
  uint32_t range(uint32_t lo, uint32_t hi) { return lo + below(hi - lo); }
};

// DatedTree bookkeeping of crates/b3/src/parameters/tree.rs
struct Tree {
  uint32_t L = 0, N = 0, root = 0;
  std::vector<uint32_t> parents, children;  // children[2 (node - L) + side]
  std::vector<double> heights;
  std::vector<uint32_t> bp, bc;
  std::vector<double> bh;
  uint32_t broot = 0;
  std::vector<uint8_t> upd_edges, upd_partials;

  bool init(uint32_t nl, const uint32_t *par, const uint32_t *ch, const double *h,
            std::string *err) {
    L = nl;
    N = 2 * L - 1;
    parents.assign(par, par + N);
    children.assign(ch, ch + 2 * (size_t)(L - 1));
    heights.assign(h, h + N);
    uint32_t roots = 0;
    for (uint32_t v = 0; v < N; v++)
      if (parents[v] == kTomb) {
        root = v;
        roots++;
      }
    if (roots != 1 || root < L) {
      *err = "the tree must have exactly one root, an internal node";
      return false;
    }
    for (uint32_t v = L; v < N; v++)
      for (int s = 0; s < 2; s++) {
        const uint32_t c = children[2 * (v - L) + s];
        if (c >= N || parents[c] != v) {
          *err = "parents and children disagree";
          return false;
        }
      }
    for (uint32_t v = 0; v < N; v++)
      if (v != root && !(heights[parents[v]] > heights[v])) {
        *err = "every node must be younger than its parent";
        return false;
      }
    upd_edges.assign(N, 0);
    upd_partials.assign(N, 0);
    save();
    return true;
  }
  void save() {
    bp = parents;
    bc = children;
    bh = heights;
    broot = root;
  }
  uint32_t child(uint32_t v, int s) const { return children[2 * (v - L) + s]; }
  double edge_length(uint32_t e) const { return heights[parents[e]] - heights[e]; }
  bool is_internal(uint32_t v) const { return v >= L; }
  // tree.rs:357-375
  void mark_edge_updated(uint32_t e) {
    upd_edges[e] = 1;
    uint32_t cur = e;
    while (parents[cur] != kTomb) {
      cur = parents[cur];
      if (upd_partials[cur]) break;
      upd_partials[cur] = 1;
    }
  }
  void mark_all() {
    std::fill(upd_edges.begin(), upd_edges.end(), 1);
    std::fill(upd_partials.begin(), upd_partials.end(), 1);
  }
  // tree.rs:398-461: BFS from the root through marked internal children,
  // reversed (root last); sorted child pairs
  void partials_lists(std::vector<uint32_t> &internals, std::vector<uint32_t> &kids) const {
    internals.clear();
    internals.push_back(root);
    for (size_t q = 0; q < internals.size(); q++) {
      const uint32_t v = internals[q];
      for (int s = 0; s < 2; s++) {
        const uint32_t c = child(v, s);
        if (c >= L && upd_partials[c]) internals.push_back(c);
      }
    }
    std::reverse(internals.begin(), internals.end());
    kids.resize(2 * internals.size());
    for (size_t i = 0; i < internals.size(); i++) {
      uint32_t a = child(internals[i], 0), b = child(internals[i], 1);
      if (a > b) std::swap(a, b);
      kids[2 * i] = a;
      kids[2 * i + 1] = b;
    }
  }
  // tree.rs:516-528
  void set_height(uint32_t v, double h) {
    heights[v] = h;
    if (parents[v] != kTomb) mark_edge_updated(v);
    if (v >= L) {
      mark_edge_updated(child(v, 0));
      mark_edge_updated(child(v, 1));
    }
  }
  // tree.rs:471-495
  void replace_child(uint32_t parent, uint32_t old_child, uint32_t new_child) {
    parents[new_child] = parent;
    uint32_t &l = children[2 * (parent - L)], &r = children[2 * (parent - L) + 1];
    if (l == old_child) l = new_child;
    else r = new_child;
    mark_edge_updated(new_child);
  }
  // tree.rs:535-550 (validity holds for the narrow exchange's choices)
  void swap_parents(uint32_t a, uint32_t b) {
    const uint32_t pa = parents[a], pb = parents[b];
    replace_child(pa, a, b);
    replace_child(pb, b, a);
  }
  // tree.rs:767-778 (two internal children); with `loose`, BEAST2's narrow
  // Exchange (at least one internal child).  Leaves are never grandparents.
  bool is_grandparent(uint32_t v, bool loose = false) const {
    if (v < L) return false;
    return loose ? (child(v, 0) >= L || child(v, 1) >= L)
                 : (child(v, 0) >= L && child(v, 1) >= L);
  }
  uint32_t num_grandparents(bool loose = false) const {
    uint32_t n = 0;
    for (uint32_t v = L; v < N; v++) n += is_grandparent(v, loose);
    return n;
  }
  double tree_length() const {
    double s = 0.0;
    for (uint32_t v = 0; v < N; v++)
      if (v != root) s += edge_length(v);
    return s;
  }
  bool changed() const {
    for (uint8_t e : upd_edges)
      if (e) return true;
    return false;
  }
  void clear() {
    std::fill(upd_edges.begin(), upd_edges.end(), 0);
    std::fill(upd_partials.begin(), upd_partials.end(), 0);
  }
  void accept() {
    save();
    clear();
  }
  void reject() {
    parents = bp;
    children = bc;
    heights = bh;
    root = broot;
    clear();
  }
};

enum Op { kNodeSlide = 0, kTreeScale, kRates, kAlpha, kNarrow, kNumOps };

}  // namespace

struct b3m_chain {
  b3m_chain_config cfg{};
  b3h_calc *calc = nullptr;
  uint32_t L = 0, P = 0, K = 4;
  Tree tree;
  Rng rng{1};
  // parameters and their accepted values
  double rates[6] = {0}, rates_acc[6] = {0};
  double alpha = 1.0, alpha_acc = 1.0;
  double freqs[4] = {0};
  double device_rates[6] = {-1, -1, -1, -1, -1, -1};  // model currently on the GPU
  // category rates of cat_alpha (the discretised gamma costs tens of
  // microseconds on the host, and alpha changes only on its own moves)
  double cat_alpha = -1.0;
  double cat_rates[64] = {0};
  // accepted values
  double loglik = 0.0, logprior = 0.0, posterior = 0.0;
  double surrogate_cur = 0.0;
  bool surrogate_valid = false;
  b3m_mlp *mlp = nullptr;
  std::vector<uint32_t> internals, kids;
  std::vector<double> times;
  std::vector<float> feat;
  // operator tuning (b3's Scheduler::tune)
  double tree_scale_f = 0.75, alpha_f = 0.75, delta_t = 0.75;
  uint64_t life_steps = 0;
  uint64_t tune_acc[5] = {0}, tune_rej[5] = {0};
};

namespace {

b3h_status fail(b3h_status s, const std::string &msg) { return b3h::host_fail(s, msg); }

// GTR Q of substitution.rs:433-604, normalised by 2 sum r_ij pi_i pi_j
void gtr_q(const double *r, const double *f, double *q) {
  const double pa = f[0], pc = f[1], pg = f[2], pt = f[3];
  const double a = r[0], b = r[1], c = r[2], d = r[3], e = r[4], ff = r[5];
  const double m[16] = {-a * pc - b * pg - c * pt, a * pc, b * pg, c * pt,
                        a * pa, -a * pa - d * pg - e * pt, d * pg, e * pt,
                        b * pa, d * pc, -b * pa - d * pc - ff * pt, ff * pt,
                        c * pa, e * pc, ff * pg, -c * pa - e * pc - ff * pg};
  const double div =
      2.0 * (a * pa * pc + b * pa * pg + c * pa * pt + d * pc * pg + e * pc * pt + ff * pg * pt);
  for (int i = 0; i < 16; i++) q[i] = m[i] / div;
}

double log_prior(const b3m_chain *c) {
  for (int i = 0; i < 6; i++)
    if (!(c->rates[i] > 0.0)) return -INFINITY;
  if (!(c->alpha > 0.0)) return -INFINITY;
  for (uint32_t v = 0; v < c->tree.N; v++)
    if (v != c->tree.root && !(c->tree.edge_length(v) > 0.0)) return -INFINITY;
  const double lam = c->cfg.birth_rate, ra = c->cfg.alpha_prior_rate;
  return (double)(c->L - 1) * std::log(lam) - lam * c->tree.tree_length() + std::log(ra) -
         ra * c->alpha;
}

// exact ln L of the current (proposed) state on the GPU
b3h_status evaluate(b3m_chain *c, double *out) {
  bool model_changed = false;
  for (int i = 0; i < 6; i++) model_changed |= c->rates[i] != c->device_rates[i];
  if (model_changed) {
    double q[16];
    gtr_q(c->rates, c->freqs, q);
    b3h_status st = b3h_set_model(c->calc, q, c->freqs, B3H_MODEL_SPECIAL_CASES);
    if (st) return st;
    for (int i = 0; i < 6; i++) c->device_rates[i] = c->rates[i];
  }
  // a parameter move changes every P(t); full_evaluations recomputes every
  // node for any move
  if (c->cfg.full_evaluations || !c->tree.changed()) c->tree.mark_all();
  c->tree.partials_lists(c->internals, c->kids);
  const size_t n = c->internals.size();
  c->times.resize(2 * n);
  for (size_t m = 0; m < 2 * n; m++) c->times[m] = c->tree.edge_length(c->kids[m]);
  b3h_status st = B3H_OK;
  if (c->alpha != c->cat_alpha) {
    if ((st = b3h_gamma_rates(c->alpha, c->K, c->cat_rates))) return st;
    c->cat_alpha = c->alpha;
  }
  st = b3h_propose_times(c->calc, (uint32_t)n, c->internals.data(), c->kids.data(),
                         c->times.data(), c->cat_rates);
  if (st) return st;
  return b3h_total(c->calc, out);
}

void features(const b3m_chain *c, float *out) {
  const Tree &t = c->tree;
  const uint32_t r0 = t.child(t.root, 0), r1 = t.child(t.root, 1);
  const uint32_t keep = std::min(r0, r1), drop = std::max(r0, r1);
  size_t k = 0;
  for (uint32_t v = 0; v < t.N; v++) {
    if (v == t.root || v == drop) continue;
    const double len = v == keep ? t.edge_length(r0) + t.edge_length(r1) : t.edge_length(v);
    out[k++] = (float)std::log(len);
  }
  for (int i = 0; i < 6; i++) out[k++] = (float)std::log(c->rates[i]);
  out[k++] = (float)std::log(c->alpha);
}

// Propose with operator op; returns the log Hastings ratio (-inf: abort).
double propose(b3m_chain *c, int op) {
  Tree &t = c->tree;
  Rng &rng = c->rng;
  switch (op) {
    case kNodeSlide: {  // _node_slide.py:26-39
      if (t.L < 3) return -INFINITY;
      uint32_t v;
      do v = rng.range(t.L, t.N);
      while (v == t.root);
      const double upper = t.heights[t.parents[v]];
      const double lower = std::max(t.heights[t.child(v, 0)], t.heights[t.child(v, 1)]);
      t.set_height(v, (upper - lower) * rng.uniform() + lower);
      return 0.0;
    }
    case kTreeScale: {  // _tree_scale.py:27-37, tree.rs:335-348
      const double f = c->tree_scale_f;
      const double s = f + (1.0 / f - f) * rng.uniform();
      for (uint32_t v = t.L; v < t.N; v++) t.set_height(v, t.heights[v] * s);
      return std::log(s) * ((double)(t.L - 1) - 2.0);
    }
    case kRates: {  // _delta_exchange.py:43-60
      const double delta = rng.uniform() * (1.0 - c->delta_t) / 10.0;
      const uint32_t d1 = rng.below(6);
      uint32_t d2 = rng.below(5);
      if (d2 >= d1) d2++;
      c->rates[d1] -= delta;
      c->rates[d2] += delta;
      return 0.0;
    }
    case kAlpha: {  // _param_scale.py:32-39
      const double f = c->alpha_f;
      const double s = f + (1.0 / f - f) * rng.uniform();
      c->alpha *= s;
      return -std::log(s);
    }
    case kNarrow: {  // _tree_exchange.py:23-57
      // b3's rule, or the reversible one (the grandparent keeps an internal
      // child, `parent`, so its own status never changes)
      const bool loose = c->cfg.narrow_reversible != 0;
      const uint32_t before = t.num_grandparents(loose);
      if (before == 0) return -INFINITY;
      uint32_t gp;
      do gp = rng.range(t.L, t.N);
      while (!t.is_grandparent(gp, loose));
      const uint32_t l = t.child(gp, 0), r = t.child(gp, 1);
      uint32_t parent, uncle;
      if (t.heights[l] > t.heights[r]) {
        parent = l;
        uncle = r;
      } else if (t.heights[r] > t.heights[l]) {
        parent = r;
        uncle = l;
      } else {
        return -INFINITY;
      }
      // b3's rule makes both children internal; under the loose one a dated
      // leaf can be the higher child, and BEAST2 aborts then
      if (parent < t.L) return -INFINITY;
      const uint32_t pre =
          (uint32_t)t.is_grandparent(parent, loose) + (uint32_t)t.is_grandparent(uncle, loose);
      const uint32_t ch = rng.uniform() < 0.5 ? t.child(parent, 0) : t.child(parent, 1);
      t.swap_parents(uncle, ch);
      const uint32_t post =
          (uint32_t)t.is_grandparent(parent, loose) + (uint32_t)t.is_grandparent(uncle, loose);
      const uint32_t after = before - pre + post;
      return std::log((double)before / (double)after);
    }
  }
  return -INFINITY;
}

void finalize(b3m_chain *c, bool accept) {  // mcmc.rs:362-393
  if (accept) {
    c->tree.accept();
    std::memcpy(c->rates_acc, c->rates, sizeof c->rates);
    c->alpha_acc = c->alpha;
    (void)b3h_accept(c->calc);
  } else {
    c->tree.reject();
    std::memcpy(c->rates, c->rates_acc, sizeof c->rates);
    c->alpha = c->alpha_acc;
    (void)b3h_reject(c->calc);
  }
}

// Scheduler::tune (operators/mod.rs:164-190,467-482) with LinearOptimizer's
// rate (optimizers.py:30-38)
void tune(b3m_chain *c) {
  const b3m_chain_config &g = c->cfg;
  const uint64_t step = c->life_steps;
  if (g.tune_every == 0 || step == 0 || step % g.tune_every != 0 || step >= g.tune_cutoff) return;
  const double ratio = (double)step / (double)g.tune_cutoff;
  const double rate = ratio < 0.9 ? g.tune_magnitude : g.tune_magnitude * 10.0 * (1.0 - ratio);
  double *param[5] = {nullptr, &c->tree_scale_f, &c->delta_t, &c->alpha_f, nullptr};
  for (int op = 0; op < kNumOps; op++) {
    const double n = (double)(c->tune_acc[op] + c->tune_rej[op]);
    if (!param[op] || n == 0.0) continue;
    const double acc = (double)c->tune_acc[op] / n;
    *param[op] = std::min(0.99, std::max(0.01, *param[op] + rate * (0.234 - acc)));
    c->tune_acc[op] = c->tune_rej[op] = 0;
  }
}

// b3's verbatim format (crates/verbatim/src/lib.rs): little-endian
// primitives, each padded to a multiple of its size from the stream start;
// tags unpadded; a slice is a u32 length, padding to the element alignment
// when non-empty, then the elements.
struct VWriter {
  std::vector<uint8_t> out;
  void pad(size_t a) {
    while (out.size() % a) out.push_back(0);
  }
  template <typename T>
  void prim(T v) {
    pad(sizeof(T));
    const uint8_t *b = reinterpret_cast<const uint8_t *>(&v);
    out.insert(out.end(), b, b + sizeof(T));
  }
  void tag(const uint8_t *t, size_t n) { out.insert(out.end(), t, t + n); }
  template <typename T>
  void slice(const T *v, uint32_t n) {
    prim<uint32_t>(n);
    if (n) pad(alignof(T));
    const uint8_t *b = reinterpret_cast<const uint8_t *>(v);
    out.insert(out.end(), b, b + (size_t)n * sizeof(T));
  }
};

struct VReader {
  const uint8_t *p;
  size_t len, pos = 0;
  bool ok = true;
  bool skip_pad(size_t a) {
    const size_t r = pos % a;
    if (r) pos += a - r;
    return pos <= len;
  }
  template <typename T>
  T prim() {
    T v{};
    if (!ok || !skip_pad(sizeof(T)) || pos + sizeof(T) > len) {
      ok = false;
      return v;
    }
    std::memcpy(&v, p + pos, sizeof(T));
    pos += sizeof(T);
    return v;
  }
  bool tag(const uint8_t *t, size_t n) {
    if (!ok || pos + n > len || std::memcmp(p + pos, t, n) != 0) return ok = false;
    pos += n;
    return true;
  }
  template <typename T>
  std::vector<T> slice() {
    const uint32_t n = prim<uint32_t>();
    std::vector<T> v;
    if (!ok) return v;
    if (n && !skip_pad(alignof(T))) {
      ok = false;
      return v;
    }
    if (pos + (size_t)n * sizeof(T) > len) {
      ok = false;
      return v;
    }
    v.resize(n);
    std::memcpy(v.data(), p + pos, (size_t)n * sizeof(T));
    pos += (size_t)n * sizeof(T);
    return v;
  }
};

// Mcmc's state header (mcmc.rs:19, lib.rs:29)
constexpr uint8_t kStateTag[4] = {0x62, 0x33, 0xb3, 0x44};
constexpr uint32_t kStateVersion = 19;

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace

extern "C" {

void b3m_chain_config_default(b3m_chain_config *cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof *cfg);
  cfg->device = 0;
  cfg->categories = 4;
  cfg->seed = 4;
  cfg->birth_rate = 1.0;
  cfg->alpha_prior_rate = 1.0;
  cfg->w_node_slide = 4.0;
  cfg->w_tree_scale = 1.0;
  cfg->w_rates = 2.0;
  cfg->w_alpha = 1.0;
  cfg->w_narrow = 2.0;
  cfg->scale_factor = 0.75;
  cfg->alpha_factor = 0.0;
  cfg->delta_tuning = 0.75;
  cfg->tune_every = 0;
  cfg->tune_cutoff = 0;
  cfg->tune_magnitude = 0.5;
  cfg->surrogate_beta = 1.0;
  cfg->narrow_reversible = 0;
}

b3h_status b3m_chain_create(const b3m_chain_config *cfg, const uint8_t *codes, uint32_t L,
                            uint32_t P, const uint32_t *weights, const uint32_t *parents,
                            const uint32_t *children, const double *heights,
                            const double *gtr_rates, const double *freqs, double alpha,
                            b3m_chain **out) {
  if (!cfg || !codes || !parents || !children || !heights || !gtr_rates || !freqs || !out)
    return fail(B3H_ERR_ARG, "null argument");
  if (L < 2 || P == 0) return fail(B3H_ERR_ARG, "need at least 2 taxa and 1 pattern");
  if (cfg->categories == 0 || cfg->categories > 64)
    return fail(B3H_ERR_ARG, "categories must be 1..64");
  if (!(cfg->birth_rate > 0.0) || !(cfg->alpha_prior_rate > 0.0) || !(alpha > 0.0) ||
      !(cfg->scale_factor > 0.0 && cfg->scale_factor < 1.0) ||
      !(cfg->alpha_factor >= 0.0 && cfg->alpha_factor < 1.0) ||
      !(cfg->delta_tuning >= 0.0 && cfg->delta_tuning < 1.0) ||
      !(cfg->surrogate_beta > 0.0 && std::isfinite(cfg->surrogate_beta)))
    return fail(B3H_ERR_ARG, "bad prior or operator setting");
  const double wsum =
      cfg->w_node_slide + cfg->w_tree_scale + cfg->w_rates + cfg->w_alpha + cfg->w_narrow;
  if (!(wsum > 0.0) || cfg->w_node_slide < 0 || cfg->w_tree_scale < 0 || cfg->w_rates < 0 ||
      cfg->w_alpha < 0 || cfg->w_narrow < 0)
    return fail(B3H_ERR_ARG, "operator weights must be non-negative with a positive sum");
  double rsum = 0.0;
  for (int i = 0; i < 6; i++) {
    if (!(gtr_rates[i] > 0.0)) return fail(B3H_ERR_ARG, "GTR rates must be positive");
    rsum += gtr_rates[i];
  }
  for (int i = 0; i < 4; i++)
    if (!(freqs[i] > 0.0)) return fail(B3H_ERR_ARG, "frequencies must be positive");
  b3m_chain *c = new b3m_chain();
  c->cfg = *cfg;
  c->L = L;
  c->P = P;
  c->K = cfg->categories;
  c->rng = Rng(cfg->seed);
  std::string err;
  if (!c->tree.init(L, parents, children, heights, &err)) {
    delete c;
    return fail(B3H_ERR_ARG, err);
  }
  for (int i = 0; i < 6; i++) c->rates[i] = c->rates_acc[i] = gtr_rates[i] / rsum;
  for (int i = 0; i < 4; i++) c->freqs[i] = freqs[i];
  c->alpha = c->alpha_acc = alpha;
  c->tree_scale_f = cfg->scale_factor;
  c->alpha_f = cfg->alpha_factor > 0.0 ? cfg->alpha_factor : cfg->scale_factor;
  c->delta_t = cfg->delta_tuning;
  c->feat.resize(2 * (size_t)L + 4);
  b3h_config hc;
  b3h_config_default(&hc);
  hc.device = cfg->device;
  b3h_status st = b3h_create(&hc, codes, L, P, c->K, &c->calc);
  if (st) {
    delete c;
    return st;
  }
  if (weights && (st = b3h_set_weights(c->calc, weights))) {
    b3m_chain_destroy(c);
    return st;
  }
  // initial full evaluation, accepted
  c->tree.mark_all();
  c->logprior = log_prior(c);
  if (!std::isfinite(c->logprior)) {
    b3m_chain_destroy(c);
    return fail(B3H_ERR_ARG, "the initial state has zero prior density");
  }
  if ((st = evaluate(c, &c->loglik))) {
    b3m_chain_destroy(c);
    return st;
  }
  c->posterior = c->loglik + c->logprior;
  finalize(c, true);
  *out = c;
  return B3H_OK;
}

b3h_status b3m_chain_destroy(b3m_chain *c) {
  if (!c) return B3H_OK;
  if (c->calc) b3h_destroy(c->calc);
  delete c;
  return B3H_OK;
}

uint32_t b3m_chain_num_features(const b3m_chain *c) { return c ? 2 * c->L + 4 : 0; }

b3h_status b3m_chain_features(const b3m_chain *c, float *out) {
  if (!c || !out) return fail(B3H_ERR_ARG, "null argument");
  features(c, out);
  return B3H_OK;
}

b3h_status b3m_chain_set_surrogate(b3m_chain *c, b3m_mlp *m) {
  if (!c) return fail(B3H_ERR_ARG, "null chain");
  c->mlp = m;
  c->surrogate_valid = false;
  return B3H_OK;
}

b3h_status b3m_chain_params(const b3m_chain *c, uint32_t *parents, uint32_t *children,
                            double *heights, double *gtr_rates, double *alpha) {
  if (!c) return fail(B3H_ERR_ARG, "null chain");
  if (parents) std::copy(c->tree.parents.begin(), c->tree.parents.end(), parents);
  if (children) std::copy(c->tree.children.begin(), c->tree.children.end(), children);
  if (heights) std::copy(c->tree.heights.begin(), c->tree.heights.end(), heights);
  if (gtr_rates) std::memcpy(gtr_rates, c->rates, sizeof c->rates);
  if (alpha) *alpha = c->alpha;
  return B3H_OK;
}

b3h_status b3m_chain_state(const b3m_chain *c, double *loglik, double *logprior,
                           double *posterior) {
  if (!c) return fail(B3H_ERR_ARG, "null chain");
  if (loglik) *loglik = c->loglik;
  if (logprior) *logprior = c->logprior;
  if (posterior) *posterior = c->posterior;
  return B3H_OK;
}

b3h_status b3m_chain_dump_state(const b3m_chain *c, uint8_t *buf, size_t cap, size_t *len) {
  if (!c || !len) return fail(B3H_ERR_ARG, "null argument");
  VWriter w;
  w.tag(kStateTag, 4);
  w.prim<uint32_t>(kStateVersion);
  w.prim<uint64_t>(c->life_steps);
  w.prim<double>(c->posterior);
  // parameters: the tree (tree.rs:1011-1016), the exchangeabilities as a
  // RealVector (real_vector.rs:77-82), alpha as a Real (real.rs:59-62)
  w.slice(c->tree.parents.data(), (uint32_t)c->tree.parents.size());
  w.slice(c->tree.heights.data(), (uint32_t)c->tree.heights.size());
  for (int i = 0; i < 6; i++) w.prim<double>(c->rates[i]);
  w.prim<double>(c->alpha);
  // scheduler (operators/mod.rs:195-217,488-507): accepts, rejects since the
  // last tune, and the tuning of operators that have one
  const double tuning[kNumOps] = {-1.0, c->tree_scale_f, c->delta_t, c->alpha_f, -1.0};
  for (int op = 0; op < kNumOps; op++) {
    w.prim<uint32_t>((uint32_t)c->tune_acc[op]);
    w.prim<uint32_t>((uint32_t)c->tune_rej[op]);
    if (tuning[op] >= 0.0) w.prim<double>(tuning[op]);
  }
  // the generator as two u128 words (rng/src/lib.rs:46-51 writes PCG64's
  // state and increment; this chain's xoshiro256** state fills the same
  // 32 bytes)
  w.prim<uint64_t>(c->rng.s[0]);
  w.prim<uint64_t>(c->rng.s[1]);
  w.prim<uint64_t>(c->rng.s[2]);
  w.prim<uint64_t>(c->rng.s[3]);
  *len = w.out.size();
  if (!buf) return B3H_OK;
  if (cap < w.out.size()) return fail(B3H_ERR_ARG, "state buffer too small");
  std::memcpy(buf, w.out.data(), w.out.size());
  return B3H_OK;
}

b3h_status b3m_chain_load_state(b3m_chain *c, const uint8_t *buf, size_t len) {
  if (!c || !buf) return fail(B3H_ERR_ARG, "null argument");
  VReader r{buf, len};
  if (!r.tag(kStateTag, 4)) return fail(B3H_ERR_ARG, "Not a b3 state file");
  const uint32_t version = r.prim<uint32_t>();
  if (r.ok && version != kStateVersion)
    return fail(B3H_ERR_ARG, "Cannot load state: protocol version v" + std::to_string(version) +
                                 " is incompatible with the supported version v" +
                                 std::to_string(kStateVersion));
  const uint64_t step = r.prim<uint64_t>();
  (void)r.prim<double>();  // the posterior is recomputed from the parameters
  std::vector<uint32_t> parents = r.slice<uint32_t>();
  std::vector<double> heights = r.slice<double>();
  double rates[6], alpha;
  for (double &x : rates) x = r.prim<double>();
  alpha = r.prim<double>();
  uint32_t acc[kNumOps], rej[kNumOps];
  double tuning[kNumOps] = {-1.0, 0.0, 0.0, 0.0, -1.0};
  for (int op = 0; op < kNumOps; op++) {
    acc[op] = r.prim<uint32_t>();
    rej[op] = r.prim<uint32_t>();
    if (tuning[op] >= 0.0) tuning[op] = r.prim<double>();
  }
  uint64_t rs[4];
  for (uint64_t &x : rs) x = r.prim<uint64_t>();
  if (!r.ok) return fail(B3H_ERR_ARG, "truncated or malformed state");
  const uint32_t N = 2 * c->L - 1;
  if (parents.size() != N || heights.size() != N)
    return fail(B3H_ERR_ARG, "the state's tree has a different number of nodes");
  // DatedTree::load_parents (tree.rs:906-930): children in node order
  std::vector<uint32_t> children(2 * (size_t)(c->L - 1), kTomb);
  for (uint32_t v = 0; v < N; v++) {
    const uint32_t p = parents[v];
    if (p == kTomb) continue;
    if (p < c->L || p >= N) return fail(B3H_ERR_ARG, "parent out of range in the state");
    uint32_t *slot = &children[2 * (size_t)(p - c->L)];
    if (slot[0] == kTomb) slot[0] = v;
    else if (slot[1] == kTomb) slot[1] = v;
    else return fail(B3H_ERR_ARG, "a node with more than two children in the state");
  }
  Tree t;
  std::string err;
  if (!t.init(c->L, parents.data(), children.data(), heights.data(), &err))
    return fail(B3H_ERR_ARG, err);
  for (double x : rates)
    if (!(x > 0.0)) return fail(B3H_ERR_ARG, "non-positive rate in the state");
  if (!(alpha > 0.0)) return fail(B3H_ERR_ARG, "non-positive alpha in the state");
  // commit, then evaluate from scratch and accept (mcmc.rs:196-213)
  c->tree = t;
  std::memcpy(c->rates, rates, sizeof rates);
  std::memcpy(c->rates_acc, rates, sizeof rates);
  c->alpha = c->alpha_acc = alpha;
  c->life_steps = step;
  for (int op = 0; op < kNumOps; op++) {
    c->tune_acc[op] = acc[op];
    c->tune_rej[op] = rej[op];
  }
  c->tree_scale_f = tuning[kTreeScale];
  c->delta_t = tuning[kRates];
  c->alpha_f = tuning[kAlpha];
  std::memcpy(c->rng.s, rs, sizeof rs);
  c->surrogate_valid = false;
  c->tree.mark_all();
  c->logprior = log_prior(c);
  b3h_status st = evaluate(c, &c->loglik);
  if (st) return st;
  c->posterior = c->loglik + c->logprior;
  finalize(c, true);
  return B3H_OK;
}

b3h_status b3m_chain_run(b3m_chain *c, uint64_t steps, uint32_t mode, uint32_t record_every,
                         double *trace, float *collect_x, float *collect_y, uint64_t collect_cap,
                         b3m_run_stats *stats) {
  if (!c) return fail(B3H_ERR_ARG, "null chain");
  if (mode != B3M_PLAIN && mode != B3M_DELAYED) return fail(B3H_ERR_ARG, "bad mode");
  if (mode == B3M_DELAYED && !c->mlp)
    return fail(B3H_ERR_STATE, "delayed acceptance needs a surrogate");
  if (record_every == 0) record_every = 1;
  b3m_run_stats s{};
  const double w[kNumOps] = {c->cfg.w_node_slide, c->cfg.w_tree_scale, c->cfg.w_rates,
                             c->cfg.w_alpha, c->cfg.w_narrow};
  double wsum = 0.0;
  for (double x : w) wsum += x;
  const uint32_t nf = b3m_chain_num_features(c);
  std::vector<float> fx(2 * (size_t)nf);
  float fy[2];
  const double t_start = now();
  b3h_status st = B3H_OK;

  auto surrogate = [&](double *out) -> b3h_status {
    const double t0 = now();
    features(c, fx.data());
    b3h_status e = b3h::mlp_predict(c->mlp, fx.data(), 1, fy);
    s.surrogate_seconds += now() - t0;
    s.surrogate_evals++;
    *out = fy[0];
    return e;
  };
  auto exact = [&](double *out) -> b3h_status {
    const double t0 = now();
    b3h_status e = evaluate(c, out);
    s.likelihood_seconds += now() - t0;
    s.exact_evals++;
    if (!e && collect_x && s.collected < collect_cap) {
      features(c, collect_x + s.collected * nf);
      collect_y[s.collected] = (float)*out;
      s.collected++;
    }
    return e;
  };

  if (mode == B3M_DELAYED && !c->surrogate_valid) {
    if ((st = surrogate(&c->surrogate_cur))) return st;
    c->surrogate_valid = true;
  }

  for (uint64_t step = 0; step < steps; step++) {
    // Scheduler::random_operator_index: weighted choice
    double u = c->rng.uniform() * wsum;
    int op = 0;
    while (op < kNumOps - 1 && u >= w[op]) u -= w[op++];
    while (w[op] == 0.0) op = (op + 1) % kNumOps;
    s.op_proposed[op]++;
    bool accept = false;
    const double hastings = propose(c, op);
    if (hastings == -INFINITY) {
      s.aborts++;
    } else {
      const double prior = log_prior(c);
      if (prior == -INFINITY) {
        s.prior_rejects++;
      } else if (mode == B3M_DELAYED && (op != kNarrow || c->cfg.screen_topology)) {
        double fnew = 0.0;
        if ((st = surrogate(&fnew))) {
          finalize(c, false);  // an error ends the run with the proposal rejected
          break;
        }
        const double beta = c->cfg.surrogate_beta;
        if (!b3m_da_stage1(beta, c->surrogate_cur, fnew, prior - c->logprior, hastings,
                           c->rng.uniform())) {
          s.stage1_rejects++;
        } else {
          double ll = 0.0;
          if ((st = exact(&ll))) {
            finalize(c, false);
            break;
          }
          // the uniform is drawn only for a finite ln L, as it always was
          if (ll != -INFINITY &&
              b3m_da_stage2(beta, c->surrogate_cur, fnew, ll - c->loglik, c->rng.uniform())) {
            accept = true;
            c->loglik = ll;
            c->logprior = prior;
            c->posterior = ll + prior;
            c->surrogate_cur = fnew;
          } else {
            s.stage2_rejects++;
          }
        }
      } else {
        double ll = 0.0;
        if ((st = exact(&ll))) {
          finalize(c, false);
          break;
        }
        if (ll != -INFINITY) {
          const double ratio = ll + prior - c->posterior + hastings;
          if (ratio > std::log(c->rng.uniform())) {  // mcmc.rs:346-359
            accept = true;
            c->loglik = ll;
            c->logprior = prior;
            c->posterior = ll + prior;
          }
        }
      }
    }
    if (accept && mode == B3M_DELAYED && op == kNarrow && !c->cfg.screen_topology) {
      if ((st = surrogate(&c->surrogate_cur))) {
        finalize(c, true);  // the move was accepted; only its surrogate value failed
        c->surrogate_valid = false;
        break;
      }
    }
    finalize(c, accept);
    if (accept) {
      s.accepted++;
      s.op_accepted[op]++;
      c->tune_acc[op]++;
    } else {
      c->tune_rej[op]++;
    }
    s.steps++;
    c->life_steps++;
    tune(c);
    if (trace && step % record_every == 0) {
      double *row = trace + (step / record_every) * B3M_TRACE_COLS;
      row[B3M_COL_STEP] = (double)step;
      row[B3M_COL_POSTERIOR] = c->posterior;
      row[B3M_COL_LIKELIHOOD] = c->loglik;
      row[B3M_COL_PRIOR] = c->logprior;
      row[B3M_COL_ROOT_HEIGHT] = c->tree.heights[c->tree.root];
      row[B3M_COL_TREE_LENGTH] = c->tree.tree_length();
      row[B3M_COL_ALPHA] = c->alpha;
      for (int i = 0; i < 6; i++) row[B3M_COL_RATE0 + i] = c->rates[i];
    }
  }
  // a plain run moves the state without the surrogate: its value is stale
  if (mode != B3M_DELAYED) c->surrogate_valid = false;
  s.seconds = now() - t_start;
  s.tuning[kNodeSlide] = s.tuning[kNarrow] = -1.0;
  s.tuning[kTreeScale] = c->tree_scale_f;
  s.tuning[kRates] = c->delta_t;
  s.tuning[kAlpha] = c->alpha_f;
  if (stats) *stats = s;
  return st;
}

int b3m_da_stage1(double beta, double f_cur, double f_new, double d_prior, double hastings,
                  double u1) {
  const double r1 = beta * (f_new - f_cur) + d_prior + hastings;
  return r1 > std::log(u1) ? 1 : 0;
}

int b3m_da_stage2(double beta, double f_cur, double f_new, double d_loglik, double u2) {
  if (d_loglik == -INFINITY || std::isnan(d_loglik)) return 0;
  return d_loglik - beta * (f_new - f_cur) > std::log(u2) ? 1 : 0;
}

}  // extern "C"
