/*
 * b3_oracle.c -- FP64 CPU restatement of b3's Felsenstein-pruning likelihood.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product.  Compiled with -ffp-contract=off so that no multiply-add is fused,
 * which matches Rust's default floating-point semantics in the reference.
 *
 * Restated reference functions and how each is pinned (tests/test_oracle_*.py):
 *   EditBuf                     crates/sk/src/editbuf.rs:48-92
 *       pinned: bit table of the reference's own comments (00->11 ...),
 *       model test against a two-slot Python model.
 *   DatedTree schedule          crates/b3/src/parameters/tree.rs:357-461,471-528
 *       pinned: brute force (node set == ancestors of marked edges),
 *       children-before-parents, root last, sorted pairs.
 *   JC / K80 closed forms       crates/b3/src/substitution.rs:131-148,208-232
 *       pinned: scipy.linalg.expm of the model's Q.
 *   HKY analytic eigensystem    substitution.rs:336-420
 *       pinned: scipy expm, special cases t>20 and t<1e-10 by value.
 *   GTR                         substitution.rs:502-588
 *       The reference decomposes Q with a Hessenberg/QR eigensolver
 *       (beast_eigen.rs:66-101); this restatement uses cyclic Jacobi on the
 *       symmetrised matrix.  Eigenvectors are not unique, P(t) is: pinned by
 *       scipy expm (1e-13), V diag V^-1 == Q, Chapman-Kolmogorov, rows sum 1.
 *   Discrete gamma              crates/b3/src/likelihood/gamma.rs:205-225
 *       The reference calls computare's Gamma::inverse_cdf (git crate, commit
 *       02193b66, not vendored).  Restated from the published series /
 *       continued-fraction expansions of the incomplete gamma function plus a
 *       safeguarded Newton inversion; pinned by scipy.stats.gamma.ppf golden
 *       vectors (tests/golden/gamma_ppf.json) to 1e-12.
 *   Cpu4Calculator              crates/b3/src/calculator/cpu4.rs:39-441
 *       pinned: brute-force sum over ancestral states on <=7 taxa, 2-taxon
 *       closed form, scaling invariance (scale_ln 3/30/300), partial update ==
 *       full recompute, reject restores, all-unknown column gives 0.
 *   S-state calculator          cpu4.rs semantics for S states (the reference
 *       n-state StateCalculator is todo!() at calculator/state.rs:35); pinned:
 *       equals calc4 at S=4 and brute force at S=61 on 3 taxa.
 *   weighted_sum/log_sum_exp    crates/b3/src/likelihood/mod.rs:134-151
 *       pinned: math.fsum references.
 *   Gamma mixing                gamma.rs:154-166 - pinned: brute force.
 *   deduplicate                 likelihood/mod.rs:74-132
 *       pinned: first-occurrence order, counts, round trip.
 *   Components from scratch     likelihood/hetero.rs:122-160, gamma.rs:126-166
 *       with per-edge clock rates (clock.rs:115-119, transitions.rs:33-35);
 *       pinned: brute force with scipy expm, strict-clock and one-category
 *       reductions.
 *   Propagation calculator      calculator/cuda_propagations/kernels.cu:73-187,
 *       mod.rs:77-164 (b3's default GPU backend) restated on the CPU: stored
 *       per-node projections P x, per-node-slot flags with scaling of any node
 *       (cherries included), the root's flag handling; pinned: equals Cpu4 on
 *       full and partial updates (1e-12), brute force, scaling invariance.
 */
#define _GNU_SOURCE
#include "b3_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* ======================================================================== */
/* EditBuf: 2 bits per slot, bit0 = active slot, bit1 = edited this epoch.  */
/* editbuf.rs:191-201: 00->11, 01->10, 10->10, 11->11                        */
void or_editbuf_set_edited(uint8_t *m, size_t index) {
  uint8_t v = m[index];
  m[index] = (uint8_t)(((v & 1u) ^ (uint8_t)~(v >> 1)) & 3u);
}
/* editbuf.rs:48-53 */
void or_editbuf_accept(uint8_t *m, size_t len) {
  for (size_t i = 0; i < len; i++) m[i] &= 1u;
}
/* editbuf.rs:55-66: 00->00 01->01 10->01 11->00 */
void or_editbuf_reject(uint8_t *m, size_t len) {
  for (size_t i = 0; i < len; i++) m[i] = (uint8_t)((m[i] ^ (m[i] >> 1)) & 1u);
}

/* ======================================================================== */
/* DatedTree                                                                */
struct or_tree {
  uint32_t L, N;
  uint32_t root, root_bak;
  uint32_t *children, *children_bak; /* 2*(L-1) */
  uint32_t *parents, *parents_bak;   /* N */
  double *heights, *heights_bak;     /* N */
  uint8_t *upd_edges, *upd_partials, *upd_props;
};

or_tree *or_tree_new(uint32_t L, const uint32_t *parents,
                     const uint32_t *children, const double *heights) {
  if (L < 2) return NULL;
  or_tree *t = calloc(1, sizeof *t);
  t->L = L;
  t->N = 2 * L - 1;
  size_t nc = 2 * (size_t)(L - 1);
  t->children = malloc(nc * 4);
  t->children_bak = malloc(nc * 4);
  t->parents = malloc(t->N * 4);
  t->parents_bak = malloc(t->N * 4);
  t->heights = malloc(t->N * 8);
  t->heights_bak = malloc(t->N * 8);
  t->upd_edges = calloc(t->N, 1);
  t->upd_partials = calloc(t->N, 1);
  t->upd_props = calloc(t->N, 1);
  memcpy(t->children, children, nc * 4);
  memcpy(t->parents, parents, t->N * 4);
  memcpy(t->heights, heights, t->N * 8);
  t->root = OR_TOMBSTONE;
  for (uint32_t n = 0; n < t->N; n++)
    if (parents[n] == OR_TOMBSTONE) t->root = n;
  memcpy(t->children_bak, t->children, nc * 4);
  memcpy(t->parents_bak, t->parents, t->N * 4);
  memcpy(t->heights_bak, t->heights, t->N * 8);
  t->root_bak = t->root;
  return t;
}

void or_tree_free(or_tree *t) {
  if (!t) return;
  free(t->children); free(t->children_bak);
  free(t->parents); free(t->parents_bak);
  free(t->heights); free(t->heights_bak);
  free(t->upd_edges); free(t->upd_partials); free(t->upd_props);
  free(t);
}

uint32_t or_tree_root(const or_tree *t) { return t->root; }

static void children_of(const or_tree *t, uint32_t node, uint32_t *l,
                        uint32_t *r) {
  size_t off = (size_t)(node - t->L) * 2; /* tree.rs:701-707 */
  *l = t->children[off];
  *r = t->children[off + 1];
}

/* tree.rs:357-375 */
void or_tree_mark_edge_updated(or_tree *t, uint32_t edge) {
  t->upd_edges[edge] = 1;
  uint32_t child = edge; /* edge_nodes: edge index == child id, tree.rs:738-743 */
  t->upd_props[child] = 1;
  uint32_t curr = child;
  while (t->parents[curr] != OR_TOMBSTONE) {
    curr = t->parents[curr];
    if (t->upd_partials[curr]) break;
    t->upd_props[curr] = 1;
    t->upd_partials[curr] = 1;
  }
}

/* tree.rs:382-386 */
void or_tree_mark_all_edges_updated(or_tree *t) {
  memset(t->upd_edges, 1, t->N);
  memset(t->upd_partials, 1, t->N);
  memset(t->upd_props, 1, t->N);
}

/* tree.rs:388-396 with edges() = every node but the root, tree.rs:833-836 */
uint32_t or_tree_edges_to_update(const or_tree *t, uint32_t *edges) {
  uint32_t k = 0;
  for (uint32_t e = 0; e < t->N; e++) {
    if (e == t->root) continue;
    if (t->upd_edges[e]) edges[k++] = e;
  }
  return k;
}

/* tree.rs:398-421: BFS from the root through marked internal children,
 * then reversed so that children precede parents and the root is last. */
static uint32_t updated_internals(const or_tree *t, const uint8_t *bitmap,
                                  uint32_t *out) {
  uint32_t *queue = malloc(t->N * 4);
  uint32_t qh = 0, qt = 0, k = 0;
  out[k++] = t->root;
  queue[qt++] = t->root;
  while (qh < qt) {
    uint32_t node = queue[qh++], l, r;
    children_of(t, node, &l, &r);
    if (bitmap[l] && l >= t->L) { out[k++] = l; queue[qt++] = l; }
    if (bitmap[r] && r >= t->L) { out[k++] = r; queue[qt++] = r; }
  }
  free(queue);
  for (uint32_t i = 0; i < k / 2; i++) {
    uint32_t tmp = out[i];
    out[i] = out[k - 1 - i];
    out[k - 1 - i] = tmp;
  }
  return k;
}

/* tree.rs:423-434: each pair sorted ascending */
static void children_list(const or_tree *t, const uint32_t *internals,
                          uint32_t n, uint32_t *children) {
  for (uint32_t i = 0; i < n; i++) {
    uint32_t l, r;
    children_of(t, internals[i], &l, &r);
    if (l < r) { children[2 * i] = l; children[2 * i + 1] = r; }
    else { children[2 * i] = r; children[2 * i + 1] = l; }
  }
}

/* tree.rs:456-461 */
uint32_t or_tree_partials_lists(const or_tree *t, uint32_t *internals,
                                uint32_t *children) {
  uint32_t n = updated_internals(t, t->upd_partials, internals);
  children_list(t, internals, n, children);
  return n;
}

/* tree.rs:436-454 */
uint32_t or_tree_propagation_lists(const or_tree *t, uint32_t *nodes,
                                   uint32_t *children,
                                   uint32_t *num_leaves_out) {
  uint32_t k = 0;
  for (uint32_t leaf = 0; leaf < t->L; leaf++)
    if (t->upd_props[leaf]) nodes[k++] = leaf;
  *num_leaves_out = k;
  uint32_t n = updated_internals(t, t->upd_props, nodes + k);
  children_list(t, nodes + k, n, children);
  return k + n;
}

/* tree.rs:516-528 */
void or_tree_set_height(or_tree *t, uint32_t node, double h) {
  t->heights[node] = h;
  if (t->parents[node] != OR_TOMBSTONE) or_tree_mark_edge_updated(t, node);
  if (node >= t->L) {
    uint32_t l, r;
    children_of(t, node, &l, &r);
    or_tree_mark_edge_updated(t, l);
    or_tree_mark_edge_updated(t, r);
  }
}

/* tree.rs:471-495 */
int or_tree_replace_child(or_tree *t, uint32_t parent, uint32_t old_child,
                          uint32_t new_child) {
  t->parents[new_child] = parent;
  uint32_t l, r;
  children_of(t, parent, &l, &r);
  size_t idx = (size_t)(parent - t->L) * 2;
  if (old_child == l) t->children[idx] = new_child;
  else if (old_child == r) t->children[idx + 1] = new_child;
  else return -1;
  or_tree_mark_edge_updated(t, new_child);
  return 0;
}

/* tree.rs:732-736 */
double or_tree_edge_length(const or_tree *t, uint32_t edge) {
  return t->heights[t->parents[edge]] - t->heights[edge];
}

/* tree.rs:976-978 */
int or_tree_is_changed(const or_tree *t) {
  for (uint32_t e = 0; e < t->N; e++)
    if (t->upd_edges[e]) return 1;
  return 0;
}

static void clear_updated(or_tree *t) { /* tree.rs:350-354 */
  memset(t->upd_edges, 0, t->N);
  memset(t->upd_partials, 0, t->N);
  memset(t->upd_props, 0, t->N);
}

/* tree.rs:980-995 (EpochBuf accept/reject == keep / restore the values) */
void or_tree_accept(or_tree *t) {
  size_t nc = 2 * (size_t)(t->L - 1);
  memcpy(t->children_bak, t->children, nc * 4);
  memcpy(t->parents_bak, t->parents, t->N * 4);
  memcpy(t->heights_bak, t->heights, t->N * 8);
  clear_updated(t);
  t->root_bak = t->root;
}

void or_tree_reject(or_tree *t) {
  size_t nc = 2 * (size_t)(t->L - 1);
  memcpy(t->children, t->children_bak, nc * 4);
  memcpy(t->parents, t->parents_bak, t->N * 4);
  memcpy(t->heights, t->heights_bak, t->N * 8);
  clear_updated(t);
  t->root = t->root_bak;
}

void or_tree_get(const or_tree *t, uint32_t *parents, uint32_t *children,
                 double *heights) {
  if (parents) memcpy(parents, t->parents, t->N * 4);
  if (children) memcpy(children, t->children, 2 * (size_t)(t->L - 1) * 4);
  if (heights) memcpy(heights, t->heights, t->N * 8);
}

/* ======================================================================== */
/* Substitution models                                                      */

/* Cyclic Jacobi eigenvalue method for a real symmetric matrix (Golub & Van
 * Loan, Matrix Computations, sec. 8.5).  A = U diag(w) U^T, U row-major with
 * eigenvectors in columns. */
int or_sym_eigen(int n, const double *A, double *w, double *U) {
  double *a = malloc((size_t)n * n * sizeof(double));
  memcpy(a, A, (size_t)n * n * sizeof(double));
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) U[i * n + j] = (i == j) ? 1.0 : 0.0;
  int sweep;
  for (sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        double v = a[i * n + j] * a[i * n + j];
        tot += v;
        if (i != j) off += v;
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    int rotated = 0;
    for (int p = 0; p < n - 1; p++) {
      for (int q = p + 1; q < n; q++) {
        double apq = a[p * n + q];
        double app = a[p * n + p], aqq = a[q * n + q];
        /* skip rotations that cannot change the diagonal at double precision */
        if (fabs(apq) <= 1e-18 * (fabs(app) + fabs(aqq)) || apq == 0.0) continue;
        rotated = 1;
        double theta = (aqq - app) / (2.0 * apq);
        double tt = (theta >= 0 ? 1.0 : -1.0) /
                    (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(tt * tt + 1.0), s = tt * c;
        for (int k = 0; k < n; k++) { /* A <- A J (columns p, q) */
          double akp = a[k * n + p], akq = a[k * n + q];
          a[k * n + p] = c * akp - s * akq;
          a[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) { /* A <- J^T A (rows p, q) */
          double apk = a[p * n + k], aqk = a[q * n + k];
          a[p * n + k] = c * apk - s * aqk;
          a[q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) { /* U <- U J */
          double ukp = U[k * n + p], ukq = U[k * n + q];
          U[k * n + p] = c * ukp - s * ukq;
          U[k * n + q] = s * ukp + c * ukq;
        }
      }
    }
    if (!rotated) break;
  }
  for (int i = 0; i < n; i++) w[i] = a[i * n + i];
  free(a);
  return sweep < 100 ? 0 : -1;
}

/* Eigensystem of a reversible Q with stationary pi:
 * S = D^1/2 Q D^-1/2 is symmetric; Q = (D^-1/2 U) diag(w) (U^T D^1/2). */
static int reversible_eigen(or_model *m) {
  int n = m->n;
  double *S = calloc((size_t)n * n, 8), *U = calloc((size_t)n * n, 8);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      double sij = m->q[i * n + j] * sqrt(m->freqs[i]) / sqrt(m->freqs[j]);
      double sji = m->q[j * n + i] * sqrt(m->freqs[j]) / sqrt(m->freqs[i]);
      S[i * n + j] = 0.5 * (sij + sji);
    }
  int rc = or_sym_eigen(n, S, m->eval, U);
  for (int i = 0; i < n; i++)
    for (int k = 0; k < n; k++) {
      m->evec[i * n + k] = U[i * n + k] / sqrt(m->freqs[i]);
      m->ievc[k * n + i] = U[i * n + k] * sqrt(m->freqs[i]);
    }
  free(S);
  free(U);
  return rc;
}

void or_model_jc(or_model *m) {
  memset(m, 0, sizeof *m);
  m->kind = 0;
  m->n = 4;
  for (int i = 0; i < 4; i++) m->freqs[i] = 0.25;
}

void or_model_k80(or_model *m, double kappa) {
  memset(m, 0, sizeof *m);
  m->kind = 1;
  m->n = 4;
  m->kappa = kappa;
  for (int i = 0; i < 4; i++) m->freqs[i] = 0.25;
}

/* substitution.rs:336-388 */
void or_model_hky(or_model *m, double kappa, const double *f) {
  memset(m, 0, sizeof *m);
  m->kind = 2;
  m->n = 4;
  m->kappa = kappa;
  for (int i = 0; i < 4; i++) m->freqs[i] = f[i];
  double a = f[0], c = f[1], g = f[2], t = f[3];
  double r = a + g, y = c + t;
  double q[4][4] = {{0.0, c, kappa * g, t},
                    {a, 0.0, g, kappa * t},
                    {kappa * a, c, 0.0, t},
                    {a, kappa * c, g, 0.0}};
  double div = 2.0 * (g * t + a * c + a * t + c * g + kappa * (a * g + c * t));
  for (int i = 0; i < 4; i++) {
    /* Rust's iter().sum() folds from 0.0 left to right (diagonal is 0 here) */
    double s = 0.0;
    for (int j = 0; j < 4; j++) s += q[i][j];
    q[i][i] = -s;
    for (int j = 0; j < 4; j++) q[i][j] /= div;
  }
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) m->q[i * 4 + j] = q[i][j];
  double p[16] = {1.0, -y / r, -g / a, 0.0, 1.0, 1.0, 0.0, -t / c,
                  1.0, -y / r, 1.0, 0.0, 1.0, 1.0, 0.0, 1.0};
  double ip[16] = {a, c, g, t, -a, c * r / y, -g, t * r / y,
                   -a / r, 0.0, a / r, 0.0, 0.0, -c / y, 0.0, c / y};
  memcpy(m->evec, p, sizeof p);
  memcpy(m->ievc, ip, sizeof ip);
  m->eval[0] = 0.0;
  m->eval[1] = -1.0 / div;
  m->eval[2] = -(y + r * kappa) / div;
  m->eval[3] = -(r + y * kappa) / div;
}

/* substitution.rs:502-555 */
void or_model_gtr(or_model *m, const double *rates, const double *f) {
  memset(m, 0, sizeof *m);
  m->kind = 3;
  m->n = 4;
  for (int i = 0; i < 4; i++) m->freqs[i] = f[i];
  double pa = f[0], pc = f[1], pg = f[2], pt = f[3];
  double a = rates[0], b = rates[1], c = rates[2], d = rates[3], e = rates[4],
         ff = rates[5];
  double g[16] = {-a * pc - b * pg - c * pt, a * pc, b * pg, c * pt,
                  a * pa, -a * pa - d * pg - e * pt, d * pg, e * pt,
                  b * pa, d * pc, -b * pa - d * pc - ff * pt, ff * pt,
                  c * pa, e * pc, ff * pg, -c * pa - e * pc - ff * pg};
  double div = 2.0 * (a * pa * pc + b * pa * pg + c * pa * pt + d * pc * pg +
                      e * pc * pt + ff * pg * pt);
  for (int i = 0; i < 16; i++) m->q[i] = g[i] / div;
  reversible_eigen(m);
}

int or_model_reversible(or_model *m, int n, const double *R, const double *pi) {
  if (n < 2 || n > 64) return -1;
  memset(m, 0, sizeof *m);
  m->kind = 4;
  m->n = n;
  for (int i = 0; i < n; i++) m->freqs[i] = pi[i];
  double mean = 0.0;
  for (int i = 0; i < n; i++) {
    double s = 0.0;
    for (int j = 0; j < n; j++) {
      if (j == i) continue;
      m->q[i * n + j] = R[i * n + j] * pi[j];
      s += m->q[i * n + j];
    }
    m->q[i * n + i] = -s;
    mean += pi[i] * s;
  }
  for (int i = 0; i < n * n; i++) m->q[i] /= mean;
  return reversible_eigen(m);
}

/* mul3 (substitution.rs:79-100): tmp = V diag; dst = tmp V^-1, each product
 * a plain row-by-column sum in index order. */
static void eigen_transition(const or_model *m, double t, double *dst) {
  int n = m->n;
  double ex[64];
  for (int k = 0; k < n; k++) ex[k] = exp(m->eval[k] * t);
  double *tmp = malloc((size_t)n * n * 8);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      double s = 0.0;
      for (int k = 0; k < n; k++) s += m->evec[i * n + k] * (k == j ? ex[j] : 0.0);
      tmp[i * n + j] = s;
    }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      double s = 0.0;
      for (int k = 0; k < n; k++) s += tmp[i * n + k] * m->ievc[k * n + j];
      dst[i * n + j] = s;
    }
  free(tmp);
}

void or_model_transition(const or_model *m, double distance, double *dst) {
  if (m->kind == 0) { /* substitution.rs:131-148 */
    double ex = exp(-4.0 / 3.0 * distance);
    double diag = 0.25 + 0.75 * ex, other = 0.25 - 0.25 * ex;
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) dst[i * 4 + j] = (i == j) ? diag : other;
    return;
  }
  if (m->kind == 1) { /* substitution.rs:208-232 */
    double kappa = m->kappa;
    double frac1 = -4.0 / (kappa + 2.0);
    double frac2 = -(2.0 * kappa + 2.0) / (kappa + 2.0);
    double arg1 = 0.25 * exp(distance * frac1);
    double arg2 = 0.5 * exp(distance * frac2);
    double dg = 0.25 + arg1 + arg2, ts = 0.25 + arg1 - arg2, tv = 0.25 - arg1;
    double out[16] = {dg, tv, ts, tv, tv, dg, tv, ts,
                      ts, tv, dg, tv, tv, ts, tv, dg};
    memcpy(dst, out, sizeof out);
    return;
  }
  if (m->kind == 2 || m->kind == 3) { /* substitution.rs:403-420,571-588 */
    if (distance > 20.0) { /* transition_freqs, :246-252 */
      for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) dst[i * 4 + j] = m->freqs[j];
      return;
    } else if (distance < 1e-10) { /* transition_q, :259-268 */
      for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) dst[i * 4 + j] = m->q[i * 4 + j] * distance;
      for (int i = 0; i < 4; i++) dst[i * 4 + i] = 1.0;
      return;
    }
  }
  eigen_transition(m, distance, dst);
}

void or_model_frequencies(const or_model *m, double *dst) {
  for (int i = 0; i < m->n; i++) dst[i] = m->freqs[i];
}

/* ======================================================================== */
/* Discrete gamma                                                           */

/* Regularised lower incomplete gamma P(a, x): power series for x < a + 1,
 * Lentz's continued fraction for Q otherwise (Press et al., Numerical
 * Recipes, 2nd ed., sec. 6.2). */
double or_gamma_p(double a, double x) {
  if (x <= 0.0) return 0.0;
  double gln = lgamma(a);
  if (x < a + 1.0) {
    double ap = a, sum = 1.0 / a, del = sum;
    for (int n = 0; n < 10000; n++) {
      ap += 1.0;
      del *= x / ap;
      sum += del;
      if (fabs(del) < fabs(sum) * 1e-17) break;
    }
    return sum * exp(-x + a * log(x) - gln);
  }
  const double tiny = 1e-300;
  double b = x + 1.0 - a, c = 1.0 / tiny, d = 1.0 / b, h = d;
  for (int i = 1; i < 10000; i++) {
    double an = -i * (i - a);
    b += 2.0;
    d = an * d + b;
    if (fabs(d) < tiny) d = tiny;
    c = b + an / c;
    if (fabs(c) < tiny) c = tiny;
    d = 1.0 / d;
    double del = d * c;
    h *= del;
    if (fabs(del - 1.0) < 1e-17) break;
  }
  return 1.0 - exp(-x + a * log(x) - gln) * h;
}

/* Quantile of Gamma(shape, scale 1).  Newton iteration on u = ln x for
 * P(a, e^u) - p (derivative x^a e^-x / Gamma(a)), kept inside a bisection
 * bracket; working in ln x resolves the tiny quantiles of small shapes. */
double or_gamma_inv_cdf(double a, double p) {
  if (p <= 0.0) return 0.0;
  if (p >= 1.0) return INFINITY;
  double ulo = -745.0, uhi = log(fmax(1.0, a));
  while (or_gamma_p(a, exp(uhi)) < p) { ulo = uhi; uhi += 1.0; }
  double gln = lgamma(a);
  /* start: leading term P(a, x) ~ x^a / Gamma(a + 1) for small x */
  double u = (log(p) + lgamma(a + 1.0)) / a;
  if (!(u > ulo && u < uhi)) u = 0.5 * (ulo + uhi);
  for (int it = 0; it < 400; it++) {
    double x = exp(u);
    double f = or_gamma_p(a, x) - p;
    if (f < 0) ulo = u; else uhi = u;
    double dfdu = exp(a * u - x - gln);
    double un = (dfdu > 0 && isfinite(dfdu)) ? u - f / dfdu : 0.5 * (ulo + uhi);
    if (!(un > ulo && un < uhi)) un = 0.5 * (ulo + uhi);
    if (fabs(un - u) <= 1e-15 * fmax(1.0, fabs(u))) { u = un; break; }
    u = un;
  }
  return exp(u);
}

/* gamma.rs:205-225: quantile midpoints (2i+1)/2K, renormalised to mean 1 */
void or_gamma_categories(double alpha, int k, double *rates) {
  double sum = 0.0;
  for (int i = 0; i < k; i++) {
    double p = (2.0 * i + 1.0) / (2.0 * k);
    rates[i] = or_gamma_inv_cdf(alpha, p);
    sum += rates[i];
  }
  double mean = sum / k;
  for (int i = 0; i < k; i++) rates[i] = rates[i] / mean;
}

/* ======================================================================== */
/* Likelihood helpers                                                       */

/* mod.rs:134-139 (serial sum in pattern order) */
double or_weighted_sum(const double *v, const uint32_t *w, size_t n) {
  double s = 0.0;
  for (size_t i = 0; i < n; i++) s += v[i] * (double)w[i];
  return s;
}

/* mod.rs:141-151 */
double or_log_sum_exp(const double *v, size_t n) {
  double mx = -INFINITY;
  for (size_t i = 0; i < n; i++) mx = fmax(mx, v[i]);
  if (mx == -INFINITY) return -INFINITY;
  double s = 0.0;
  for (size_t i = 0; i < n; i++) s += exp(v[i] - mx);
  return mx + log(s);
}

/* gamma.rs:154-166 */
double or_gamma_mix(int k, size_t p, const double *const *site, const uint32_t *w,
                    double *mixed) {
  double lik = 0.0;
  double div = log(1.0 / (double)k);
  double sums[64];
  for (size_t i = 0; i < p; i++) {
    for (int c = 0; c < k; c++) sums[c] = site[c][i];
    double s = or_log_sum_exp(sums, (size_t)k) + div;
    if (mixed) mixed[i] = s;
    lik += s * (double)w[i];
  }
  return lik;
}

/* ======================================================================== */
/* Thread helper: split [0, n) into contiguous slices, one per thread,
 * like fork_union's for_slices used at cpu4.rs:202. */
typedef void (*slice_fn)(void *ctx, size_t start, size_t count);
typedef struct { slice_fn fn; void *ctx; size_t start, count; } slice_job;

static void *slice_trampoline(void *arg) {
  slice_job *j = arg;
  if (j->count) j->fn(j->ctx, j->start, j->count);
  return NULL;
}

static void for_slices(int threads, size_t n, slice_fn fn, void *ctx) {
  if (threads <= 1 || n < 2) { fn(ctx, 0, n); return; }
  if ((size_t)threads > n) threads = (int)n;
  pthread_t tid[256];
  slice_job jobs[256];
  if (threads > 256) threads = 256;
  size_t base = n / threads, rem = n % threads, s = 0;
  for (int t = 0; t < threads; t++) {
    size_t cnt = base + ((size_t)t < rem ? 1 : 0);
    jobs[t] = (slice_job){fn, ctx, s, cnt};
    s += cnt;
  }
  for (int t = 1; t < threads; t++)
    pthread_create(&tid[t], NULL, slice_trampoline, &jobs[t]);
  slice_trampoline(&jobs[0]);
  for (int t = 1; t < threads; t++) pthread_join(tid[t], NULL);
}

/* ======================================================================== */
/* Cpu4Calculator (cpu4.rs)                                                 */
struct or_calc4 {
  uint32_t L, I;
  size_t P;
  uint8_t *sel;     /* EditBuf over internals */
  uint8_t *samples; /* [L x P], remapped to 0..4 */
  double (*partials)[4];
  double *likelihoods;
  uint8_t *scales;
  uint32_t *scale_sums, *scale_sums_backup;
  uint32_t scale_ln;
  double thr, mult;
  int threads;
};

static long host_cores(void) {
  long n = sysconf(_SC_NPROCESSORS_ONLN);
  return n > 0 ? n : 1;
}

/* cpu4.rs:108-165 */
or_calc4 *or_calc4_new(size_t P, const uint8_t *samples, uint32_t L,
                       uint32_t scale_ln, int num_threads) {
  or_calc4 *c = calloc(1, sizeof *c);
  c->L = L;
  c->I = L - 1;
  c->P = P;
  c->sel = calloc(c->I, 1);
  c->samples = malloc((size_t)L * P);
  for (size_t i = 0; i < (size_t)L * P; i++) { /* cpu4.rs:117-125 */
    switch (samples[i]) {
      case 1: c->samples[i] = 0; break;
      case 2: c->samples[i] = 1; break;
      case 4: c->samples[i] = 2; break;
      case 8: c->samples[i] = 3; break;
      default: c->samples[i] = 4;
    }
  }
  if (num_threads == 0) { /* cpu4.rs:127-132 */
    long cores = host_cores();
    size_t want = (P + 999) / 1000;
    num_threads = (int)((size_t)cores < want ? (size_t)cores : want);
  }
  c->threads = num_threads < 1 ? 1 : num_threads;
  size_t slots = (size_t)c->I * P * 2;
  c->partials = aligned_alloc(32, ((slots * 32) + 31) / 32 * 32);
  memset(c->partials, 0, slots * 32);
  c->scales = calloc(slots, 1);
  c->scale_sums = calloc(P, 4);
  c->scale_sums_backup = calloc(P, 4);
  c->likelihoods = malloc(P * 8);
  for (size_t i = 0; i < P; i++) c->likelihoods[i] = NAN;
  c->scale_ln = scale_ln;
  double s = (double)scale_ln;
  c->mult = exp(s);
  c->thr = exp(-s);
  return c;
}

void or_calc4_free(or_calc4 *c) {
  if (!c) return;
  free(c->sel); free(c->samples); free(c->partials); free(c->likelihoods);
  free(c->scales); free(c->scale_sums); free(c->scale_sums_backup);
  free(c);
}

int or_calc4_num_threads(const or_calc4 *c) { return c->threads; }

typedef double tm5[4][5];

typedef struct {
  or_calc4 *c;
  uint32_t n;
  const uint32_t *internals, *children;
  const tm5 *tms;
} p4_ctx;

/* cpu4.rs:403-433 */
static inline void scale4(double prod[4], uint8_t *scales_new,
                          const uint8_t *scales_old, uint32_t *scale_sum,
                          double thr, double mult) {
  uint8_t old = *scales_old;
  if (prod[0] < thr && prod[1] < thr && prod[2] < thr && prod[3] < thr) {
    for (int j = 0; j < 4; j++) prod[j] *= mult;
    if (!old) *scale_sum += 1;
    *scales_new = 1;
  } else {
    if (old) *scale_sum -= 1;
    *scales_new = 0;
  }
}

/* the slice body of cpu4.rs:202-275 */
static void p4_slice(void *vctx, size_t start, size_t count) {
  p4_ctx *x = vctx;
  or_calc4 *c = x->c;
  size_t P = c->P;
  uint32_t L = c->L;
#define OFF(idx) (((size_t)(idx) * 2 + (c->sel[idx] & 1u)) * P)
#define OFF_OLD(idx) (((size_t)(idx) * 2 + ((c->sel[idx] & 1u) ^ 1u)) * P)
  for (uint32_t i = 0; i < x->n; i++) {
    uint32_t node = x->internals[i];
    uint32_t left = x->children[2 * i], right = x->children[2 * i + 1];
    uint32_t idx = node - L;
    double(*out)[4] = c->partials + OFF(idx) + start;
    uint8_t *sn = c->scales + OFF(idx) + start;
    uint8_t *so = c->scales + OFF_OLD(idx) + start;
    uint32_t *ss = c->scale_sums + start;
    const tm5 *tl = &x->tms[2 * i], *tr = &x->tms[2 * i + 1];
    if (right < L) { /* calc_sample_sample, cpu4.rs:288-318 */
      const uint8_t *sl = c->samples + (size_t)left * P + start;
      const uint8_t *sr = c->samples + (size_t)right * P + start;
      for (size_t p = 0; p < count; p++) {
        if (so[p]) ss[p] -= 1;
        sn[p] = 0;
        for (int j = 0; j < 4; j++) out[p][j] = (*tl)[j][sl[p]] * (*tr)[j][sr[p]];
      }
    } else if (left < L) { /* calc_sample_partial, cpu4.rs:321-358 */
      const uint8_t *sl = c->samples + (size_t)left * P + start;
      const double(*pr)[4] = c->partials + OFF(right - L) + start;
      for (size_t p = 0; p < count; p++) {
        double y[4] = {pr[p][0], pr[p][1], pr[p][2], pr[p][3]};
        double rr[4] = {0, 0, 0, 0};
        for (int row = 0; row < 4; row++)
          for (int col = 0; col < 4; col++) rr[row] += (*tr)[row][col] * y[col];
        double prod[4];
        for (int j = 0; j < 4; j++) prod[j] = (*tl)[j][sl[p]] * rr[j];
        scale4(prod, &sn[p], &so[p], &ss[p], c->thr, c->mult);
        for (int j = 0; j < 4; j++) out[p][j] = prod[j];
      }
    } else { /* calc_partial_partial, cpu4.rs:361-401 */
      const double(*pl)[4] = c->partials + OFF(left - L) + start;
      const double(*pr)[4] = c->partials + OFF(right - L) + start;
      for (size_t p = 0; p < count; p++) {
        double rl[4] = {0, 0, 0, 0}, rr[4] = {0, 0, 0, 0};
        double yl[4] = {pl[p][0], pl[p][1], pl[p][2], pl[p][3]};
        double yr[4] = {pr[p][0], pr[p][1], pr[p][2], pr[p][3]};
        for (int row = 0; row < 4; row++)
          for (int col = 0; col < 4; col++) {
            rl[row] += (*tl)[row][col] * yl[col];
            rr[row] += (*tr)[row][col] * yr[col];
          }
        double prod[4];
        for (int j = 0; j < 4; j++) prod[j] = rl[j] * rr[j];
        scale4(prod, &sn[p], &so[p], &ss[p], c->thr, c->mult);
        for (int j = 0; j < 4; j++) out[p][j] = prod[j];
      }
    }
  }
#undef OFF
#undef OFF_OLD
}

/* cpu4.rs:39-76 and 170-285 */
void or_calc4_propose(or_calc4 *c, uint32_t n, const uint32_t *internals,
                      const uint32_t *children, const double *mats,
                      const double *freqs) {
  if (n == 0) return;
  tm5 *tms = malloc((size_t)2 * n * sizeof(tm5));
  for (uint32_t m = 0; m < 2 * n; m++) /* 5th column of 1.0, cpu4.rs:53-61 */
    for (int r = 0; r < 4; r++) {
      for (int k = 0; k < 4; k++) tms[m][r][k] = mats[(size_t)m * 16 + r * 4 + k];
      tms[m][r][4] = 1.0;
    }
  for (uint32_t i = 0; i < n; i++) /* set_selectors, cpu4.rs:101-106 */
    or_editbuf_set_edited(c->sel, internals[i] - c->L);
  p4_ctx ctx = {c, n, internals, children, (const tm5 *)tms};
  for_slices(c->threads, c->P, p4_slice, &ctx);
  uint32_t ridx = internals[n - 1] - c->L; /* root, cpu4.rs:277-284 */
  const double(*root)[4] =
      c->partials + ((size_t)ridx * 2 + (c->sel[ridx] & 1u)) * c->P;
  for (size_t i = 0; i < c->P; i++) {
    double prod[4];
    for (int j = 0; j < 4; j++) prod[j] = root[i][j] * freqs[j];
    c->likelihoods[i] = log(((prod[0] + prod[1]) + prod[2]) + prod[3]);
  }
  for (size_t i = 0; i < c->P; i++) /* cpu4.rs:69-73 */
    c->likelihoods[i] -= (double)(uint32_t)(c->scale_sums[i] * c->scale_ln);
  free(tms);
}

void or_calc4_accept(or_calc4 *c) { /* cpu4.rs:78-83 */
  memcpy(c->scale_sums_backup, c->scale_sums, c->P * 4);
  or_editbuf_accept(c->sel, c->I);
}

void or_calc4_reject(or_calc4 *c) { /* cpu4.rs:85-89 */
  memcpy(c->scale_sums, c->scale_sums_backup, c->P * 4);
  or_editbuf_reject(c->sel, c->I);
}

const double *or_calc4_likelihoods(const or_calc4 *c) { return c->likelihoods; }

/* ======================================================================== */
/* S-state calculator: cpu4.rs semantics with S states.  Tip codes are state
 * indices 0..S-1; any value >= S is fully unknown (tip partial all ones). */
struct or_calcn {
  uint32_t L, I;
  int S;
  size_t P;
  uint8_t *sel;
  uint8_t *samples;
  double *partials; /* [I*2*P][S] */
  double *likelihoods;
  uint8_t *scales;
  uint32_t *scale_sums, *scale_sums_backup;
  uint32_t scale_ln;
  double thr, mult;
  int threads;
};

or_calcn *or_calcn_new(size_t P, const uint8_t *samples, uint32_t L, int S,
                       uint32_t scale_ln, int num_threads) {
  or_calcn *c = calloc(1, sizeof *c);
  c->L = L;
  c->I = L - 1;
  c->S = S;
  c->P = P;
  c->sel = calloc(c->I, 1);
  c->samples = malloc((size_t)L * P);
  memcpy(c->samples, samples, (size_t)L * P);
  if (num_threads == 0) {
    long cores = host_cores();
    size_t want = (P + 999) / 1000;
    num_threads = (int)((size_t)cores < want ? (size_t)cores : want);
  }
  c->threads = num_threads < 1 ? 1 : num_threads;
  size_t slots = (size_t)c->I * P * 2;
  c->partials = calloc(slots * (size_t)S, 8);
  c->scales = calloc(slots, 1);
  c->scale_sums = calloc(P, 4);
  c->scale_sums_backup = calloc(P, 4);
  c->likelihoods = malloc(P * 8);
  for (size_t i = 0; i < P; i++) c->likelihoods[i] = NAN;
  c->scale_ln = scale_ln;
  c->mult = exp((double)scale_ln);
  c->thr = exp(-(double)scale_ln);
  return c;
}

void or_calcn_free(or_calcn *c) {
  if (!c) return;
  free(c->sel); free(c->samples); free(c->partials); free(c->likelihoods);
  free(c->scales); free(c->scale_sums); free(c->scale_sums_backup);
  free(c);
}

typedef struct {
  or_calcn *c;
  uint32_t n;
  const uint32_t *internals, *children;
  const double *mats;
} pn_ctx;

static void pn_slice(void *vctx, size_t start, size_t count) {
  pn_ctx *x = vctx;
  or_calcn *c = x->c;
  size_t P = c->P;
  int S = c->S;
  uint32_t L = c->L;
  double *rl = malloc(S * 8), *rr = malloc(S * 8), *prod = malloc(S * 8);
  for (uint32_t i = 0; i < x->n; i++) {
    uint32_t node = x->internals[i];
    uint32_t left = x->children[2 * i], right = x->children[2 * i + 1];
    uint32_t idx = node - L;
    size_t off = ((size_t)idx * 2 + (c->sel[idx] & 1u)) * P;
    size_t off_old = ((size_t)idx * 2 + ((c->sel[idx] & 1u) ^ 1u)) * P;
    const double *tl = x->mats + (size_t)2 * i * S * S;
    const double *tr = tl + (size_t)S * S;
    for (size_t p = start; p < start + count; p++) {
      /* left projection */
      if (left < L) {
        uint8_t s = c->samples[(size_t)left * P + p];
        for (int j = 0; j < S; j++) rl[j] = s < S ? tl[j * S + s] : 1.0;
      } else {
        uint32_t ci = left - L;
        const double *y =
            c->partials + (((size_t)ci * 2 + (c->sel[ci] & 1u)) * P + p) * S;
        for (int j = 0; j < S; j++) {
          double acc = 0.0;
          for (int k = 0; k < S; k++) acc += tl[j * S + k] * y[k];
          rl[j] = acc;
        }
      }
      if (right < L) {
        uint8_t s = c->samples[(size_t)right * P + p];
        for (int j = 0; j < S; j++) rr[j] = s < S ? tr[j * S + s] : 1.0;
      } else {
        uint32_t ci = right - L;
        const double *y =
            c->partials + (((size_t)ci * 2 + (c->sel[ci] & 1u)) * P + p) * S;
        for (int j = 0; j < S; j++) {
          double acc = 0.0;
          for (int k = 0; k < S; k++) acc += tr[j * S + k] * y[k];
          rr[j] = acc;
        }
      }
      for (int j = 0; j < S; j++) prod[j] = rl[j] * rr[j];
      uint8_t old = c->scales[off_old + p];
      if (right < L) { /* both tips: never scaled (cpu4.rs:307-311) */
        if (old) c->scale_sums[p] -= 1;
        c->scales[off + p] = 0;
      } else {
        int all = 1;
        for (int j = 0; j < S; j++) all &= prod[j] < c->thr;
        if (all) {
          for (int j = 0; j < S; j++) prod[j] *= c->mult;
          if (!old) c->scale_sums[p] += 1;
          c->scales[off + p] = 1;
        } else {
          if (old) c->scale_sums[p] -= 1;
          c->scales[off + p] = 0;
        }
      }
      memcpy(c->partials + (off + p) * S, prod, S * 8);
    }
  }
  free(rl); free(rr); free(prod);
}

void or_calcn_propose(or_calcn *c, uint32_t n, const uint32_t *internals,
                      const uint32_t *children, const double *mats,
                      const double *freqs) {
  if (n == 0) return;
  for (uint32_t i = 0; i < n; i++) or_editbuf_set_edited(c->sel, internals[i] - c->L);
  pn_ctx ctx = {c, n, internals, children, mats};
  for_slices(c->threads, c->P, pn_slice, &ctx);
  uint32_t ridx = internals[n - 1] - c->L;
  size_t off = ((size_t)ridx * 2 + (c->sel[ridx] & 1u)) * c->P;
  for (size_t i = 0; i < c->P; i++) {
    double s = 0.0;
    for (int j = 0; j < c->S; j++) s += c->partials[(off + i) * c->S + j] * freqs[j];
    c->likelihoods[i] = log(s) - (double)(uint32_t)(c->scale_sums[i] * c->scale_ln);
  }
}

void or_calcn_accept(or_calcn *c) {
  memcpy(c->scale_sums_backup, c->scale_sums, c->P * 4);
  or_editbuf_accept(c->sel, c->I);
}

void or_calcn_reject(or_calcn *c) {
  memcpy(c->scale_sums, c->scale_sums_backup, c->P * 4);
  or_editbuf_reject(c->sel, c->I);
}

const double *or_calcn_likelihoods(const or_calcn *c) { return c->likelihoods; }

/* ======================================================================== */
/* GammaLikelihood (gamma.rs:97-203) over K Cpu4Calculators                 */
struct or_gamma_lik {
  or_tree *tree;
  or_model model, model_bak;
  int K;
  double alpha, alpha_bak;
  int alpha_changed, model_changed;
  double rates[64], rates_bak[64];
  double clock;
  size_t P;
  uint32_t *weights;
  or_calc4 **calcs;
  int parallel;
  double cache, last;
  int launched;
};

or_gamma_lik *or_gamma_lik_new(or_tree *tree, const or_model *model,
                               const uint8_t *codes, uint32_t L, size_t P,
                               const uint32_t *weights, int K, double alpha,
                               double clock, uint32_t scale_ln,
                               int threads_per_cat, int parallel) {
  if (K < 1 || K > 64) return NULL;
  or_gamma_lik *g = calloc(1, sizeof *g);
  g->tree = tree;
  g->model = *model;
  g->model_bak = *model;
  g->K = K;
  g->alpha = alpha;
  g->alpha_bak = alpha;
  g->clock = clock;
  g->P = P;
  g->weights = malloc(P * 4);
  memcpy(g->weights, weights, P * 4);
  g->calcs = calloc(K, sizeof(or_calc4 *));
  for (int c = 0; c < K; c++)
    g->calcs[c] = or_calc4_new(P, codes, L, scale_ln, threads_per_cat);
  g->parallel = parallel;
  if (K == 1) g->rates[0] = 1.0; /* DnaLikelihood: one category, rate 1 */
  else or_gamma_categories(alpha, K, g->rates);
  memcpy(g->rates_bak, g->rates, sizeof g->rates);
  /* constructor: mark all, likelihood, accept (gamma.rs:89-93) */
  or_tree_mark_all_edges_updated(tree);
  or_gamma_lik_likelihood(g);
  or_gamma_lik_accept(g);
  or_tree_accept(tree);
  return g;
}

void or_gamma_lik_free(or_gamma_lik *g) {
  if (!g) return;
  for (int c = 0; c < g->K; c++) or_calc4_free(g->calcs[c]);
  free(g->calcs);
  free(g->weights);
  free(g);
}

void or_gamma_lik_set_model(or_gamma_lik *g, const or_model *m) {
  g->model = *m;
  g->model_changed = 1;
}

void or_gamma_lik_set_alpha(or_gamma_lik *g, double alpha) {
  g->alpha = alpha;
  g->alpha_changed = 1;
}

typedef struct {
  or_calc4 *c;
  uint32_t n;
  const uint32_t *internals, *children;
  const double *mats;
  const double *freqs;
} cat_job;

static void *cat_job_run(void *arg) {
  cat_job *j = arg;
  or_calc4_propose(j->c, j->n, j->internals, j->children, j->mats, j->freqs);
  return NULL;
}

double or_gamma_lik_likelihood(or_gamma_lik *g) {
  or_tree *t = g->tree;
  if (g->model_changed) or_tree_mark_all_edges_updated(t); /* gamma.rs:106-108 */
  int changed = or_tree_is_changed(t) || g->alpha_changed;
  if (!changed) { g->last = g->cache; return g->cache; }
  if (g->alpha_changed) { /* gamma.rs:121-124 */
    or_tree_mark_all_edges_updated(t);
    if (g->K > 1) or_gamma_categories(g->alpha, g->K, g->rates);
  }
  uint32_t *internals = malloc(t->N * 4), *children = malloc(2 * t->N * 4);
  uint32_t n = or_tree_partials_lists(t, internals, children);
  int K = g->K;
  double *mats = malloc((size_t)K * 2 * n * 16 * 8);
  double freqs[4];
  or_model_frequencies(&g->model, freqs);
  for (int c = 0; c < K; c++) /* transitions: length * clock * r_c */
    for (uint32_t m = 0; m < 2 * n; m++) {
      double len = or_tree_edge_length(t, children[m]) * (g->clock * g->rates[c]);
      or_model_transition(&g->model, len, mats + ((size_t)c * 2 * n + m) * 16);
    }
  g->launched = 1;
  cat_job jobs[64];
  pthread_t tid[64];
  for (int c = 0; c < K; c++)
    jobs[c] = (cat_job){g->calcs[c], n, internals, children,
                        mats + (size_t)c * 2 * n * 16, freqs};
  if (g->parallel && K > 1) {
    for (int c = 1; c < K; c++) pthread_create(&tid[c], NULL, cat_job_run, &jobs[c]);
    cat_job_run(&jobs[0]);
    for (int c = 1; c < K; c++) pthread_join(tid[c], NULL);
  } else {
    for (int c = 0; c < K; c++) cat_job_run(&jobs[c]);
  }
  double lik;
  if (K == 1) { /* DnaLikelihood::likelihood, dna.rs:96-97 */
    lik = or_weighted_sum(or_calc4_likelihoods(g->calcs[0]), g->weights, g->P);
  } else {
    const double *site[64];
    for (int c = 0; c < K; c++) site[c] = or_calc4_likelihoods(g->calcs[c]);
    lik = or_gamma_mix(K, g->P, site, g->weights, NULL);
  }
  free(internals); free(children); free(mats);
  g->last = lik;
  return lik;
}

void or_gamma_lik_accept(or_gamma_lik *g) { /* gamma.rs:171-186 */
  g->cache = g->last;
  if (g->launched) {
    memcpy(g->rates_bak, g->rates, sizeof g->rates);
    for (int c = 0; c < g->K; c++) or_calc4_accept(g->calcs[c]);
    g->model_bak = g->model;
    g->alpha_bak = g->alpha;
  }
  g->alpha_changed = 0;
  g->model_changed = 0;
  g->launched = 0;
}

void or_gamma_lik_reject(or_gamma_lik *g) { /* gamma.rs:188-202 */
  if (g->launched) {
    memcpy(g->rates, g->rates_bak, sizeof g->rates);
    for (int c = 0; c < g->K; c++) or_calc4_reject(g->calcs[c]);
    g->model = g->model_bak;
    g->alpha = g->alpha_bak;
  }
  g->alpha_changed = 0;
  g->model_changed = 0;
  g->launched = 0;
}

const double *or_gamma_lik_site(const or_gamma_lik *g, int c) {
  return or_calc4_likelihoods(g->calcs[c]);
}

const double *or_gamma_lik_rates(const or_gamma_lik *g) { return g->rates; }

/* ======================================================================== */
/* Propagation calculator (cuda_propagations): every node n stores, in one of
 * two slots chosen by its selector (an EditBuf over all nodes), the
 * projection P(edge n) x_n of its likelihood x_n (kernels.cu:22-59).  An
 * update list from DatedTree::propagation_lists (tree.rs:436-454) names the
 * updated leaves, then the internal nodes children first, root last; one
 * transition per listed node but the root (mod.rs:83-86).  Internal node:
 * x = proj(left) * proj(right); if all four values < e^-scale_ln the node is
 * scaled by e^scale_ln and flagged (kernels.cu:110-136); then
 * proj = P x (kernels.cu:138-150).  Root: ln(pi . x) - scale_sum, with the
 * root's own flag cleared (kernels.cu:155-187). */
struct or_prop4 {
  size_t P;
  uint32_t L, N, scale_ln;
  double threshold, mult;
  uint8_t *leaves;     /* [L][P] */
  double *proj;        /* [2N][P][4] */
  uint8_t *scales;     /* [2N][P] */
  uint32_t *ss, *ss_bak;
  uint8_t *sel;        /* [N] */
  double *lik;         /* [P] */
};

or_prop4 *or_prop4_new(size_t P, const uint8_t *samples, uint32_t L, uint32_t scale_ln) {
  if (L < 2 || P == 0) return NULL;
  or_prop4 *c = calloc(1, sizeof *c);
  c->P = P;
  c->L = L;
  c->N = 2 * L - 1;
  c->scale_ln = scale_ln;
  c->threshold = exp(-(double)scale_ln); /* mod.rs:267-268 */
  c->mult = exp((double)scale_ln);
  c->leaves = malloc((size_t)L * P);
  memcpy(c->leaves, samples, (size_t)L * P);
  c->proj = calloc((size_t)2 * c->N * P * 4, sizeof(double));
  c->scales = calloc((size_t)2 * c->N * P, 1);
  c->ss = calloc(P, 4);
  c->ss_bak = calloc(P, 4);
  c->sel = calloc(c->N, 1);
  c->lik = calloc(P, 8);
  return c;
}

void or_prop4_free(or_prop4 *c) {
  if (!c) return;
  free(c->leaves); free(c->proj); free(c->scales);
  free(c->ss); free(c->ss_bak); free(c->sel); free(c->lik);
  free(c);
}

static inline size_t prop_slot(const or_prop4 *c, uint32_t node) {
  return (size_t)node * 2 + (c->sel[node] & 1u);
}

/* kernels.cu:5-7: a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w, left to right */
static inline double dot4(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
}

void or_prop4_propose(or_prop4 *c, uint32_t num_nodes, const uint32_t *nodes,
                      uint32_t num_leaves, const uint32_t *children,
                      const double *mats, const double *freqs) {
  const size_t P = c->P;
  for (uint32_t i = 0; i < num_nodes; i++) or_editbuf_set_edited(c->sel, nodes[i]);
  for (uint32_t i = 0; i < num_leaves; i++) { /* CALCULATE_LEAF_PROJECTION */
    const uint32_t leaf = nodes[i];
    const double *M = mats + (size_t)i * 16;
    double *dst = c->proj + prop_slot(c, leaf) * P * 4;
    for (size_t p = 0; p < P; p++) {
      const uint8_t v = c->leaves[(size_t)leaf * P + p];
      for (int s = 0; s < 4; s++) {
        double x = 1.0;
        if (v == 1) x = M[s * 4 + 0];
        else if (v == 2) x = M[s * 4 + 1];
        else if (v == 4) x = M[s * 4 + 2];
        else if (v == 8) x = M[s * 4 + 3];
        dst[p * 4 + s] = x;
      }
    }
  }
  for (uint32_t i = num_leaves; i + 1 < num_nodes; i++) {
    const uint32_t node = nodes[i];
    const uint32_t l = children[2 * (i - num_leaves)], r = children[2 * (i - num_leaves) + 1];
    const double *M = mats + (size_t)i * 16;
    const size_t snew = prop_slot(c, node), sold = snew ^ 1u;
    const double *pl = c->proj + prop_slot(c, l) * P * 4;
    const double *pr = c->proj + prop_slot(c, r) * P * 4;
    double *dst = c->proj + snew * P * 4;
    for (size_t p = 0; p < P; p++) {
      double x[4];
      for (int s = 0; s < 4; s++) x[s] = pl[p * 4 + s] * pr[p * 4 + s];
      const int scale = x[0] < c->threshold && x[1] < c->threshold &&
                        x[2] < c->threshold && x[3] < c->threshold;
      if (scale)
        for (int s = 0; s < 4; s++) x[s] *= c->mult;
      const uint8_t old = c->scales[sold * P + p];
      c->scales[snew * P + p] = (uint8_t)scale;
      if (old == 0 && scale) c->ss[p] += c->scale_ln;
      else if (old == 1 && !scale) c->ss[p] -= c->scale_ln;
      for (int s = 0; s < 4; s++) dst[p * 4 + s] = dot4(M + s * 4, x);
    }
  }
  { /* update_likelihoods */
    const uint32_t root = nodes[num_nodes - 1];
    const uint32_t l = children[2 * (num_nodes - 1 - num_leaves)];
    const uint32_t r = children[2 * (num_nodes - 1 - num_leaves) + 1];
    const size_t snew = prop_slot(c, root), sold = snew ^ 1u;
    const double *pl = c->proj + prop_slot(c, l) * P * 4;
    const double *pr = c->proj + prop_slot(c, r) * P * 4;
    for (size_t p = 0; p < P; p++) {
      double x[4];
      for (int s = 0; s < 4; s++) x[s] = pl[p * 4 + s] * pr[p * 4 + s];
      double lk = dot4(x, freqs);
      c->scales[snew * P + p] = 0;
      if (c->scales[sold * P + p]) c->ss[p] -= c->scale_ln;
      c->lik[p] = log(lk) - (double)c->ss[p];
    }
  }
}

void or_prop4_accept(or_prop4 *c) { /* mod.rs:143-153 */
  memcpy(c->ss_bak, c->ss, c->P * 4);
  or_editbuf_accept(c->sel, c->N);
}

void or_prop4_reject(or_prop4 *c) { /* mod.rs:155-164 */
  memcpy(c->ss, c->ss_bak, c->P * 4);
  or_editbuf_reject(c->sel, c->N);
}

const double *or_prop4_likelihoods(const or_prop4 *c) { return c->lik; }

/* ======================================================================== */
/* Per-category models with per-edge clock rates, evaluated from scratch.
 * The likelihood layers' totals over K calculators:
 *   HeteroLikelihood (hetero.rs:122-160): calculator c proposes with its own
 *     transitions (model c, rate of clock c per edge); pattern i adds
 *     w_i * lnL_{classes[i]}(i), summed serially in pattern order.
 *   GammaLikelihood (gamma.rs:126-166) when classes is NULL: LSE over the
 *     categories plus ln(1/K), weighted (or_gamma_mix); K == 1 is
 *     DnaLikelihood's weighted sum (dna.rs:96-97).
 * Branch length of edge e in category c: edge_length(e) * edge_rates[c][e]
 * (transitions.rs:33-35, with clock.get_rate(edge) [* category rate]
 * already formed by the caller, clock.rs:115-119 / gamma.rs:131-133).
 * Every internal node is recomputed (all edges marked), so the value does
 * not depend on any earlier state. */
double or_components_eval(const or_tree *t, int K, const or_model *models,
                          const double *edge_rates, const uint8_t *codes,
                          size_t P, const uint32_t *weights,
                          const uint32_t *classes, uint32_t scale_ln,
                          double *site) {
  if (K < 1 || K > 64) return NAN;
  uint8_t *all = malloc(t->N);
  memset(all, 1, t->N);
  uint32_t *internals = malloc(t->N * 4), *children = malloc(2 * (size_t)t->N * 4);
  uint32_t n = updated_internals(t, all, internals);
  children_list(t, internals, n, children);
  double *mats = malloc((size_t)2 * n * 16 * 8);
  double *sites = malloc((size_t)K * P * 8);
  for (int c = 0; c < K; c++) {
    for (uint32_t m = 0; m < 2 * n; m++) {
      uint32_t e = children[m];
      double len = or_tree_edge_length(t, e) * edge_rates[(size_t)c * t->N + e];
      or_model_transition(&models[c], len, mats + (size_t)m * 16);
    }
    double freqs[4];
    or_model_frequencies(&models[c], freqs);
    or_calc4 *calc = or_calc4_new(P, codes, t->L, scale_ln, 1);
    or_calc4_propose(calc, n, internals, children, mats, freqs);
    memcpy(sites + (size_t)c * P, or_calc4_likelihoods(calc), P * 8);
    or_calc4_free(calc);
  }
  double total = 0.0;
  if (classes) {
    for (size_t i = 0; i < P; i++)
      total += sites[(size_t)classes[i] * P + i] * (double)weights[i];
  } else if (K == 1) {
    total = or_weighted_sum(sites, weights, P);
  } else {
    const double *rows[64];
    for (int c = 0; c < K; c++) rows[c] = sites + (size_t)c * P;
    total = or_gamma_mix(K, P, rows, weights, NULL);
  }
  if (site) memcpy(site, sites, (size_t)K * P * 8);
  free(all); free(internals); free(children); free(mats); free(sites);
  return total;
}

/* ======================================================================== */
/* deduplicate (likelihood/mod.rs:74-132).  Columns are compared by content
 * (the reference compares blake3 digests of the column bytes); patterns are
 * numbered in order of first occurrence. */
typedef struct { uint64_t h; size_t site; } hent;

static uint64_t col_hash(const uint8_t *msa, uint32_t ns, size_t nsites,
                         size_t site) {
  uint64_t h = 1469598103934665603ull; /* FNV-1a */
  for (uint32_t s = 0; s < ns; s++) {
    h ^= msa[(size_t)s * nsites + site];
    h *= 1099511628211ull;
  }
  return h;
}

static int col_eq(const uint8_t *msa, uint32_t ns, size_t nsites, size_t a,
                  size_t b) {
  for (uint32_t s = 0; s < ns; s++)
    if (msa[(size_t)s * nsites + a] != msa[(size_t)s * nsites + b]) return 0;
  return 1;
}

size_t or_deduplicate(uint32_t ns, size_t nsites, const uint8_t *msa,
                      uint8_t *leaves, uint32_t *weights,
                      size_t *sites_to_patterns) {
  size_t cap = 1;
  while (cap < 2 * nsites + 2) cap <<= 1;
  size_t *table = malloc(cap * sizeof(size_t)); /* site+1 of first occurrence */
  memset(table, 0, cap * sizeof(size_t));
  size_t *first = malloc((nsites + 1) * sizeof(size_t));
  size_t np = 0;
  for (size_t site = 0; site < nsites; site++) {
    uint64_t h = col_hash(msa, ns, nsites, site);
    size_t slot = h & (cap - 1);
    for (;;) {
      if (table[slot] == 0) {
        table[slot] = site + 1;
        first[np] = site;
        weights[np] = 1;
        sites_to_patterns[site] = np;
        np++;
        break;
      }
      size_t other = table[slot] - 1;
      if (col_eq(msa, ns, nsites, other, site)) {
        size_t pat = sites_to_patterns[other];
        weights[pat] += 1;
        sites_to_patterns[site] = pat;
        break;
      }
      slot = (slot + 1) & (cap - 1);
    }
  }
  for (uint32_t s = 0; s < ns; s++)
    for (size_t p = 0; p < np; p++) {
      uint8_t ch = msa[(size_t)s * nsites + first[p]];
      leaves[(size_t)s * np + p] = ch == 0 ? 0x0F : ch; /* gap -> 0b1111 */
    }
  free(table);
  free(first);
  return np;
}
