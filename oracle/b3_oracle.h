/*
 * b3_oracle.h -- FP64 CPU restatement of b3's Felsenstein-pruning likelihood.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * by nothing else.  The product (aspartik_amd/) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the aspartik repository root).  Parity status per function is listed in the
 * header comment of b3_oracle.c.
 */
#ifndef B3_ORACLE_H
#define B3_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_TOMBSTONE 0xFFFFFFFFu

/* ---- EditBuf (crates/sk/src/editbuf.rs:8-105) ---------------------------- */
void or_editbuf_set_edited(uint8_t *m, size_t index);
void or_editbuf_accept(uint8_t *m, size_t len);
void or_editbuf_reject(uint8_t *m, size_t len);

/* ---- DatedTree update schedule (crates/b3/src/parameters/tree.rs) ------- */
typedef struct or_tree or_tree;

/* parents[n] is the parent of node n, OR_TOMBSTONE for the root.  children
 * are given as (L-1) pairs in the order of internal nodes L..2L-2, exactly as
 * DatedTree stores them (tree.rs:701-707).  heights[n] per node. */
or_tree *or_tree_new(uint32_t num_leaves, const uint32_t *parents,
                     const uint32_t *children, const double *heights);
void or_tree_free(or_tree *t);
uint32_t or_tree_root(const or_tree *t);
void or_tree_mark_edge_updated(or_tree *t, uint32_t edge);
void or_tree_mark_all_edges_updated(or_tree *t);
/* returns count; writes edges (<= N) */
uint32_t or_tree_edges_to_update(const or_tree *t, uint32_t *edges);
/* returns n; writes internals[n], children[2n] */
uint32_t or_tree_partials_lists(const or_tree *t, uint32_t *internals,
                                uint32_t *children);
/* returns n; writes nodes[n] (updated leaves first), children[2*internals],
 * *num_leaves_out (tree.rs:436-454) */
uint32_t or_tree_propagation_lists(const or_tree *t, uint32_t *nodes,
                                   uint32_t *children,
                                   uint32_t *num_leaves_out);
void or_tree_set_height(or_tree *t, uint32_t node, double h);
int or_tree_replace_child(or_tree *t, uint32_t parent, uint32_t old_child,
                          uint32_t new_child);
double or_tree_edge_length(const or_tree *t, uint32_t edge);
int or_tree_is_changed(const or_tree *t);
void or_tree_accept(or_tree *t);
void or_tree_reject(or_tree *t);
void or_tree_get(const or_tree *t, uint32_t *parents, uint32_t *children,
                 double *heights);

/* ---- Substitution models (crates/b3/src/substitution.rs) ----------------- */
typedef struct or_model {
  int kind;      /* 0 JC, 1 K80, 2 HKY, 3 GTR, 4 general reversible */
  int n;         /* states */
  double kappa;
  double freqs[64];
  double q[64 * 64];    /* normalised rate matrix (row-major) */
  double evec[64 * 64]; /* V   */
  double ievc[64 * 64]; /* V^-1 */
  double eval[64];      /* eigenvalues */
} or_model;

void or_model_jc(or_model *m);
void or_model_k80(or_model *m, double kappa);
void or_model_hky(or_model *m, double kappa, const double *freqs4);
void or_model_gtr(or_model *m, const double *rates6, const double *freqs4);
/* general reversible model from an (unnormalised) symmetric exchangeability
 * matrix R (n x n, diagonal ignored) and frequencies: q_ij = R_ij pi_j,
 * normalised to mean rate 1.  Used for the 61-state codon model. */
int or_model_reversible(or_model *m, int n, const double *R, const double *pi);
/* P(t) = exp(Q t) written row-major n x n */
void or_model_transition(const or_model *m, double distance, double *dst);
void or_model_frequencies(const or_model *m, double *dst);

/* symmetric eigendecomposition (cyclic Jacobi): A = U diag(w) U^T */
int or_sym_eigen(int n, const double *A, double *w, double *U);

/* ---- Discrete gamma rates (crates/b3/src/likelihood/gamma.rs:205-225) --- */
double or_gamma_p(double a, double x); /* regularised lower incomplete gamma */
double or_gamma_inv_cdf(double shape, double p);
void or_gamma_categories(double alpha, int k, double *rates);

/* ---- Likelihood helpers (crates/b3/src/likelihood/mod.rs:134-151) ------- */
double or_weighted_sum(const double *values, const uint32_t *weights, size_t n);
double or_log_sum_exp(const double *values, size_t n);
/* gamma.rs:154-166: serial per-pattern LSE over K plus ln(1/K), weighted */
double or_gamma_mix(int k, size_t p, const double *const *site_lls,
                    const uint32_t *weights, double *mixed_out);

/* ---- 4-state FP64 calculator (crates/b3/src/calculator/cpu4.rs) -------- */
typedef struct or_calc4 or_calc4;
or_calc4 *or_calc4_new(size_t num_patterns, const uint8_t *samples,
                       uint32_t num_leaves, uint32_t scale_ln,
                       int num_threads);
void or_calc4_free(or_calc4 *c);
/* mats: 2n row-major 4x4 f64 in children order */
void or_calc4_propose(or_calc4 *c, uint32_t n, const uint32_t *internals,
                      const uint32_t *children, const double *mats,
                      const double *freqs);
void or_calc4_accept(or_calc4 *c);
void or_calc4_reject(or_calc4 *c);
const double *or_calc4_likelihoods(const or_calc4 *c);
int or_calc4_num_threads(const or_calc4 *c);

/* ---- S-state FP64 calculator (cpu4.rs semantics generalised to S states;
 * the reference's n-state StateCalculator is a todo!() stub) -------------- */
typedef struct or_calcn or_calcn;
or_calcn *or_calcn_new(size_t num_patterns, const uint8_t *samples,
                       uint32_t num_leaves, int states, uint32_t scale_ln,
                       int num_threads);
void or_calcn_free(or_calcn *c);
void or_calcn_propose(or_calcn *c, uint32_t n, const uint32_t *internals,
                      const uint32_t *children, const double *mats,
                      const double *freqs);
void or_calcn_accept(or_calcn *c);
void or_calcn_reject(or_calcn *c);
const double *or_calcn_likelihoods(const or_calcn *c);

/* ---- Gamma likelihood over K 4-state calculators (gamma.rs:97-203) ------ */
typedef struct or_gamma_lik or_gamma_lik;
/* codes: [L x P] IUPAC byte codes by leaf; weights[P]; model copied. */
or_gamma_lik *or_gamma_lik_new(or_tree *tree, const or_model *model,
                               const uint8_t *codes, uint32_t num_leaves,
                               size_t num_patterns, const uint32_t *weights,
                               int num_categories, double alpha,
                               double clock_rate, uint32_t scale_ln,
                               int threads_per_category, int parallel_categories);
void or_gamma_lik_free(or_gamma_lik *g);
/* one likelihood() evaluation for whatever the tree marks as dirty */
double or_gamma_lik_likelihood(or_gamma_lik *g);
void or_gamma_lik_accept(or_gamma_lik *g);
void or_gamma_lik_reject(or_gamma_lik *g);
void or_gamma_lik_set_model(or_gamma_lik *g, const or_model *model);
void or_gamma_lik_set_alpha(or_gamma_lik *g, double alpha);
const double *or_gamma_lik_site(const or_gamma_lik *g, int category);
const double *or_gamma_lik_rates(const or_gamma_lik *g);

/* ---- Propagation calculator (calculator/cuda_propagations) ------------- */
typedef struct or_prop4 or_prop4;
/* samples: [L x P] IUPAC bytes (no remapping) */
or_prop4 *or_prop4_new(size_t num_patterns, const uint8_t *samples,
                       uint32_t num_leaves, uint32_t scale_ln);
void or_prop4_free(or_prop4 *c);
/* DatedTree::propagation_lists: nodes[num_nodes] (updated leaves first, then
 * internals children first, root last), children[2 (num_nodes -
 * num_leaves)]; mats: [num_nodes - 1][4][4], node order */
void or_prop4_propose(or_prop4 *c, uint32_t num_nodes, const uint32_t *nodes,
                      uint32_t num_leaves, const uint32_t *children,
                      const double *mats, const double *freqs);
void or_prop4_accept(or_prop4 *c);
void or_prop4_reject(or_prop4 *c);
const double *or_prop4_likelihoods(const or_prop4 *c);

/* ---- Per-category models and per-edge clock rates, from scratch -------- */
/* HeteroLikelihood's total (hetero.rs:122-160) when classes != NULL, else
 * GammaLikelihood's mix (gamma.rs:126-166).  models[K]; edge_rates[K][N]
 * (N = 2L - 1 nodes, edge e = node e); codes [L x P]; writes the K x P site
 * log likelihoods to site when non-NULL. */
double or_components_eval(const or_tree *tree, int num_categories,
                          const or_model *models, const double *edge_rates,
                          const uint8_t *codes, size_t num_patterns,
                          const uint32_t *weights, const uint32_t *classes,
                          uint32_t scale_ln, double *site);

/* ---- Pattern deduplication (crates/b3/src/likelihood/mod.rs:74-132) ----- */
/* msa: [num_seqs x num_sites] byte codes (seq-major).  Returns P; writes
 * leaves[num_seqs x P], weights[P], sites_to_patterns[num_sites]. */
size_t or_deduplicate(uint32_t num_seqs, size_t num_sites, const uint8_t *msa,
                      uint8_t *leaves, uint32_t *weights,
                      size_t *sites_to_patterns);

#ifdef __cplusplus
}
#endif
#endif
