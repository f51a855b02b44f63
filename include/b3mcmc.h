/* MCMC around the MI355X likelihood: the step driver, the effective sample
 * size estimator and the neural surrogate used by delayed-acceptance MH.
 *
 * b3's MCMC loop is Rust (crates/b3/src/mcmc.rs:270-393) with Python
 * operators and priors (python/aspartik/b3/operators/); the likelihood is
 * the only part on the GPU.  This header is the native counterpart used to
 * measure ESS per second: it mirrors Mcmc::step / finalize_step
 * (mcmc.rs:314-393) over a GTR+Gamma likelihood evaluated by libb3hip's
 * FP32 calculator, with the operators NodeSlide (_node_slide.py:26-39),
 * TreeScale (_tree_scale.py:27-37), DeltaExchange (_delta_exchange.py:43-60),
 * ScaleReal (_param_scale.py:32-39) and NarrowExchange
 * (_tree_exchange.py:23-57).  ESS and the surrogate have no counterpart in
 * the reference.
 *
 * Every call returns a b3h_status; b3h_last_error() (b3hip.h) gives the
 * message.  Nothing here aborts the process.
 */
#ifndef B3MCMC_H
#define B3MCMC_H

#include <stddef.h>
#include <stdint.h>

#include "b3hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- effective sample size ------------------------------------------- */

/* ESS of a scalar trace by Geyer's initial monotone sequence estimator
 * (Geyer 1992): autocovariances by direct summation, pairs
 * Gamma_k = rho_2k + rho_2k+1 summed while positive and forced
 * non-increasing; ESS = n / (-1 + 2 sum_k Gamma_k).  Host only.  Returns 0
 * for n < 4 or a constant trace. */
double b3m_ess(const double *x, size_t n);

/* ---- the delayed-acceptance rule -------------------------------------- */

/* The two tests of one delayed-acceptance MH step (Christen & Fox 2005), as
 * the chain applies them; host only, no reference counterpart (b3 has no
 * surrogate).  f is the surrogate's ln L, tempered by beta in the first
 * stage: the first stage accepts iff
 *     beta (f_new - f_cur) + d_prior + hastings > ln u1,
 * i.e. with probability min(1, e^{beta df + d_prior + h}); the exact stage
 * then accepts iff
 *     d_loglik - beta (f_new - f_cur) > ln u2
 * (d_loglik = -inf rejects).  Together the chain targets the exact
 * posterior.  u1, u2 in (0, 1). */
int b3m_da_stage1(double beta, double f_cur, double f_new, double d_prior, double hastings,
                  double u1);
int b3m_da_stage2(double beta, double f_cur, double f_new, double d_loglik, double u2);

/* ---- MLP surrogate ---------------------------------------------------- */

/* inputs -> hidden -> hidden -> 1, GELU (tanh form) after both hidden
 * layers.  Master weights are FP32 (Adam state FP32); the forward and
 * backward passes use BF16 copies of the weights with FP32 accumulation.
 * Inputs and output are standardised with b3m_mlp_set_scaling. */
typedef struct b3m_mlp b3m_mlp;

b3h_status b3m_mlp_create(int device, uint32_t inputs, uint32_t hidden, uint64_t seed,
                          b3m_mlp **out);
b3h_status b3m_mlp_destroy(b3m_mlp *m);

/* x' = (x - x_mean) / x_std per input, y = y' * y_std + y_mean */
b3h_status b3m_mlp_set_scaling(b3m_mlp *m, const float *x_mean, const float *x_std,
                               float y_mean, float y_std);

/* y[batch] from x[batch][inputs] (host arrays).  A batch of one (with
 * inputs <= 512 and hidden <= 1024) is one launch whose result is polled
 * from pinned host memory; it sums in another order than larger batches,
 * so the two agree to FP32 rounding.  Calls on one handle must not overlap:
 * they share the handle's device buffers. */
b3h_status b3m_mlp_forward(b3m_mlp *m, const float *x, uint32_t batch, float *y);

/* Mean squared error of the standardised output and its gradient with
 * respect to every parameter (as FP32, at the BF16 weights), for the
 * gradient check.  Layout of grads (and of b3m_mlp_get_params):
 * W1[inputs][hidden], b1[hidden], W2[hidden][hidden], b2[hidden],
 * w3[hidden], b3[1].  The number of floats is b3m_mlp_num_params(). */
size_t b3m_mlp_num_params(const b3m_mlp *m);
b3h_status b3m_mlp_loss_grad(b3m_mlp *m, const float *x, const float *y, uint32_t batch,
                             float *loss, float *grads);
b3h_status b3m_mlp_get_params(const b3m_mlp *m, float *params);
b3h_status b3m_mlp_set_params(b3m_mlp *m, const float *params);

/* Adam on minibatches of the standardised MSE: `epochs` passes over the n
 * samples in a seeded shuffled order; returns the last epoch's mean loss. */
b3h_status b3m_mlp_train(b3m_mlp *m, const float *x, const float *y, uint32_t n,
                         uint32_t epochs, uint32_t batch, float lr, float *final_loss);

/* ---- chain ------------------------------------------------------------ */

typedef struct b3m_chain b3m_chain;

typedef struct {
  int32_t device;
  uint32_t categories;       /* Gamma categories (4) */
  uint64_t seed;
  double birth_rate;         /* Yule prior on the node heights */
  double alpha_prior_rate;   /* Exponential prior on the Gamma shape */
  /* operator weights (mcmc.rs picks one operator per step in proportion) */
  double w_node_slide, w_tree_scale, w_rates, w_alpha, w_narrow;
  double scale_factor;       /* TreeScale factor f: scale ~ U(f, 1/f) */
  double alpha_factor;       /* ScaleReal factor on alpha (0: scale_factor) */
  double delta_tuning;       /* DeltaExchange tuning t: delta ~ U(0, (1 - t) / 10) */
  /* Operator tuning as b3's Scheduler::tune with a LinearOptimizer
   * (operators/mod.rs:164-190,467-482; optimizers.py:13-38): every
   * tune_every steps each tunable operator's parameter moves by
   * rate * (0.234 - acceptance since the last tune), clamped to [0.01, 0.99];
   * rate = tune_magnitude until 0.9 tune_cutoff, then falls linearly to 0 at
   * tune_cutoff (steps counted over the chain's life).  tune_every = 0: off. */
  uint32_t tune_every;
  uint64_t tune_cutoff;
  double tune_magnitude;
  /* 1: every exact evaluation recomputes every node, as a run whose tree is
   * marked whole before each likelihood call; 0 (default): only the changed
   * edges' paths to the root (tree.rs:357-375) */
  uint32_t full_evaluations;
  /* delayed acceptance: 1: topology moves (NarrowExchange) pass the
   * surrogate's first stage like every other operator (the features are a
   * function of the whole labelled state); 0 (default): they go straight to
   * the exact MH test */
  uint32_t screen_topology;
  /* delayed acceptance: the first stage targets exp(beta f) x prior, f the
   * surrogate's ln L; the second accepts with min(1, exp(dlnL - beta df)).
   * beta < 1 lets more moves reach the exact test when f is noisy; any
   * beta > 0 keeps the exact posterior (default 1) */
  double surrogate_beta;
  /* NarrowExchange's notion of a grandparent.  0 (default): b3's, an
   * internal node with two internal children (tree.rs:767-770,
   * _tree_exchange.py:23-57).  Exchanging a leaf child of `parent` then
   * leaves the grandparent with a leaf child, so the reverse move cannot be
   * proposed and the chain does not keep its posterior (on 4 taxa a
   * balanced tree can become a caterpillar but never return).  1: BEAST2's
   * narrow Exchange, an internal node with at least one internal child; the
   * grandparent keeps that status, every move has its reverse, and the
   * Hastings term ln(V_before / V_after) counts such nodes. */
  uint32_t narrow_reversible;
} b3m_chain_config;

void b3m_chain_config_default(b3m_chain_config *cfg);

/* codes [L][P] as b3h_create; tree in b3's numbering (parents[N] with the
 * root's 0xFFFFFFFF, children[L-1][2], heights[N]); GTR exchangeabilities
 * (AC AG AT CG CT GT, any positive scale; kept summing to 1), fixed base
 * frequencies, Gamma shape. */
b3h_status b3m_chain_create(const b3m_chain_config *cfg, const uint8_t *codes, uint32_t L,
                            uint32_t P, const uint32_t *weights, const uint32_t *parents,
                            const uint32_t *children, const double *heights,
                            const double *gtr_rates, const double *freqs, double alpha,
                            b3m_chain **out);
b3h_status b3m_chain_destroy(b3m_chain *c);

/* Surrogate input of the current state: the 2L-3 log unrooted branch
 * lengths (the root's two edges summed, in edge order), log exchangeabilities
 * (6) and log alpha: 2L + 4 values (404 for 200 taxa). */
uint32_t b3m_chain_num_features(const b3m_chain *c);
b3h_status b3m_chain_features(const b3m_chain *c, float *out);

enum {
  B3M_PLAIN = 0,    /* exact likelihood every step */
  B3M_DELAYED = 1   /* surrogate first stage, exact second (screen_topology) */
};

/* Trace columns, one row per recorded step */
enum {
  B3M_COL_STEP = 0, B3M_COL_POSTERIOR, B3M_COL_LIKELIHOOD, B3M_COL_PRIOR, B3M_COL_ROOT_HEIGHT,
  B3M_COL_TREE_LENGTH, B3M_COL_ALPHA, B3M_COL_RATE0, /* 6 exchangeabilities */
  B3M_TRACE_COLS = B3M_COL_RATE0 + 6
};

typedef struct {
  uint64_t steps, accepted;
  uint64_t exact_evals;        /* likelihood evaluations on the GPU */
  uint64_t surrogate_evals;    /* forward passes of the surrogate */
  uint64_t stage1_rejects;     /* proposals the surrogate turned away */
  uint64_t stage2_rejects;     /* proposals rejected by the exact second stage */
  uint64_t aborts, prior_rejects;
  uint64_t op_proposed[5], op_accepted[5];  /* node slide, tree scale, rates, alpha, narrow */
  uint64_t collected;          /* (features, logL) pairs written */
  double seconds, likelihood_seconds, surrogate_seconds;
  double tuning[5];            /* operator parameters at the end (-1: not tunable) */
} b3m_run_stats;

/* Runs `steps` steps.  Every `record_every` steps one trace row is written
 * (trace may be null).  If collect_x is non-null, the features and exact
 * log-likelihood of every exact evaluation (accepted or not) are stored
 * until collect_cap pairs.  B3M_DELAYED needs b3m_chain_set_surrogate. */
b3h_status b3m_chain_run(b3m_chain *c, uint64_t steps, uint32_t mode, uint32_t record_every,
                         double *trace, float *collect_x, float *collect_y, uint64_t collect_cap,
                         b3m_run_stats *stats);

/* The chain keeps a reference; the surrogate must outlive its use. */
b3h_status b3m_chain_set_surrogate(b3m_chain *c, b3m_mlp *m);

/* Current tree (parents[2L-1] with the root's 0xFFFFFFFF, children
 * [L-1][2], heights[2L-1]), exchangeabilities [6] and alpha; any output may
 * be null. */
b3h_status b3m_chain_params(const b3m_chain *c, uint32_t *parents, uint32_t *children,
                            double *heights, double *gtr_rates, double *alpha);

/* Mcmc::dump / load_state (mcmc.rs:183-268) in b3's verbatim format
 * (crates/verbatim/src/lib.rs): tag 62 33 b3 44, protocol version 19, the
 * step count (u64) and the posterior (f64); then the parameters in order --
 * the tree as DatedTree::dump writes it (a u32 slice of parents, an f64
 * slice of heights), the six exchangeabilities as a RealVector, alpha as a
 * Real; then the scheduler, per operator (node slide, tree scale, rates,
 * alpha, narrow exchange) accepts and rejects since the last tune (u32 each)
 * and the tuning of the three tunable ones (f64); then the generator state
 * in 32 bytes where b3 writes PCG64's two u128 words.
 * dump: with buf null only *len is set; otherwise cap must hold *len bytes.
 * load: replaces the chain's state, rebuilds children from parents as
 * DatedTree::load_parents does, evaluates the likelihood from scratch and
 * accepts (the stored posterior is not trusted). */
b3h_status b3m_chain_dump_state(const b3m_chain *c, uint8_t *buf, size_t cap, size_t *len);
b3h_status b3m_chain_load_state(b3m_chain *c, const uint8_t *buf, size_t len);

/* current log likelihood / prior / posterior */
b3h_status b3m_chain_state(const b3m_chain *c, double *loglik, double *logprior,
                           double *posterior);

#ifdef __cplusplus
}
#endif

#endif
