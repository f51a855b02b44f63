/*
 * b3hip.h -- C ABI of the MI355X (gfx950) Felsenstein-pruning calculator.
 *
 * This is the drop-in boundary for b3's `Calculator<f64>` trait
 * (crates/b3/src/calculator/mod.rs:54-76).  One handle plays the role of the
 * K calculators that GammaLikelihood keeps, one per rate category
 * (crates/b3/src/likelihood/gamma.rs:57-66); with K = 1 it is exactly the one
 * calculator DnaLikelihood keeps (crates/b3/src/likelihood/dna.rs:41-42).
 *
 * Conventions
 *  - Plain C types only.  Every call returns a b3h_status; no call aborts or
 *    throws.  On failure b3h_last_error() gives a message (thread-local).
 *  - Node numbering is b3's: leaves 0..L-1, internal nodes L..2L-2; the edge
 *    above node n has index n (crates/b3/src/parameters/tree.rs:728-743).
 *  - A handle owns its device buffers and one HIP stream.  Calls on one handle
 *    must be serialised by the caller (b3 holds each calculator behind a
 *    mutex, dna.rs:18); distinct handles may be used from distinct threads
 *    concurrently (gamma.rs:143-152).  No global mutable state.
 *  - Protocol (calculator/mod.rs:44-53): propose -> (likelihoods | total) ->
 *    exactly one of accept / reject.
 */
#ifndef B3HIP_H
#define B3HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum b3h_status {
  B3H_OK = 0,
  B3H_ERR_ARG = 1,         /* invalid argument (null, size, node id, ...)   */
  B3H_ERR_NO_DEVICE = 2,   /* no HIP device / device ordinal out of range   */
  B3H_ERR_OOM = 3,         /* device or pinned-host allocation failed       */
  B3H_ERR_HIP = 4,         /* a HIP runtime call failed                     */
  B3H_ERR_STATE = 5,       /* call out of protocol order                    */
  B3H_ERR_UNSUPPORTED = 6  /* combination not implemented                   */
} b3h_status;

/* Arithmetic of the pruning kernels. */
typedef enum b3h_precision {
  /* FP32 partials with exact per-node power-of-two rescaling, FP64
   * exponents / root / mixing / weighted sum.  The product path. */
  B3H_F32_SCALED = 0,
  /* FP64 partials, one lane per (pattern, category), children read from
   * HBM, e^scale_ln threshold scaling with double-buffered flags: b3's own
   * GPU algorithm (cuda_partials/kernels.cu:36-140) at FP64, kept as the
   * plain-FP64 baseline. */
  B3H_F64_PLAIN = 1,
  /* Codon / S-state models: partials FP16 in HBM, P x partials on MFMA
   * (fp16 in, fp32 accumulate), power-of-two scaling. */
  B3H_F16_MFMA = 2,
  /* Propagation storage, b3's default GPU algorithm (cuda_propagations/
   * kernels.cu:73-187) at FP64: every node keeps the projection P x of its
   * likelihood in two selector slots, e^scale_ln flag scaling of any node.
   * Takes propagation lists (b3h_propose_propagations*). */
  B3H_F64_PROPAGATION = 3,
  /* Propagation storage with FP32 projections and power-of-two exponents,
   * FP64 root and sums: the storage alternative to B3H_F32_SCALED. */
  B3H_F32_PROPAGATION = 4
} b3h_precision;

typedef struct b3h_config {
  int32_t device;          /* HIP device ordinal                              */
  uint32_t precision;      /* b3h_precision                                   */
  uint32_t states;         /* 4 (DNA) or 5..64 (e.g. 61 codons)               */
  uint32_t scale_ln;       /* calculator/mod.rs:108,149-155 (default 30); the
                              F64_PLAIN path scales by e^scale_ln.  The FP32
                              and FP16 paths scale by exact powers of two and
                              only check it is nonzero.                       */
  uint32_t tile_patterns;  /* 64-pattern rows per workgroup x 64; 0 = tuned  */
  uint32_t lane_patterns;  /* patterns per lane (1 or 2); 0 = tuned           */
  uint32_t lane_states;    /* FP32 partials: states per lane, 4 (a lane holds
                              a pattern's four states) or 1 (four lanes per
                              pattern, 16 patterns per wavefront: for pattern
                              counts too small to fill the chip); 0 = 4.
                              Values are bit-identical either way.           */
  uint32_t reserved[1];
} b3h_config;

typedef struct b3h_calc b3h_calc;

/* Fill *cfg with defaults: device 0, B3H_F32_SCALED, 4 states, scale_ln 30. */
void b3h_config_default(b3h_config *cfg);

/*
 * Create a calculator.  Replaces CalculatorConfig::make4 (calculator/mod.rs:
 * 159-200) and Cpu4Calculator::new (calculator/cpu4.rs:108-165).
 *   samples: [num_leaves x num_patterns] bytes, row-major by leaf, as
 *            deduplicate() produces them (likelihood/mod.rs:124-129).  For
 *            4 states these are IUPAC bit codes: A=1 C=2 G=4 T=8; any other
 *            value is a fully unknown state (cpu4.rs:117-125).  For S > 4
 *            they are state indices 0..S-1; values >= S are unknown.
 *   num_categories: K rate categories evaluated by this one handle.
 */
b3h_status b3h_create(const b3h_config *cfg, const uint8_t *samples,
                      uint32_t num_leaves, uint32_t num_patterns,
                      uint32_t num_categories, b3h_calc **out);

b3h_status b3h_destroy(b3h_calc *calc);

/* Pattern weights used by b3h_total (likelihood/mod.rs:134-139 and
 * gamma.rs:154-166).  Default: all ones. */
b3h_status b3h_set_weights(b3h_calc *calc, const uint32_t *weights);

/*
 * Calculator::propose (calculator/mod.rs:56-60; cpu4.rs:39-76).
 *   internals[n], children[2n]: DatedTree::partials_lists() (tree.rs:456-461):
 *       dirty internal nodes, children before parents, root last; each
 *       child pair sorted ascending.
 *   matrices: [K][2n][S][S] f64 row-major, category-major; within a
 *       category, Transitions::write_matrices(children) order
 *       (transitions.rs:62-72).
 *   frequencies: [S] f64 (Transitions::frequencies, transitions.rs:50-52).
 * Enqueues the work and returns; results are read by b3h_likelihoods or
 * b3h_total.
 */
b3h_status b3h_propose(b3h_calc *calc, uint32_t n, const uint32_t *internals,
                       const uint32_t *children, const double *matrices,
                       const double *frequencies);

/*
 * Substitution model on the device: the eigensystem of Q is computed on the
 * GPU in FP64 (replaces the host-side SubstitutionModel::update and
 * DefaultEigenSystem::decompose_matrix, substitution.rs:336-388,502-555 and
 * beast_eigen.rs:66-101).
 *   q: [S][S] f64, a reversible rate matrix normalised to mean rate 1.
 *   frequencies: [S] f64 stationary distribution.
 *   flags: B3H_MODEL_SPECIAL_CASES applies HKY/GTR's t > 20 -> rows = pi and
 *          t < 1e-10 -> I + Qt with unit diagonal (substitution.rs:245-268,
 *          408-414, 576-582).
 */
#define B3H_MODEL_SPECIAL_CASES 1u
b3h_status b3h_set_model(b3h_calc *calc, const double *q,
                         const double *frequencies, uint32_t flags);

/*
 * propose() with P(t) built on the GPU from the device eigensystem
 * (replaces Transitions::update + write_matrices, transitions.rs:26-72):
 *   edge_times[2n]: time lengths of the child edges in children order
 *       (tree.rs:732-736, h_parent - h_child).
 *   category_rates[K]: clock_rate x r_c per category (gamma.rs:126-134).
 *   The branch length of child m in category c is edge_times[m] *
 *   category_rates[c], exactly as transitions.rs:34-35 forms it.
 */
b3h_status b3h_propose_times(b3h_calc *calc, uint32_t n,
                             const uint32_t *internals,
                             const uint32_t *children,
                             const double *edge_times,
                             const double *category_rates);

/*
 * Per-category models and branch lengths: HeteroLikelihood and relaxed
 * clocks on one handle.
 *
 * b3's HeteroLikelihood (likelihood/hetero.rs:97-212) keeps one calculator,
 * one substitution model and one clock per category, and each pattern's
 * class (a ClassVector) picks which category's ln L enters the total.  Here
 * the K categories of one handle play those calculators, so one launch
 * evaluates them all and the class-select sum runs on the device.
 *
 * b3h_set_category_model: the device eigensystem and root frequencies of
 *   category k alone (4 states; FP32 and FP64 paths).  The first call
 *   gives every other category the current shared model (or this one, if
 *   none was set); b3h_set_model returns to one model for all categories.
 *   b3h_propose (host matrices) still takes one frequency vector.
 */
b3h_status b3h_set_category_model(b3h_calc *calc, uint32_t category,
                                  const double *q, const double *frequencies,
                                  uint32_t flags);

/*
 * propose with P built on the GPU from branch lengths the caller formed:
 *   lengths[K][2n]: category-major, children order within a category; the
 *   length of child m in category c as Transitions::update forms it
 *   (transitions.rs:33-35): time length x the clock's rate for that edge,
 *   which for a relaxed clock is the rate of the edge's rate category
 *   (clock.rs:115-119), times the category rate for GammaLikelihood
 *   (gamma.rs:131-133).
 */
b3h_status b3h_propose_lengths(b3h_calc *calc, uint32_t n,
                               const uint32_t *internals,
                               const uint32_t *children,
                               const double *lengths);

/*
 * Per-pattern classes (hetero.rs:150-160): with classes set, the total is
 * sum_i w_i lnL_{classes[i]}(i) and b3h_site_mixed returns those values;
 * null returns to the category mean (GammaLikelihood).  classes[P], each
 * < K.  If a result exists (pending, else accepted) its mixed values and
 * total are recomputed from its site values without a propose: the
 * hetero.rs:117-120,150-160 path when only the classes changed.
 */
b3h_status b3h_set_classes(b3h_calc *calc, const uint32_t *classes);

/*
 * CudaPropagationsCalculator::propose (calculator/cuda_propagations/mod.rs:
 * 77-141), for handles created with a propagation precision.
 *   nodes[num_nodes], children[num_nodes - num_leaves][2], num_leaves:
 *       DatedTree::propagation_lists() (tree.rs:436-454): the updated
 *       leaves, then the internal nodes whose projections change, children
 *       first, the root last; child pairs sorted.
 *   matrices: [K][num_nodes - 1][4][4] f64, one per listed node but the
 *       root, Transitions::write_matrices(&nodes[..len - 1]) order.
 *   frequencies: [4].
 * The _times form builds the matrices on the GPU from the device model:
 *   edge_times[num_nodes - 1] (the time length of each listed node's edge)
 *   times category_rates[K].
 * Other propose calls return B3H_ERR_UNSUPPORTED on such handles, and these
 * return it on the others.
 */
b3h_status b3h_propose_propagations(b3h_calc *calc, uint32_t num_nodes,
                                    const uint32_t *nodes, uint32_t num_leaves,
                                    const uint32_t *children,
                                    const double *matrices,
                                    const double *frequencies);
b3h_status b3h_propose_propagations_times(b3h_calc *calc, uint32_t num_nodes,
                                          const uint32_t *nodes,
                                          uint32_t num_leaves,
                                          const uint32_t *children,
                                          const double *edge_times,
                                          const double *category_rates);

/* Calculator::accept / reject (calculator/mod.rs:62-69; cpu4.rs:78-89).
 * Results are kept in two slots like the likelihood layer's buffer
 * (hetero.rs:30,170-212): after accept the proposal's results are current;
 * after reject the previously accepted results are. */
b3h_status b3h_accept(b3h_calc *calc);
b3h_status b3h_reject(b3h_calc *calc);

/* Calculator::likelihoods (calculator/mod.rs:71-72): per-pattern ln L of
 * category k, net of scaling, of the pending propose (else of the accepted
 * state).  Blocks until the last propose finishes. */
b3h_status b3h_likelihoods(b3h_calc *calc, uint32_t category, double *out);

/* Per-pattern mixed value LSE_c(lnL_c) + ln(1/K) (gamma.rs:159-163); for
 * K = 1 this equals b3h_likelihoods(calc, 0, ...). */
b3h_status b3h_site_mixed(b3h_calc *calc, double *out);

/* The likelihood layer's total: sum_i w_i * mixed_i, reduced in FP64 on the
 * device (dna.rs:96-97, gamma.rs:156-166).  Blocks until the last propose
 * has produced it: the reducing kernel stores the value and then a sequence
 * number in pinned host memory, and this call polls for the number (after
 * 20 ms it waits for the stream instead, which also reports launch errors).
 * No copy is issued. */
b3h_status b3h_total(b3h_calc *calc, double *out);

/* Calculator::num_patterns (calculator/mod.rs:74-75). */
uint32_t b3h_num_patterns(const b3h_calc *calc);

/* Message of the last failed call on this thread ("" if none). */
const char *b3h_last_error(void);

/* Number of visible HIP devices (0 when none; never fails). */
int b3h_device_count(void);

/* Kernel timing for measurement: when enabled, every pruning-kernel launch
 * is bracketed by HIP events on the handle's stream.  read() synchronises
 * and returns the summed kernel time and the launch count since enable. */
b3h_status b3h_profile_enable(b3h_calc *calc, int enable);
b3h_status b3h_profile_read(b3h_calc *calc, double *kernel_ms,
                            uint64_t *launches);
/* The same launches one by one: ms[i] is the i-th timed launch since enable
 * (up to `capacity` values); *launches is the count.  For medians. */
b3h_status b3h_profile_launch_times(b3h_calc *calc, double *ms,
                                    uint64_t capacity, uint64_t *launches);

/* Discrete-gamma category rates (likelihood/gamma.rs:205-225): the
 * Gamma(alpha, 1) quantiles at (2i+1)/(2K), i = 0..K-1, divided by their
 * mean.  Host-side FP64; K = 1 gives rate 1. */
b3h_status b3h_gamma_rates(double alpha, uint32_t num_categories,
                           double *rates);

/* Bytes of device memory held by the handle. */
uint64_t b3h_device_bytes(const b3h_calc *calc);

/* ------------------------------------------------------------------------
 * Batched evaluation: many independent instances (trees, alignments,
 * models) of uneven size evaluated by one persistent kernel launch, which
 * takes (instance, pattern tile) work items from a device queue, longest
 * trees first.  b3 has no batched calculator; this serves many chains or
 * proposals at once.  Every instance has its own update list, EditBuf slots,
 * model and weights; all share K and the configuration.
 * ---------------------------------------------------------------------- */
typedef struct b3h_batch b3h_batch;

typedef struct b3h_instance {
  const uint8_t *samples;   /* [num_leaves x num_patterns] codes, as b3h_create */
  uint32_t num_leaves;
  uint32_t num_patterns;
  const uint32_t *weights;  /* [num_patterns] or NULL (all ones)            */
} b3h_instance;

typedef struct b3h_update {
  uint32_t n;                      /* list length; 0 = instance not updated  */
  const uint32_t *internals;       /* [n]   as b3h_propose_times            */
  const uint32_t *children;        /* [2n]                                   */
  const double *edge_times;        /* [2n]                                   */
  const double *category_rates;    /* [K]                                    */
} b3h_update;

/* flags for b3h_batch_create */
#define B3H_BATCH_SINGLE_SLOT 1u  /* one partial slot per node: half the
                                     memory; b3h_batch_reject is unsupported */

b3h_status b3h_batch_create(const b3h_config *cfg, uint32_t count,
                            const b3h_instance *instances,
                            uint32_t num_categories, uint32_t flags,
                            b3h_batch **out);
b3h_status b3h_batch_destroy(b3h_batch *batch);
/* Model of one instance (see b3h_set_model). */
b3h_status b3h_batch_set_model(b3h_batch *batch, uint32_t instance,
                               const double *q, const double *frequencies,
                               uint32_t flags);
/* Evaluate every instance whose update has n > 0, in one launch.  With
 * B3H_BATCH_GRAPH set in `flags`, the launch sequence is captured into a HIP
 * graph and replayed while the update shapes repeat. */
#define B3H_BATCH_GRAPH 1u
b3h_status b3h_batch_propose_times(b3h_batch *batch, const b3h_update *updates,
                                   uint32_t flags);
/* Per-instance totals of the last evaluation (blocks). */
b3h_status b3h_batch_totals(b3h_batch *batch, double *out);
b3h_status b3h_batch_likelihoods(b3h_batch *batch, uint32_t instance,
                                 uint32_t category, double *out);
b3h_status b3h_batch_accept(b3h_batch *batch);
b3h_status b3h_batch_reject(b3h_batch *batch);
b3h_status b3h_batch_profile_enable(b3h_batch *batch, int enable);
b3h_status b3h_batch_profile_read(b3h_batch *batch, double *kernel_ms,
                                  uint64_t *launches);
uint64_t b3h_batch_device_bytes(const b3h_batch *batch);

#ifdef __cplusplus
}
#endif
#endif
