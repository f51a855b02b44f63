"""DNA and gamma likelihoods over one HIP calculator handle.

Mirrors b3's ``DNALikelihood`` (crates/b3/src/likelihood/dna.rs:16-124) and
``GammaLikelihood`` (likelihood/gamma.rs:19-225): clock/substitution/alpha
change -> mark all edges; unchanged tree -> cached value; otherwise the
dirty update list goes to the calculator and the FP64 total comes back.
One handle evaluates all K rate categories (b3 keeps K calculators,
gamma.rs:57-66) and P(t) is built on the device from the device eigensystem.
"""
from __future__ import annotations

import numpy as np

from .calculator import Calculator, gamma_rates
from .tree import DatedTree


def deduplicate(msa):
    """likelihood/mod.rs:74-132: unique columns in first-occurrence order,
    counts as weights, gap (0) -> 0b1111.  msa: [num_seqs, num_sites] u8."""
    msa = np.ascontiguousarray(msa, np.uint8)
    cols = np.ascontiguousarray(msa.T)
    view = cols.view(np.dtype((np.void, cols.shape[1])))[:, 0]
    _, first, inverse, counts = np.unique(view, return_index=True, return_inverse=True,
                                          return_counts=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    leaves = msa[:, first[order]].copy()
    leaves[leaves == 0] = 0x0F
    return leaves, counts[order].astype(np.uint32), rank[inverse.reshape(-1)].astype(np.int64)


class GammaLikelihood:
    """``GammaLikelihood(codes, weights, tree, substitution, num_categories,
    alpha, clock_rate, calculator)``; K = 1 behaves as ``DNALikelihood``."""

    def __init__(self, codes, weights, tree: DatedTree, substitution, num_categories: int = 4,
                 alpha: float = 0.5, clock_rate: float = 1.0,
                 calculator: Calculator | None = None):
        codes = np.ascontiguousarray(codes, np.uint8)
        self.P = codes.shape[-1] if codes.ndim == 2 else codes.size // tree.num_leaves
        self.K = int(num_categories)
        self.tree = tree
        self.substitution = substitution
        self.alpha = float(alpha)
        self.clock_rate = float(clock_rate)
        self.calculator = (calculator or Calculator.HIP()).make(codes, self.P, self.K)
        self.weights = np.ascontiguousarray(weights, np.uint32)
        self.calculator.set_weights(self.weights)
        self._rates = self._category_rates(self.alpha)
        self._accepted = (self.alpha, self.clock_rate, substitution.version, self._rates.copy())
        self._model_version = None
        self._upload_model()
        self._cache = float("nan")
        self._last = float("nan")
        self._launched = False
        self._alpha_changed = False
        self._clock_changed = False
        tree.mark_all_edges_updated()
        self.likelihood()
        self.accept()
        tree.accept()

    def _category_rates(self, alpha):
        return gamma_rates(alpha, self.K) if self.K > 1 else np.ones(1)

    def _upload_model(self):
        s = self.substitution
        self.calculator.set_model(s.q_matrix(), s.frequencies(), s.special_cases)
        self._model_version = s.version

    # parameter edits (b3 tracks these with Parameter::is_changed)
    def set_alpha(self, alpha):
        self.alpha = float(alpha)
        self._alpha_changed = True

    def set_clock_rate(self, rate):
        self.clock_rate = float(rate)
        self._clock_changed = True

    def likelihood(self) -> float:
        """gamma.rs:97-169 / dna.rs:68-100"""
        tree = self.tree
        if self._clock_changed:
            tree.mark_all_edges_updated()
        if self.substitution.version != self._model_version:
            self._upload_model()
            tree.mark_all_edges_updated()
        if not (tree.is_changed() or self._alpha_changed):
            self._last = self._cache
            return self._cache
        if self._alpha_changed:
            tree.mark_all_edges_updated()
            self._rates = self._category_rates(self.alpha)
        internals, children = tree.partials_lists()
        times = tree.heights[tree.parents[children.reshape(-1).astype(np.int64)]] - \
            tree.heights[children.reshape(-1).astype(np.int64)]
        self._launched = True
        self.calculator.propose_times(internals, children, times, self.clock_rate * self._rates)
        self._last = self.calculator.total()
        return self._last

    def site_likelihoods(self):
        return self.calculator.site_mixed()

    def accept(self):
        self._cache = self._last
        if self._launched:
            self.calculator.accept()
        self._accepted = (self.alpha, self.clock_rate, self.substitution.version,
                          self._rates.copy())
        self._launched = False
        self._alpha_changed = False
        self._clock_changed = False

    def reject(self):
        if self._launched:
            self.calculator.reject()
        alpha, clock, _, rates = self._accepted
        self.alpha, self.clock_rate, self._rates = alpha, clock, rates.copy()
        self._launched = False
        self._alpha_changed = False
        self._clock_changed = False

    def restore_model(self, substitution):
        """After a rejected substitution move the caller restores the model
        parameters; the device eigensystem is rebuilt on the next evaluation
        (substitution.rs:428-430,598-603)."""
        self.substitution = substitution


def DNALikelihood(codes, weights, tree, substitution, clock_rate=1.0, calculator=None):
    """dna.rs:32-66: one rate category."""
    return GammaLikelihood(codes, weights, tree, substitution, num_categories=1, alpha=1.0,
                           clock_rate=clock_rate, calculator=calculator)
