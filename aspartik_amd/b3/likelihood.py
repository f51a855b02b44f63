"""DNA, gamma and heterotachy likelihoods over one HIP calculator handle.

Mirrors b3's ``DNALikelihood`` (crates/b3/src/likelihood/dna.rs:16-124),
``GammaLikelihood`` (likelihood/gamma.rs:19-225) and ``HeteroLikelihood``
(likelihood/hetero.rs:16-212): clock/substitution/alpha change -> mark all
edges; unchanged tree -> cached value; otherwise the dirty update list goes
to the calculator and the FP64 total comes back.  One handle evaluates all
K categories (b3 keeps K calculators, gamma.rs:57-66, hetero.rs:57-66) and
P(t) is built on the device from the device eigensystem(s).
"""
from __future__ import annotations

import numpy as np

from .calculator import Calculator, gamma_rates
from .clock import Clock
from .parameters import ClassVector
from .tree import DatedTree


def deduplicate(msa, device=None):
    """likelihood/mod.rs:74-132: unique columns in first-occurrence order,
    counts as weights, gap (0) -> 0b1111.  msa: [num_seqs, num_sites] u8 or
    an ``aspartik_amd.data.MSA``; device None: host, an int: that GPU."""
    from ..data import deduplicate as _dedup
    return _dedup(msa, device)


def _child_times(tree, children):
    ch = np.asarray(children).reshape(-1).astype(np.int64)
    return tree.heights[tree.parents[ch]] - tree.heights[ch], ch


class GammaLikelihood:
    """``GammaLikelihood(codes, weights, tree, substitution, num_categories,
    alpha, clock_rate, calculator, clock)``; K = 1 behaves as
    ``DNALikelihood``.  ``clock`` (a ``Clock.Strict`` or ``Clock.Relaxed``)
    replaces the fixed ``clock_rate`` when given."""

    def __init__(self, codes, weights, tree: DatedTree, substitution, num_categories: int = 4,
                 alpha: float = 0.5, clock_rate: float = 1.0,
                 calculator: Calculator | None = None, clock=None):
        codes = np.ascontiguousarray(codes, np.uint8)
        self.P = codes.shape[-1] if codes.ndim == 2 else codes.size // tree.num_leaves
        self.K = int(num_categories)
        self.tree = tree
        self.substitution = substitution
        self.alpha = float(alpha)
        self.clock_rate = float(clock_rate)
        self.clock = clock
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
        if self.clock is not None:
            self.clock.update(tree)  # gamma.rs:104
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
        self._launched = True
        if self.calculator.propagations:
            # b3's propagation calculator takes its own lists and one matrix
            # per listed node but the root (cuda_propagations/mod.rs:77-86)
            if self.clock is not None and not self.clock.strict:
                raise NotImplementedError("propagation storage takes a strict clock")
            nodes, pchildren, nl = tree.propagation_lists()
            times, _ = _child_times(tree, nodes[:-1])
            rate = self.clock_rate if self.clock is None else self.clock.get_rate()
            self.calculator.propose_propagations_times(nodes, pchildren, nl, times,
                                                       rate * self._rates)
            self._last = self.calculator.total()
            return self._last
        internals, children = tree.partials_lists()
        times, ch = _child_times(tree, children)
        if self.clock is None or self.clock.strict:
            # time x (clock rate x category rate), gamma.rs:128-134
            rate = self.clock_rate if self.clock is None else self.clock.get_rate()
            self.calculator.propose_times(internals, children, times, rate * self._rates)
        else:
            # per-edge clock rate: time x (rate of the edge x category rate)
            er = self.clock.edge_rates(tree.num_nodes)[ch]
            lengths = times[None, :] * (er[None, :] * self._rates[:, None])
            self.calculator.propose_lengths(internals, children, lengths)
        self._last = self.calculator.total()
        return self._last

    def site_likelihoods(self):
        return self.calculator.site_mixed()

    def accept(self):
        self._cache = self._last
        if self._launched:
            self.calculator.accept()
            if self.clock is not None:
                self.clock.accept()
        self._accepted = (self.alpha, self.clock_rate, self.substitution.version,
                          self._rates.copy())
        self._launched = False
        self._alpha_changed = False
        self._clock_changed = False

    def reject(self):
        if self._launched:
            self.calculator.reject()
            if self.clock is not None:
                self.clock.reject()
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


def DNALikelihood(codes, weights, tree, substitution, clock_rate=1.0, calculator=None,
                  clock=None):
    """dna.rs:32-66: one rate category."""
    return GammaLikelihood(codes, weights, tree, substitution, num_categories=1, alpha=1.0,
                           clock_rate=clock_rate, calculator=calculator, clock=clock)


class HeteroLikelihood:
    """``HeteroLikelihood(codes, weights, tree, categories, substitutions,
    clocks, calculator)`` (hetero.rs:39-96).

    Category c has its own substitution model and clock; ``categories`` is a
    ``ClassVector`` whose entry i picks the category whose ln L pattern i
    contributes (hetero.rs:150-160).  The K = len(clocks) categories are
    the K categories of one device handle: one launch evaluates them all,
    each with its own eigensystem, root frequencies and per-edge branch
    lengths, and the class-select weighted sum runs on the device.
    """

    def __init__(self, codes, weights, tree: DatedTree, categories: ClassVector,
                 substitutions, clocks, calculator: Calculator | None = None,
                 sites_to_patterns=None):
        codes = np.ascontiguousarray(codes, np.uint8)
        self.P = codes.shape[-1] if codes.ndim == 2 else codes.size // tree.num_leaves
        self.K = len(clocks)
        if len(substitutions) != self.K:
            raise ValueError("one substitution model per clock")
        if self.K < 1 or self.K > 16:
            raise ValueError("1..16 categories")
        if len(categories) < self.P or categories.num_classes > self.K:
            raise ValueError("categories must cover every pattern with classes < len(clocks)")
        self.tree = tree
        self.categories = categories
        self.substitutions = list(substitutions)
        self.clocks = list(clocks)
        self.sites_to_patterns = sites_to_patterns
        self.calculator = (calculator or Calculator.HIP()).make(codes, self.P, self.K)
        self.weights = np.ascontiguousarray(weights, np.uint32)
        self.calculator.set_weights(self.weights)
        self._versions = [None] * self.K
        for c in range(self.K):
            self._upload_model(c)
        self._classes = None
        self._sync_classes()
        self._cache = float("nan")
        self._last = float("nan")
        self._launched = False
        tree.mark_all_edges_updated()
        self.likelihood()
        self.accept()
        tree.accept()

    def _upload_model(self, c):
        s = self.substitutions[c]
        self.calculator.set_category_model(c, s.q_matrix(), s.frequencies(), s.special_cases)
        self._versions[c] = s.version

    def _sync_classes(self):
        """The device's classes follow the ClassVector (also after its reject)."""
        cls = self.categories.values[:self.P]
        if self._classes is None or not np.array_equal(cls, self._classes):
            self.calculator.set_classes(cls)
            self._classes = cls

    def likelihood(self) -> float:
        """hetero.rs:97-168"""
        tree = self.tree
        for clock in self.clocks:
            clock.update(tree)
        for c, s in enumerate(self.substitutions):
            if s.version != self._versions[c]:
                self._upload_model(c)
                tree.mark_all_edges_updated()
        if not tree.is_changed() and not self.categories.is_changed():
            self._last = self._cache
            return self._cache
        # classes first: without a propose this recomputes the total from
        # the current site values (the categories-only path)
        self._sync_classes()
        if tree.is_changed():
            internals, children = tree.partials_lists()
            times, ch = _child_times(tree, children)
            lengths = np.stack([times * clock.edge_rates(tree.num_nodes)[ch]
                                for clock in self.clocks])
            self._launched = True
            self.calculator.propose_lengths(internals, children, lengths)
        self._last = self.calculator.total()
        return self._last

    def site_likelihoods(self):
        """Per-pattern ln L of each pattern's class."""
        return self.calculator.site_mixed()

    def category_likelihoods(self, c):
        return self.calculator.likelihoods(c)

    def accept(self):
        """hetero.rs:170-190"""
        self._cache = self._last
        if self._launched:
            self.calculator.accept()
            for clock in self.clocks:
                clock.accept()
            self._launched = False

    def reject(self):
        """hetero.rs:192-212"""
        if self._launched:
            self.calculator.reject()
            for clock in self.clocks:
                clock.reject()
            self._launched = False
