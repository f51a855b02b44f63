"""Molecular clocks: per-edge substitution rates for the P(t) build.

Mirrors crates/b3/src/clock.rs:
* ``Clock.Strict(rate)`` (clock.rs:11-39): one rate for every edge; a
  changed rate marks every edge.
* ``Clock.Relaxed(rate_categories, distribution)`` (clock.rs:41-134): edge
  e's rate is rates[rate_categories[e]], where rates[i] is the
  distribution's quantile at (i + 1/2) / n for n categories, accumulated
  as b3 does (quantile = step / 2, then += step).  A changed distribution
  recomputes the rates and marks every edge; changed category indices mark
  those edges; reject after a distribution change recomputes the rates
  from the restored parameters.
"""
from __future__ import annotations

import numpy as np

from .parameters import ClassVector


class Clock:
    def __init__(self):
        raise TypeError("use Clock.Strict(rate) or Clock.Relaxed(rate_categories, distribution)")

    @classmethod
    def Strict(cls, rate) -> "StrictClock":  # noqa: N802 - b3's constructor name
        return StrictClock(rate)

    @classmethod
    def Relaxed(cls, rate_categories: ClassVector, distribution) -> "RelaxedClock":  # noqa: N802
        return RelaxedClock(rate_categories, distribution)


class StrictClock:
    strict = True

    def __init__(self, rate):
        self.rate = rate
        self._cached = float(rate)

    def update(self, tree) -> None:
        """clock.rs:17-24"""
        changed = getattr(self.rate, "is_changed", lambda: False)()
        if changed:
            self._cached = float(self.rate)
            tree.mark_all_edges_updated()

    def get_rate(self, edge=None) -> float:
        return self._cached

    def edge_rates(self, num_nodes: int) -> np.ndarray:
        return np.full(num_nodes, self._cached)

    def accept(self) -> None:
        pass

    def reject(self) -> None:
        """clock.rs:36-38: re-read the (restored) rate."""
        self._cached = float(self.rate)


class RelaxedClock:
    strict = False

    def __init__(self, rate_categories: ClassVector, distribution):
        self.rate_categories = rate_categories
        self.distribution = distribution
        self._updated = False
        self.rates = np.zeros(rate_categories.num_classes)
        self._update_rates()

    def _update_rates(self) -> None:
        """clock.rs:66-85"""
        n = self.rate_categories.num_classes
        step = 1.0 / n
        quantile = step / 2.0
        for i in range(n):
            self.rates[i] = float(self.distribution.inverse_cdf(quantile))
            quantile += step

    def update(self, tree) -> None:
        """clock.rs:87-113"""
        if self.distribution.is_changed():
            self._update_rates()
            self._updated = True
            tree.mark_all_edges_updated()
            return
        for i in self.rate_categories.changed_indices():
            tree.mark_edge_updated(i)

    def get_rate(self, edge: int) -> float:
        """clock.rs:115-119"""
        return float(self.rates[self.rate_categories[edge]])

    def edge_rates(self, num_nodes: int) -> np.ndarray:
        cats = self.rate_categories.values[:num_nodes].astype(np.int64)
        return self.rates[cats]

    def accept(self) -> None:
        self._updated = False

    def reject(self) -> None:
        if self._updated:
            self._update_rates()
            self._updated = False
