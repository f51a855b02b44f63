"""Measured choice of storage layout and tile shape for one workload.

b3 picks its GPU calculator by configuration (calculator/mod.rs:94-156; the
propagation calculator is the default "cuda" backend).  Here the choice is
measured: every candidate handle (partial storage with each valid tile shape
and patterns per lane, and propagation storage) evaluates the caller's own
tree, alignment and model -- full evaluations (every edge dirty) and partial
ones (one random edge and its ancestors, as an MCMC move produces) -- and
the candidate with the lowest expected time per evaluation under the given
mix wins.  All candidates compute the same likelihood; only the schedule
and storage differ.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from .calculator import Calculator
from .tree import DatedTree


@dataclass
class Trial:
    name: str
    config: Calculator
    full_ms: float
    partial_ms: float
    score_ms: float
    full_kernel_us: float = 0.0
    partial_kernel_us: float = 0.0


def _lists(tree: DatedTree, propagations: bool):
    if propagations:
        nodes, children, nl = tree.propagation_lists()
        flat = nodes[:-1].astype(np.int64)
        return ("p", nodes, children, nl,
                tree.heights[tree.parents[flat]] - tree.heights[flat])
    internals, children = tree.partials_lists()
    flat = children.reshape(-1).astype(np.int64)
    return ("i", internals, children, tree.heights[tree.parents[flat]] - tree.heights[flat])


def _propose(calc, lst, rates):
    if lst[0] == "p":
        calc.propose_propagations_times(lst[1], lst[2], lst[3], lst[4], rates)
    else:
        calc.propose_times(lst[1], lst[2], lst[3], rates)


def default_candidates(num_categories: int, device: int = 0):
    """Partial storage with every tile that fits a 1024-thread workgroup: four
    states per lane with 1 or 2 patterns per lane, and one state per lane
    (four lanes per pattern); plus propagation storage."""
    out = []
    t = 1
    while 64 * num_categories * t <= 1024:
        for ppl in (1, 2):
            out.append((f"partials T={t} ppl={ppl}",
                        Calculator.HIP(device, tile_patterns=64 * t, lane_patterns=ppl)))
        out.append((f"partials T={t} states/lane=1",
                    Calculator.HIP(device, tile_patterns=64 * t, lane_states=1)))
        t *= 2
    out.append(("propagations", Calculator.HIPPropagations(device)))
    return out


def autotune(codes, weights, tree: DatedTree, substitution, category_rates,
             partial_fraction: float = 0.9, evals: int = 20, candidates=None, seed: int = 0,
             warm_s: float = 0.0):
    """Returns (best Calculator, [Trial, ...] sorted by score).

    partial_fraction: share of partial evaluations in the expected mix (an
    MCMC chain over a fixed alignment is mostly partial updates).
    warm_s: seconds of untimed full evaluations before each candidate's
    timing (at least three are always run), so that every candidate is
    timed at the device's steady clocks."""
    codes = np.ascontiguousarray(codes, np.uint8)
    P = codes.shape[1]
    rates = np.ascontiguousarray(category_rates, np.float64)
    K = rates.size
    rng = np.random.default_rng(seed)
    edges = [int(e) for e in rng.choice(tree.num_nodes, size=16) if int(e) != tree.root]
    trials = []
    for name, cfg in candidates or default_candidates(K):
        calc = cfg.make(codes, P, K)
        calc.set_weights(weights)
        calc.set_model(substitution.q_matrix(), substitution.frequencies(),
                       substitution.special_cases)
        t = tree.copy()
        t.mark_all_edges_updated()
        full = _lists(t, cfg.propagations)
        t.accept()
        parts = []
        for e in edges:
            t.mark_edge_updated(e)
            parts.append(_lists(t, cfg.propagations))
            t.accept()
        t_end = time.perf_counter() + warm_s
        nwarm = 0
        while nwarm < 3 or time.perf_counter() < t_end:
            _propose(calc, full, rates)
            calc.total()
            calc.accept()
            nwarm += 1
        calc.profile_enable(True)
        t0 = time.perf_counter()
        for _ in range(evals):
            _propose(calc, full, rates)
            calc.total()
            calc.accept()
        full_ms = (time.perf_counter() - t0) / evals * 1e3
        fk, fn = calc.profile_read()
        calc.profile_enable(True)
        t0 = time.perf_counter()
        for i in range(evals):
            _propose(calc, parts[i % len(parts)], rates)
            calc.total()
            calc.accept()
        partial_ms = (time.perf_counter() - t0) / evals * 1e3
        pk, pn = calc.profile_read()
        calc.profile_enable(False)
        calc.close()
        score = (1.0 - partial_fraction) * full_ms + partial_fraction * partial_ms
        trials.append(Trial(name, cfg, full_ms, partial_ms, score,
                            fk / max(1, fn) * 1e3, pk / max(1, pn) * 1e3))
    trials.sort(key=lambda tr: tr.score_ms)
    return trials[0].config, trials


# ---- batched evaluation ------------------------------------------------------

@dataclass
class BatchTrial:
    tile_patterns: int   # patterns per lane row of a workgroup's tile (64 x sub-tiles)
    lane_patterns: int   # patterns per lane
    lane_states: int     # states per lane (4, or 1: four lanes per pattern)
    ms: float            # per batch evaluation: captured-graph replay, synchronous
    kernel_ms: float     # the batched pruning kernel (HIP events)


def batch_candidates(num_categories: int):
    """Tile shapes of the batched kernel (workgroups of at most 512 threads)
    as (tile_patterns, lane_patterns, lane_states): every power-of-two
    number of wavefronts per category, four states per lane with one or two
    patterns per lane, and one state per lane."""
    t = max(1, 512 // (64 * num_categories))
    out = []
    while t >= 1:
        out += [(64 * t, 1, 4), (64 * t, 2, 4), (64 * t, 1, 1)]
        t //= 2
    return out


def autotune_batch(instances, models, updates, num_categories: int, evals: int = 8,
                   candidates=None, single_slot: bool = True, device: int = 0):
    """Measured tile choice for one batch: each candidate (tile_patterns,
    lane_patterns, lane_states) evaluates the caller's own instances -- one
    eager evaluation, then ``evals`` captured-graph replays -- and the
    fastest wins.  instances: [(codes, weights)]; models: [(q, frequencies,
    special_cases)] per instance; updates: one update per instance, as for
    ``Batch.propose_times``.  Returns ((tile_patterns, lane_patterns,
    lane_states), trials sorted fastest first).  Every shape computes the
    same per-site values; only the schedule differs."""
    from .batch import Batch
    trials = []
    for tp, lp, ls in (candidates or batch_candidates(num_categories)):
        b = Batch(instances, num_categories, device, single_slot, tile_patterns=tp,
                  lane_patterns=lp, lane_states=ls)
        try:
            for i, (q, f, special) in enumerate(models):
                b.set_model(i, q, f, special)
            b.propose_times(updates)
            b.totals()
            b.accept()
            b.propose_times(updates, graph=True)
            b.totals()
            b.accept()
            t0 = time.perf_counter()
            for _ in range(evals):
                b.propose_times(updates, graph=True)
                b.totals()
                b.accept()
            ms = (time.perf_counter() - t0) / evals * 1e3
            b.profile_enable(True)
            b.propose_times(updates)
            b.totals()
            b.accept()
            kms, kn = b.profile_read()
            b.profile_enable(False)
        finally:
            b.close()
        trials.append(BatchTrial(tp, lp, ls, ms, kms / max(1, kn)))
    trials.sort(key=lambda tr: tr.ms)
    best = trials[0]
    return (best.tile_patterns, best.lane_patterns, best.lane_states), trials
