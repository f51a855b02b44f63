"""CPU FP64 MCMC chain over the oracle (test infrastructure and the CPU
baseline of the 16-taxon HKY chain; never part of the product).

A restatement of b3's step for a tree-only chain with a fixed substitution
model:
  - Mcmc::step / finalize_step (crates/b3/src/mcmc.rs:314-393): Abort when
    the Hastings ratio is -inf, PriorReject when the prior is -inf, Reject
    when ln L is -inf, accept iff new_post - old_post + hastings > ln U;
    then accept or reject the tree and the likelihood;
  - NodeSlide (python/aspartik/b3/operators/_node_slide.py:26-39),
    TreeScale (_tree_scale.py:27-37, ratio ln(s) (internals - 2)) and
    NarrowExchange (_tree_exchange.py:23-57, DatedTree::swap_parents
    tree.rs:535-549, is_grandparent / num_grandparents tree.rs:767-778) at
    b3's default weights 4 : 1 : 2;
  - the prior: Yule with birth rate lambda as the native chain states it,
    (L - 1) ln lambda - lambda * tree length, -inf for a non-positive edge;
  - narrow_reversible=True: the exchange's grandparent is an internal node
    with at least one internal child (BEAST2's narrow Exchange), the
    reversible form of b3's rule (two internal children), as the native
    chain's setting of the same name.
The likelihood is the oracle's FP64 GammaLikelihood (cpu4.rs restated) with
every node recomputed at each step, as the measured configuration asks.
"""
from __future__ import annotations

import math
import random
import time

import numpy as np

from . import oracle as O

TOMBSTONE = 0xFFFFFFFF


class Chain:
    def __init__(self, num_leaves, parents, children, heights, model, codes, weights,
                 birth_rate=1.0, seed=1, op_weights=(4.0, 1.0, 2.0), tree_scale_factor=0.75,
                 narrow_reversible=False):
        self.L = int(num_leaves)
        self.N = 2 * self.L - 1
        self.parents = [int(p) for p in np.asarray(parents)]
        self.children = [tuple(int(c) for c in pair)
                         for pair in np.asarray(children).reshape(-1, 2)]
        self.heights = [float(x) for x in np.asarray(heights)]
        self.root = self.parents.index(TOMBSTONE)
        self.tree = O.Tree(self.L, parents, children, heights)
        codes = np.ascontiguousarray(codes, np.uint8)
        self.lik = O.GammaLikelihood(self.tree, model, codes, codes.shape[1],
                                     np.ascontiguousarray(weights, np.uint32), num_categories=1,
                                     threads_per_category=1)
        self.lam = float(birth_rate)
        self.rng = random.Random(seed)
        self.w = [float(x) for x in op_weights]
        self.factor = float(tree_scale_factor)
        self.loose = bool(narrow_reversible)
        self.tree.mark_all_edges_updated()
        self.loglik = self.lik.likelihood()
        self.lik.accept()
        self.tree.accept()
        self.prior = self.log_prior()
        self.posterior = self.loglik + self.prior

    # ---- tree mirror -------------------------------------------------------
    def _internal(self, v):
        return v >= self.L

    def _is_grandparent(self, v):
        if not self._internal(v):
            return False
        a, b = self.children[v - self.L]
        if self.loose:
            return self._internal(a) or self._internal(b)
        return self._internal(a) and self._internal(b)

    def _num_grandparents(self):
        return sum(1 for v in range(self.L, self.N) if self._is_grandparent(v))

    def _set_height(self, v, h):
        self.heights[v] = h
        self.tree.set_height(v, h)

    def _replace_child(self, parent, old, new):
        a, b = self.children[parent - self.L]
        pair = (new, b) if a == old else (a, new)
        self.children[parent - self.L] = tuple(sorted(pair))
        self.parents[new] = parent
        self.tree.replace_child(parent, old, new)

    def tree_length(self):
        return sum(self.heights[self.parents[v]] - self.heights[v]
                   for v in range(self.N) if v != self.root)

    def log_prior(self):
        for v in range(self.N):
            if v != self.root and not (self.heights[self.parents[v]] - self.heights[v] > 0.0):
                return -math.inf
        return (self.L - 1) * math.log(self.lam) - self.lam * self.tree_length()

    # ---- operators ---------------------------------------------------------
    def _node_slide(self):
        while True:
            v = self.rng.randrange(self.L, self.N)
            if v != self.root:
                break
        a, b = self.children[v - self.L]
        upper = self.heights[self.parents[v]]
        lower = max(self.heights[a], self.heights[b])
        self._set_height(v, (upper - lower) * self.rng.random() + lower)
        return 0.0

    def _tree_scale(self):
        f = self.factor
        s = f + (1.0 / f - f) * self.rng.random()
        for v in range(self.L, self.N):
            self._set_height(v, self.heights[v] * s)
        return math.log(s) * ((self.L - 1) - 2)

    def _narrow_exchange(self):
        before = self._num_grandparents()
        if before == 0:
            return -math.inf
        while True:
            gp = self.rng.randrange(self.L, self.N)
            if self._is_grandparent(gp):
                break
        left, right = self.children[gp - self.L]
        if self.heights[left] > self.heights[right]:
            parent, uncle = left, right
        elif self.heights[right] > self.heights[left]:
            parent, uncle = right, left
        else:
            return -math.inf
        if not self._internal(parent):
            return -math.inf
        pre = int(self._is_grandparent(parent)) + int(self._is_grandparent(uncle))
        pc = self.children[parent - self.L]
        child = pc[0] if self.rng.random() < 0.5 else pc[1]
        # swap_parents(uncle, child): uncle's parent is gp, child's is parent
        self._replace_child(gp, uncle, child)
        self._replace_child(parent, child, uncle)
        post = int(self._is_grandparent(parent)) + int(self._is_grandparent(uncle))
        return math.log(before / (before - pre + post))

    # ---- steps -------------------------------------------------------------
    def run(self, steps, trace=None):
        """Run `steps` steps; returns a dict of counts and the seconds taken.
        With a list `trace`, the node heights after every step are appended."""
        ops = (self._node_slide, self._tree_scale, self._narrow_exchange)
        wsum = sum(self.w)
        stats = dict(steps=0, accepted=0, aborts=0, prior_rejects=0, exact_evals=0,
                     op_proposed=[0, 0, 0], op_accepted=[0, 0, 0])
        t0 = time.perf_counter()
        for _ in range(int(steps)):
            u = self.rng.random() * wsum
            op = 0
            while op < 2 and u >= self.w[op]:
                u -= self.w[op]
                op += 1
            stats["op_proposed"][op] += 1
            saved = (list(self.parents), list(self.children), list(self.heights))
            hastings = ops[op]()
            accept = False
            new_prior = new_ll = None
            if hastings == -math.inf:
                stats["aborts"] += 1
            else:
                new_prior = self.log_prior()
                if new_prior == -math.inf:
                    stats["prior_rejects"] += 1
                else:
                    self.tree.mark_all_edges_updated()
                    new_ll = self.lik.likelihood()
                    stats["exact_evals"] += 1
                    if new_ll != -math.inf:
                        ratio = new_ll + new_prior - self.posterior + hastings
                        accept = ratio > math.log(self.rng.random())
            if accept:
                self.lik.accept()
                self.tree.accept()
                self.loglik, self.prior = new_ll, new_prior
                self.posterior = new_ll + new_prior
                stats["accepted"] += 1
                stats["op_accepted"][op] += 1
            else:
                if new_ll is not None:
                    self.lik.reject()
                self.tree.reject()
                self.parents, self.children, self.heights = saved
            stats["steps"] += 1
            if trace is not None:
                trace.append(tuple(self.heights))
        stats["seconds"] = time.perf_counter() - t0
        return stats

    def state(self):
        """(parents, children [L-1, 2], heights) of the current state."""
        p = np.array(self.parents, np.uint32)
        c = np.array(self.children, np.uint32).reshape(-1, 2)
        return p, c, np.array(self.heights)
