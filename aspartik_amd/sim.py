"""Seeded synthetic inputs: Yule trees, sequences simulated along a tree,
unique-pattern alignments and pattern weights.

b3 itself has no tree or sequence simulator (its MSA.random draws iid
bases, crates/data/src/msa/mod.rs:182-191); these generators make the
workloads of the measured configurations.  Trees use b3's node numbering
(crates/b3/src/parameters/tree.rs:653-743): leaves 0..L-1, internal nodes
L..2L-2, the edge above node n has index n, children stored as pairs in
internal-node order, parents[root] = TOMBSTONE.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg

TOMBSTONE = 0xFFFFFFFF


def yule_tree(num_leaves: int, rng: np.random.Generator, birth_rate: float = 1.0,
              mean_branch: float = 0.1):
    """Yule (pure-birth) tree.

    Backwards in time, while k lineages remain the waiting time to the next
    split is Exp(k * birth_rate) and the two lineages that merge are chosen
    uniformly.  Heights are then rescaled so that the mean branch length is
    ``mean_branch``.  Returns (parents[N] u32, children[L-1, 2] u32,
    heights[N] f64); the root is node 2L-2.
    """
    L = int(num_leaves)
    if L < 2:
        raise ValueError("need at least 2 leaves")
    N = 2 * L - 1
    parents = np.full(N, TOMBSTONE, np.uint32)
    children = np.zeros((L - 1, 2), np.uint32)
    heights = np.zeros(N, np.float64)
    lineages = list(range(L))
    t = 0.0
    nxt = L
    while len(lineages) > 1:
        k = len(lineages)
        t += rng.exponential(1.0 / (k * birth_rate))
        i, j = rng.choice(k, 2, replace=False)
        a, b = lineages[i], lineages[j]
        node = nxt
        nxt += 1
        heights[node] = t
        parents[a] = node
        parents[b] = node
        children[node - L] = (a, b)
        for idx in sorted((i, j), reverse=True):
            lineages.pop(idx)
        lineages.append(node)
    lengths = heights[parents[:-1].astype(np.int64)] - heights[:-1]
    heights *= mean_branch / lengths.mean()
    return parents, children, heights


def simulate_states(parents, children, heights, transition, freqs, rates, num_sites,
                    rng: np.random.Generator, categories=None):
    """Simulate states at the leaves along the tree.

    The root state is drawn from ``freqs``; each site draws a rate category
    uniformly from ``rates`` (unless ``categories`` is given); a child's
    state is drawn from row (parent state) of transition(t * r_c).  Returns
    int16 states [L, num_sites] and the site categories.
    """
    N = parents.size
    L = (N + 1) // 2
    S = len(freqs)
    K = len(rates)
    if categories is None:
        categories = rng.integers(0, K, size=num_sites)
    states = np.zeros((N, num_sites), np.int16)
    root = N - 1 if parents[N - 1] == TOMBSTONE else int(np.nonzero(parents == TOMBSTONE)[0][0])
    cf = np.cumsum(freqs)
    states[root] = np.minimum(np.searchsorted(cf, rng.random(num_sites) * cf[-1], side="right"), S - 1)
    order = sorted(range(L, N), key=lambda n: -heights[n])  # parents before children
    cat_masks = [np.nonzero(categories == c)[0] for c in range(K)]
    for node in order:
        for child in children[node - L]:
            child = int(child)
            length = heights[node] - heights[child]
            out = np.empty(num_sites, np.int16)
            for c in range(K):
                idx = cat_masks[c]
                if idx.size == 0:
                    continue
                P = np.clip(transition(length * rates[c]), 0.0, None)
                cum = np.cumsum(P, axis=1)
                rows = cum[states[node, idx]]
                u = rng.random(idx.size) * rows[:, -1]
                out[idx] = np.minimum((u[:, None] >= rows).sum(1), S - 1)
            states[child] = out
    return states[:L], categories


def unique_columns(states, seen=None):
    """Columns of ``states`` [L, n] that are not in ``seen`` (a set of bytes),
    first occurrence order; updates ``seen``."""
    if seen is None:
        seen = set()
    keep = []
    cols = np.ascontiguousarray(states.T)
    for i in range(cols.shape[0]):
        key = cols[i].tobytes()
        if key not in seen:
            seen.add(key)
            keep.append(i)
    return np.asarray(keep, np.int64)


def simulate_unique_patterns(parents, children, heights, transition, freqs, rates,
                             num_patterns, rng: np.random.Generator, batch=None):
    """Simulate sites along the tree until ``num_patterns`` distinct columns
    exist; returns int16 states [L, num_patterns] in first-occurrence order."""
    L = (parents.size + 1) // 2
    seen = set()
    parts = []
    have = 0
    if batch is None:
        batch = max(1024, int(num_patterns * 1.1))
    rounds = 0
    while have < num_patterns:
        st, _ = simulate_states(parents, children, heights, transition, freqs, rates, batch, rng)
        keep = unique_columns(st, seen)
        take = keep[: num_patterns - have]
        parts.append(st[:, take])
        have += take.size
        rounds += 1
        if rounds > 1000:
            raise RuntimeError("could not reach the requested number of unique patterns")
    return np.concatenate(parts, axis=1)[:, :num_patterns].reshape(L, num_patterns)


def dna_codes(states):
    """State indices 0..3 -> b3's nucleotide bit codes A=1 C=2 G=4 T=8."""
    return (np.uint8(1) << states.astype(np.uint8)).astype(np.uint8)


def expm_transition(q):
    """P(t) = expm(Q t) (scipy) for simulation."""
    q = np.asarray(q, np.float64)
    return lambda t: scipy.linalg.expm(q * t)


def dirichlet(rng, alpha, n):
    return rng.dirichlet(np.full(n, float(alpha)))
