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

import ctypes as C
import os

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


_SIM = None


def _simlib():
    global _SIM
    if _SIM is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libb3sim.so")
        lib = C.CDLL(path)
        u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
        lib.b3sim_states.restype = C.c_int
        lib.b3sim_states.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, u32p, u32p, f64p, f64p,
                                     C.c_uint64, C.c_uint64, C.c_uint32, u8p]
        _SIM = lib
    return _SIM


def reversible_transitions(q, freqs, times):
    """P(t) for many t at once from the eigensystem of the symmetrised
    reversible Q: returns [len(times), S, S]."""
    q = np.asarray(q, np.float64)
    pi = np.asarray(freqs, np.float64)
    d = np.sqrt(pi)
    Ssym = (d[:, None] * q / d[None, :])
    Ssym = 0.5 * (Ssym + Ssym.T)
    w, U = np.linalg.eigh(Ssym)
    V = U / d[:, None]
    Vi = U.T * d[None, :]
    ex = np.exp(np.outer(np.asarray(times, np.float64), w))
    return np.einsum("ik,tk,kj->tij", V, ex, Vi)


def simulate_states_fast(parents, children, heights, q, freqs, rates, num_sites, seed,
                         threads=None):
    """Same process as simulate_states (root ~ freqs, per-site category
    uniform over ``rates``, child ~ row of P(t r_c)) in the multithreaded
    host simulator; returns uint8 states [L, num_sites]."""
    parents = np.asarray(parents)
    N = parents.size
    L = (N + 1) // 2
    S = len(freqs)
    K = len(rates)
    root = int(np.nonzero(parents == TOMBSTONE)[0][0])
    lengths = np.zeros(N)
    nonroot = np.arange(N) != root
    lengths[nonroot] = heights[parents[nonroot].astype(np.int64)] - heights[nonroot]
    t = (lengths[:, None] * np.asarray(rates)[None, :]).reshape(-1)
    P = np.clip(reversible_transitions(q, freqs, t), 0.0, None)
    cum = np.ascontiguousarray(np.cumsum(P, axis=2).reshape(-1))
    order = np.ascontiguousarray(sorted(range(L, N), key=lambda n: -heights[n]), np.uint32)
    out = np.empty((L, num_sites), np.uint8)
    rc = _simlib().b3sim_states(L, S, K, order,
                                np.ascontiguousarray(children, np.uint32).reshape(-1), cum,
                                np.cumsum(np.asarray(freqs, np.float64)), int(num_sites),
                                int(seed) & (2**64 - 1), int(threads or os.cpu_count() or 1),
                                out.reshape(-1))
    if rc != 0:
        raise ValueError("b3sim_states rejected its arguments")
    return out
