"""Shared test helpers: seeded trees and alignments, brute-force likelihoods."""
from __future__ import annotations

import itertools

import numpy as np

from aspartik_amd import sim
from oracle import oracle as O

TOMB = 0xFFFFFFFF


def random_tree(L, rng, mean_branch=0.1):
    return sim.yule_tree(L, rng, 1.0, mean_branch)


def random_codes(L, P, rng, unknown_frac=0.0, ambiguous=False):
    codes = (np.uint8(1) << rng.integers(0, 4, size=(L, P)).astype(np.uint8)).astype(np.uint8)
    if unknown_frac > 0:
        m = rng.random((L, P)) < unknown_frac
        vals = rng.choice([0, 15, 3, 5, 9, 6], size=m.sum()) if ambiguous else 15
        codes[m] = vals
    return codes


def model_matrices(model, tree: O.Tree, children, rates):
    """[K][2n][4][4] transition matrices for a children list, as b3's
    Transitions build them: P(length * rate_c)."""
    K = len(rates)
    n = children.shape[0]
    out = np.zeros((K, 2 * n, model.n, model.n))
    flat = children.reshape(-1)
    for c in range(K):
        for m in range(2 * n):
            out[c, m] = model.transition(tree.edge_length(int(flat[m])) * rates[c])
    return out


def tip_vector(code, S=4):
    """cpu4.rs:117-125: {1,2,4,8} one-hot, anything else all ones."""
    if S == 4:
        m = {1: 0, 2: 1, 4: 2, 8: 3}
        v = np.zeros(4)
        if code in m:
            v[m[code]] = 1.0
        else:
            v[:] = 1.0
        return v
    v = np.zeros(S)
    if code < S:
        v[code] = 1.0
    else:
        v[:] = 1.0
    return v


def brute_force_site_ll(parents, heights, codes_col, transition, freqs, S=4, edge_rates=None):
    """ln sum over all assignments of internal states of
    pi(root) prod_edges P(t)[parent, child]; tips fixed (unknown: summed).
    With edge_rates, edge n's length is its time length times edge_rates[n].
    Pure-Python enumeration, only for tiny trees."""
    N = parents.size
    L = (N + 1) // 2
    root = int(np.nonzero(parents == TOMB)[0][0])
    internals = list(range(L, N))
    Pm = {}
    for n in range(N):
        if n == root:
            continue
        t = heights[parents[n]] - heights[n]
        Pm[n] = transition(t if edge_rates is None else t * edge_rates[n])
    tipv = [tip_vector(int(codes_col[i]), S) for i in range(L)]
    total = 0.0
    for assign in itertools.product(range(S), repeat=len(internals)):
        st = dict(zip(internals, assign))
        prob = freqs[st[root]]
        for n in internals:
            if n == root:
                continue
            prob *= Pm[n][st[int(parents[n])], st[n]]
        for leaf in range(L):
            ps = st[int(parents[leaf])]
            prob *= float(Pm[leaf][ps] @ tipv[leaf])
        total += prob
    return np.log(total)
