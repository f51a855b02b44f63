"""Workload generators: Yule trees and the sequence simulator.  The
simulator is pinned by the exact site-pattern probabilities of a 3-taxon
tree (the FP64 oracle's mixture likelihood of each of the 64 patterns)."""
import numpy as np
import scipy.linalg

from aspartik_amd import sim
from aspartik_amd.b3 import HKY, gamma_rates
from oracle import oracle as O


def test_yule_tree_structure():
    rng = np.random.default_rng(0)
    for L in [2, 3, 17, 200]:
        par, ch, h = sim.yule_tree(L, rng)
        N = 2 * L - 1
        assert (par == sim.TOMBSTONE).sum() == 1 and par[N - 1] == sim.TOMBSTONE
        for n in range(L, N):
            a, b = ch[n - L]
            assert par[a] == n and par[b] == n and h[n] > h[a] and h[n] > h[b]
        lengths = h[par[:-1].astype(np.int64)] - h[:-1]
        assert abs(lengths.mean() - 0.1) < 1e-12
        assert (lengths > 0).all()


def test_reversible_transitions_match_expm():
    m = HKY([0.1, 0.2, 0.3, 0.4], 3.0)
    q = m.q_matrix()
    ts = np.array([1e-6, 0.05, 0.7, 4.0])
    P = sim.reversible_transitions(q, m.frequencies(), ts)
    for i, t in enumerate(ts):
        np.testing.assert_allclose(P[i], scipy.linalg.expm(q * t), atol=1e-13)


def test_simulator_matches_exact_pattern_probabilities():
    rng = np.random.default_rng(1)
    L = 3
    par, ch, h = sim.yule_tree(L, rng, mean_branch=0.4)
    f = np.array([0.1, 0.2, 0.3, 0.4])
    m = HKY(f, 2.5)
    K = 4
    rates = gamma_rates(0.5, K)
    n = 2_000_000
    states = sim.simulate_states_fast(par, ch, h, m.q_matrix(), f, rates, n, seed=7, threads=4)
    idx = states[0].astype(np.int64) * 16 + states[1] * 4 + states[2]
    counts = np.bincount(idx, minlength=64)
    # exact probability of each pattern under the mixture (oracle likelihood)
    codes = np.array([[1 << (p // 16), 1 << ((p // 4) % 4), 1 << (p % 4)] for p in range(64)],
                     np.uint8).T.copy()
    tree = O.Tree(L, par, ch, h)
    lik = O.GammaLikelihood(tree, O.Model.hky(2.5, f), codes, 64, np.ones(64, np.uint32),
                            num_categories=K, alpha=0.5)
    site = np.array([lik.site(c) for c in range(K)])
    prob = np.exp(site).mean(0)
    assert abs(prob.sum() - 1.0) < 1e-12
    expect = prob * n
    z = (counts - expect) / np.sqrt(expect * (1 - prob))
    assert np.abs(z).max() < 5.5, np.abs(z).max()
    chi2 = ((counts - expect) ** 2 / expect).sum()
    assert chi2 < 63 + 6 * np.sqrt(2 * 63)
    # the stream does not depend on the thread count
    again = sim.simulate_states_fast(par, ch, h, m.q_matrix(), f, rates, 10_000, seed=7, threads=1)
    np.testing.assert_array_equal(again, states[:, :10_000])
