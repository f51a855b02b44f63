"""Pins the oracle's calculators (crates/b3/src/calculator/cpu4.rs) and
likelihood layer (likelihood/mod.rs, gamma.rs) against brute-force
enumeration, closed forms and the invariants b3's own tests rely on."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import brute_force_site_ll, model_matrices, random_codes, random_tree


def _full_lists(tree: O.Tree):
    tree.mark_all_edges_updated()
    return tree.partials_lists()


def _calc_site(tree, model, codes, rates=(1.0,), scale_ln=30, threads=1):
    P = codes.shape[1]
    internals, children = _full_lists(tree)
    mats = model_matrices(model, tree, children, rates)
    out = []
    for c in range(len(rates)):
        calc = O.Calc4(codes, P, scale_ln, threads)
        calc.propose(internals, children, mats[c], model.frequencies())
        out.append(calc.likelihoods())
    return np.array(out)


@pytest.mark.parametrize("L", [2, 3, 4, 5, 6])
@pytest.mark.parametrize("seed", [0, 1])
def test_calc4_matches_brute_force(L, seed):
    rng = np.random.default_rng(seed * 10 + L)
    parents, children, heights = random_tree(L, rng, mean_branch=0.3)
    tree = O.Tree(L, parents, children, heights)
    f = rng.dirichlet([5] * 4)
    model = O.Model.gtr(rng.dirichlet([1] * 6), f)
    codes = random_codes(L, 24, rng, unknown_frac=0.2, ambiguous=True)
    site = _calc_site(tree, model, codes)[0]
    for i in range(codes.shape[1]):
        bf = brute_force_site_ll(parents, heights, codes[:, i], model.transition, f)
        assert abs(site[i] - bf) <= 1e-12 * max(1.0, abs(bf))


def test_two_taxon_closed_form_and_pulley():
    """L = sum_ij pi_i P_ij(t1 + t2) for a reversible Q; depends only on t1+t2."""
    rng = np.random.default_rng(3)
    f = rng.dirichlet([3] * 4)
    model = O.Model.gtr(rng.dirichlet([1] * 6), f)
    codes = np.array([[1, 1, 2, 4, 8, 15], [1, 2, 8, 4, 2, 1]], np.uint8)
    idx = {1: 0, 2: 1, 4: 2, 8: 3}
    for t1, t2 in [(0.1, 0.2), (0.25, 0.05), (0.01, 0.29)]:
        parents = np.array([2, 2, O.TOMBSTONE], np.uint32)
        heights = np.array([0.3 - t1, 0.3 - t2, 0.3])
        tree = O.Tree(2, parents, np.array([[0, 1]], np.uint32), heights)
        site = _calc_site(tree, model, codes)[0]
        P = model.transition(t1 + t2)
        for i in range(codes.shape[1]):
            a, b = codes[0, i], codes[1, i]
            if a == 15 and b == 15:
                expect = 0.0
            elif a == 15 or b == 15:  # sum_i pi_i P_ij = pi_j (stationarity)
                expect = math.log(f[idx[b if a == 15 else a]])
            else:
                expect = math.log(f[idx[a]] * P[idx[a], idx[b]])
            assert abs(site[i] - expect) < 1e-13


def test_all_unknown_column_is_zero():
    rng = np.random.default_rng(4)
    L = 20
    parents, children, heights = random_tree(L, rng)
    tree = O.Tree(L, parents, children, heights)
    model = O.Model.hky(2.0, [0.25] * 4)
    codes = np.full((L, 3), 15, np.uint8)
    codes[:, 1] = 0  # gap code, also unknown
    site = _calc_site(tree, model, codes)[0]
    np.testing.assert_allclose(site, 0.0, atol=1e-13)


@pytest.mark.parametrize("seed", range(3))
def test_scaling_invariance(seed):
    """b3's own check (python/tests/b3/likelihoods_comp.py:10,26-33): results
    do not depend on scale_ln in {3, 30, 300}."""
    rng = np.random.default_rng(seed)
    L = 120
    parents, children, heights = random_tree(L, rng, mean_branch=0.5)
    tree = O.Tree(L, parents, children, heights)
    model = O.Model.hky(3.0, rng.dirichlet([10] * 4))
    codes = random_codes(L, 300, rng, unknown_frac=0.05)
    a = _calc_site(tree, model, codes, scale_ln=3)[0]
    b = _calc_site(tree, model, codes, scale_ln=30)[0]
    c = _calc_site(tree, model, codes, scale_ln=300)[0]
    assert np.min(b) < -150  # deep enough that e^-30 and e^-3 scaling both fire
    np.testing.assert_allclose(a, b, rtol=1e-12)
    np.testing.assert_allclose(c, b, rtol=1e-12)


def test_threads_do_not_change_results():
    rng = np.random.default_rng(9)
    L, P = 30, 5000
    parents, children, heights = random_tree(L, rng)
    tree = O.Tree(L, parents, children, heights)
    model = O.Model.hky(2.0, rng.dirichlet([10] * 4))
    codes = random_codes(L, P, rng)
    a = _calc_site(tree, model, codes, threads=1)[0]
    b = _calc_site(tree, model, codes, threads=7)[0]
    np.testing.assert_array_equal(a, b)
    calc = O.Calc4(codes, P, 30, 0)
    assert 1 <= calc.num_threads <= 5  # min(cores, ceil(P/1000)), cpu4.rs:127-132


def _random_move(tree: O.Tree, rng):
    """A height move on a random internal node inside its valid range."""
    p, c, h = tree.get()
    L = tree.L
    node = int(rng.integers(L, 2 * L - 1))
    kids = c[node - L]
    lo = max(h[kids[0]], h[kids[1]])
    hi = h[p[node]] if p[node] != O.TOMBSTONE else lo + 1.0
    tree.set_height(node, lo + (hi - lo) * rng.uniform(0.05, 0.95))


@pytest.mark.parametrize("seed", range(4))
def test_partial_update_equals_full_and_reject_restores(seed):
    """Partial re-evaluation through the EditBuf slots equals a full
    recompute; reject restores the previous values bit-exactly."""
    rng = np.random.default_rng(seed)
    L, P = 40, 500
    parents, children, heights = random_tree(L, rng, mean_branch=0.4)
    tree = O.Tree(L, parents, children, heights)
    model = O.Model.hky(2.5, rng.dirichlet([10] * 4))
    codes = random_codes(L, P, rng, unknown_frac=0.05)
    f = model.frequencies()
    calc = O.Calc4(codes, P, 30, 1)
    internals, ch = _full_lists(tree)
    calc.propose(internals, ch, model_matrices(model, tree, ch, [1.0])[0], f)
    calc.accept()
    tree.accept()
    accepted = calc.likelihoods()
    for step in range(30):
        _random_move(tree, rng)
        internals, ch = tree.partials_lists()
        assert internals.size < L - 1 or step == 0 or True
        calc.propose(internals, ch, model_matrices(model, tree, ch, [1.0])[0], f)
        got = calc.likelihoods()
        # full recompute on a fresh calculator
        p2, c2, h2 = tree.get()
        fresh_tree = O.Tree(L, p2, c2, h2)
        full = _calc_site(fresh_tree, model, codes)[0]
        np.testing.assert_allclose(got, full, rtol=1e-13, atol=1e-13)
        if rng.random() < 0.5:
            calc.accept()
            tree.accept()
            accepted = got
        else:
            calc.reject()
            tree.reject()
            # re-propose the identical state: must reproduce the accepted values
            tree.mark_all_edges_updated()
            internals, ch = tree.partials_lists()
            calc.propose(internals, ch, model_matrices(model, tree, ch, [1.0])[0], f)
            np.testing.assert_array_equal(calc.likelihoods(), accepted)
            calc.accept()
            tree.accept()


def test_calcn_equals_calc4_at_four_states():
    rng = np.random.default_rng(12)
    L, P = 25, 400
    parents, children, heights = random_tree(L, rng)
    tree = O.Tree(L, parents, children, heights)
    model = O.Model.gtr(rng.dirichlet([1] * 6), rng.dirichlet([10] * 4))
    codes = random_codes(L, P, rng, unknown_frac=0.1)
    idx = np.full(256, 4, np.uint8)
    idx[[1, 2, 4, 8]] = [0, 1, 2, 3]
    states = idx[codes]
    internals, ch = _full_lists(tree)
    mats = model_matrices(model, tree, ch, [1.0])[0]
    c4 = O.Calc4(codes, P, 30, 1)
    c4.propose(internals, ch, mats, model.frequencies())
    cn = O.CalcN(states, P, 4, 30, 1)
    cn.propose(internals, ch, mats, model.frequencies())
    np.testing.assert_allclose(cn.likelihoods(), c4.likelihoods(), rtol=1e-14)


def test_calcn_61_states_brute_force():
    from aspartik_amd.b3 import GY94
    rng = np.random.default_rng(5)
    gy = GY94(0.3, 2.0, np.stack([rng.dirichlet([10] * 4) for _ in range(3)]))
    pi = gy.frequencies()
    model = O.Model.reversible(gy.exchangeability(), pi)
    L, P = 3, 12
    parents, children, heights = random_tree(L, rng, mean_branch=0.3)
    tree = O.Tree(L, parents, children, heights)
    states = rng.integers(0, 61, size=(L, P)).astype(np.uint8)
    states[0, 3] = 61  # unknown
    internals, ch = _full_lists(tree)
    mats = model_matrices(model, tree, ch, [1.0])[0]
    cn = O.CalcN(states, P, 61, 30, 1)
    cn.propose(internals, ch, mats, pi)
    got = cn.likelihoods()
    for i in range(P):
        bf = brute_force_site_ll(parents, heights, states[:, i], model.transition, pi, S=61)
        assert abs(got[i] - bf) < 1e-11 * max(1, abs(bf))


def test_weighted_sum_and_lse():
    rng = np.random.default_rng(1)
    v = rng.normal(-50, 10, size=1000)
    w = rng.integers(1, 6, size=1000).astype(np.uint32)
    assert abs(O.weighted_sum(v, w) - math.fsum(v * w)) < 1e-9
    x = np.array([-1000.0, -1001.0, -999.5])
    m = x.max()
    assert abs(O.log_sum_exp(x) - (m + math.log(math.fsum(np.exp(x - m))))) < 1e-12
    assert O.log_sum_exp(np.array([-np.inf, -np.inf])) == -np.inf


@pytest.mark.parametrize("K", [1, 4])
def test_gamma_likelihood_mixing_brute_force(K):
    """gamma.rs:154-166 against a direct mixture over categories."""
    rng = np.random.default_rng(21 + K)
    L, P = 5, 30
    parents, children, heights = random_tree(L, rng, mean_branch=0.3)
    tree = O.Tree(L, parents, children, heights)
    f = rng.dirichlet([5] * 4)
    model = O.Model.gtr(rng.dirichlet([1] * 6), f)
    codes = random_codes(L, P, rng)
    w = rng.integers(1, 6, size=P).astype(np.uint32)
    g = O.GammaLikelihood(tree, model, codes, P, w, num_categories=K, alpha=0.5)
    total = g.likelihood()  # cached after construction
    rates = g.rates()
    expect = 0.0
    for i in range(P):
        lk = [math.exp(brute_force_site_ll(parents, heights, codes[:, i],
                                           lambda t, r=r: model.transition(t * r), f))
              for r in rates]
        expect += w[i] * math.log(sum(lk) / K)
    assert abs(total - expect) < 1e-10 * abs(expect)
