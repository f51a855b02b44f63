"""CPU pins for the heterotachy / relaxed-clock oracle and host logic.

* oracle.components_eval (per-category models, per-edge clock rates, class
  select or gamma mix) against brute-force enumeration of internal states
  with P(t) = expm(Q t) from scipy, on 4-taxon trees: a dropped clock rate,
  a wrong category's model, frequencies or class fail it.
* Reductions: identical categories make every class assignment equal the
  one-category value; all edge rates r equal a strict clock r; classes all
  c equal category c alone.
* Relaxed-clock rates: quantile midpoints (i + 1/2)/n of LogNormal,
  Exponential and Gamma against scipy.stats and closed forms.
* ClassVector (EpochBuf semantics) and the clocks' dirty-edge marking.
"""
import math

import numpy as np
import pytest
from scipy import stats
from scipy.linalg import expm

from aspartik_amd import distributions as D
from aspartik_amd.b3 import ClassVector, Clock, DatedTree, Real
from oracle import oracle as O
from tests.helpers import brute_force_site_ll, random_codes, random_tree


def _gtr(rng):
    rates = rng.dirichlet([2] * 6)
    freqs = rng.dirichlet([5] * 4)
    return O.Model.gtr(rates, freqs), rates, freqs


def _q(model):
    return model.q


def _case(L=4, P=12, K=3, seed=0):
    rng = np.random.default_rng(seed)
    par, ch, h = random_tree(L, rng, mean_branch=0.3)
    codes = random_codes(L, P, rng, unknown_frac=0.1)
    models = [_gtr(rng)[0] for _ in range(K)]
    er = rng.uniform(0.2, 3.0, size=(K, 2 * L - 1))
    w = rng.integers(1, 5, size=P).astype(np.uint32)
    return rng, par, ch, h, codes, models, er, w


def _brute(par, h, codes, models, er):
    K, P = len(models), codes.shape[1]
    out = np.zeros((K, P))
    for c, m in enumerate(models):
        Q = _q(m)
        tr = lambda t, Q=Q: expm(Q * t)  # noqa: E731
        for i in range(P):
            out[c, i] = brute_force_site_ll(par, h, codes[:, i], tr, m.frequencies(),
                                            edge_rates=er[c])
    return out


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_hetero_total_matches_brute_force(seed):
    rng, par, ch, h, codes, models, er, w = _case(seed=seed)
    t = O.Tree(4, par, ch, h)
    cls = rng.integers(0, 3, size=codes.shape[1]).astype(np.uint32)
    tot, site = O.components_eval(t, models, er, codes, w, classes=cls)
    bf = _brute(par, h, codes, models, er)
    np.testing.assert_allclose(site, bf, rtol=1e-10, atol=1e-10)
    ref = float(sum(bf[cls[i], i] * w[i] for i in range(codes.shape[1])))
    assert abs(tot - ref) <= 1e-10 * abs(ref)


def test_gamma_mix_with_relaxed_clock_matches_brute_force():
    rng, par, ch, h, codes, models, er, w = _case(seed=7, K=1)
    t = O.Tree(4, par, ch, h)
    r = O.gamma_categories(0.7, 4)
    edge = er[0]
    rates = edge[None, :] * r[:, None]      # clock(e) * r_c
    tot, _ = O.components_eval(t, [models[0]] * 4, rates, codes, w)
    bf = _brute(par, h, codes, [models[0]] * 4, rates)
    mixed = np.log(np.exp(bf).mean(0))
    ref = float((mixed * w).sum())
    assert abs(tot - ref) <= 1e-10 * abs(ref)


def test_reductions():
    rng, par, ch, h, codes, models, er, w = _case(L=9, P=80, seed=3)
    t = O.Tree(9, par, ch, h)
    # identical categories: every class assignment gives the one-category value
    same = [models[0]] * 3
    er_same = np.repeat(er[:1], 3, axis=0)
    one, _ = O.components_eval(t, [models[0]], er[:1], codes, w)
    cls = rng.integers(0, 3, size=80).astype(np.uint32)
    tot, _ = O.components_eval(t, same, er_same, codes, w, classes=cls)
    assert abs(tot - one) <= 1e-12 * abs(one)
    # uniform edge rates r == strict clock r (GammaLikelihood with clock_rate r)
    g = O.GammaLikelihood(t, models[1], codes, 80, w, num_categories=1, clock_rate=1.7)
    strict, _ = O.components_eval(t, [models[1]], np.full((1, 17), 1.7), codes, w)
    assert abs(strict - g.likelihood()) <= 1e-12 * abs(strict)
    # all patterns in class c == category c alone
    for c in range(3):
        tot, _ = O.components_eval(t, models, er, codes, w, classes=np.full(80, c, np.uint32))
        alone, _ = O.components_eval(t, [models[c]], er[c:c + 1], codes, w)
        assert tot == alone


def test_lognormal_quantiles():
    d = D.LogNormal(0.3, 0.5)
    assert abs(d.inverse_cdf(0.5) - math.exp(0.3)) < 1e-15
    assert abs(d.inverse_cdf(stats.norm.cdf(1.0)) - math.exp(0.8)) < 1e-12
    for p in [1e-6, 0.01, 0.3, 0.77, 0.999]:
        assert abs(d.inverse_cdf(p) - stats.lognorm.ppf(p, 0.5, scale=math.exp(0.3))) \
            <= 1e-12 * d.inverse_cdf(p)
    e = D.Exponential(2.0)
    assert abs(e.inverse_cdf(0.5) - math.log(2) / 2) < 1e-15
    g = D.Gamma(2.5, 0.4)
    assert abs(g.inverse_cdf(0.3) - stats.gamma.ppf(0.3, 2.5, scale=0.4)) < 1e-12
    u = D.Uniform(1.0, 3.0)
    assert u.inverse_cdf(0.25) == 1.5


def test_relaxed_clock_rates_are_quantile_midpoints():
    loc, scale = Real(-0.2), Real(0.4)
    cats = ClassVector(5, 7)
    clock = Clock.Relaxed(cats, D.LogNormal(loc, scale))
    ref = stats.lognorm.ppf((np.arange(5) + 0.5) / 5, 0.4, scale=math.exp(-0.2))
    np.testing.assert_allclose(clock.rates, ref, rtol=1e-12)
    for i in range(7):
        cats.set(i, i % 5)
    np.testing.assert_array_equal(clock.edge_rates(7), clock.rates[np.arange(7) % 5])
    assert clock.get_rate(6) == clock.rates[1]


def test_class_vector_epochbuf_semantics():
    v = ClassVector(3, 5)
    assert not v.is_changed()
    v.set(2, 1)
    v.set(2, 1)          # unchanged value still records the index
    v.set(4, 2)
    assert v.changed_indices() == [2, 2, 4] and v.is_changed()
    v.reject()
    assert list(v.values) == [0] * 5 and not v.is_changed()
    v.set(1, 2)
    v.accept()
    assert v[1] == 2 and not v.is_changed()
    with pytest.raises(ValueError):
        v.set(0, 3)
    with pytest.raises(ValueError):
        ClassVector(1, 4)


def _tree(seed=0, L=8):
    rng = np.random.default_rng(seed)
    par, ch, h = random_tree(L, rng)
    return DatedTree(L, par, ch, h)


def test_clock_marking():
    tree = _tree()
    tree.accept()
    rate = Real(1.0)
    strict = Clock.Strict(rate)
    strict.update(tree)
    assert not tree.is_changed()
    rate.set(2.0)
    strict.update(tree)
    assert tree.is_changed() and strict.get_rate() == 2.0
    rate.reject()
    strict.reject()
    assert strict.get_rate() == 1.0
    tree.accept()

    cats = ClassVector(4, tree.num_nodes)
    scale = Real(0.5)
    relaxed = Clock.Relaxed(cats, D.LogNormal(0.0, scale))
    relaxed.update(tree)
    assert not tree.is_changed()
    leaf = 3
    cats.set(leaf, 2)
    relaxed.update(tree)
    internals, _ = tree.partials_lists()
    # the leaf's edge and every ancestor are dirty
    anc, n = [], leaf
    while tree.parent_of(n) is not None:
        n = tree.parent_of(n)
        anc.append(n)
    assert tree.edges_to_update() == [leaf]
    assert set(int(i) for i in internals) == set(anc)
    cats.accept()
    tree.accept()
    before = relaxed.rates.copy()
    scale.set(0.9)
    relaxed.update(tree)
    assert len(tree.edges_to_update()) == tree.num_nodes - 1
    assert not np.array_equal(before, relaxed.rates)
    scale.reject()
    relaxed.reject()          # rates recomputed from the restored scale
    np.testing.assert_array_equal(before, relaxed.rates)
