"""HeteroLikelihood and relaxed clocks on the GPU against the FP64 oracle.

Every likelihood() of a run of moves (node heights, per-pattern classes,
relaxed-clock rate categories and distribution parameters, a substitution
model), each accepted or rejected at random, is compared with the oracle's
from-scratch evaluation of the same state (oracle.components_eval: one
calculator per category with its own model and per-edge clock rates,
hetero.rs:122-160).  Tolerances: FP32 path 1e-6 relative on the total (the
stated target); FP64 plain path 1e-12.  The C-ABI pieces (per-category
models, branch lengths, classes, result slots after reject) are checked
directly, with the error paths.
"""
import ctypes as C

import numpy as np
import pytest

from aspartik_amd import _lib
from aspartik_amd import distributions as D
from aspartik_amd.b3 import (GTR, Calculator, ClassVector, Clock, DatedTree, GammaLikelihood,
                             HeteroLikelihood, Real, gamma_rates)
from oracle import oracle as O
from tests.helpers import random_codes, random_tree

pytestmark = pytest.mark.gpu

TOL = {"f32": 1e-6, "f64": 1e-12}


def _calc(precision):
    return Calculator.HIP() if precision == "f32" else Calculator.HIPPlainF64()


def _oracle_total(tree, subs, clocks, codes, w, classes=None, cat_rates=None):
    t = O.Tree(tree.L, tree.parents_u32(), tree.children.astype(np.uint32), tree.heights)
    models = [O.Model.gtr(s.rates, s.freqs) for s in subs]
    er = np.stack([c.edge_rates(tree.num_nodes) for c in clocks])
    if cat_rates is not None:  # gamma: clock(e) * r_c
        er = er[0][None, :] * np.asarray(cat_rates)[:, None]
    tot, site = O.components_eval(t, models, er, codes, w, classes=classes)
    return tot, site


def _state(L=24, P=900, K=3, seed=0):
    rng = np.random.default_rng(seed)
    par, ch, h = random_tree(L, rng, mean_branch=0.08)
    codes = random_codes(L, P, rng, unknown_frac=0.05)
    w = rng.integers(1, 4, size=P).astype(np.uint32)
    tree = DatedTree(L, par, ch, h)
    subs = [GTR(rng.dirichlet([8] * 4), rng.dirichlet([3] * 6)) for _ in range(K)]
    cats = ClassVector(K, P, rng.integers(0, K, size=P))
    clocks, params = [Clock.Strict(Real(1.0))], []
    for _ in range(K - 1):
        rc = ClassVector(4, tree.num_nodes, rng.integers(0, 4, size=tree.num_nodes))
        loc, scale = Real(float(rng.normal(0, 0.2))), Real(float(rng.uniform(0.2, 0.6)))
        clocks.append(Clock.Relaxed(rc, D.LogNormal(loc, scale)))
        params.append((rc, loc, scale))
    return rng, tree, codes, w, subs, cats, clocks, params


def _height_move(rng, tree):
    cand = [n for n in range(tree.L, tree.N) if n != tree.root]
    n = int(rng.choice(cand))
    lo = max(tree.heights[c] for c in tree.children_of(n))
    hi = tree.heights[tree.parent_of(n)]
    tree.set_height(n, float(rng.uniform(lo, hi)))


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_hetero_likelihood_through_moves(hip_device, precision):
    rng, tree, codes, w, subs, cats, clocks, params = _state(seed=1 if precision == "f32" else 2)
    lik = HeteroLikelihood(codes, w, tree, cats, subs, clocks, calculator=_calc(precision))
    ref, _ = _oracle_total(tree, subs, clocks, codes, w, cats.values)
    assert abs(lik.likelihood() - ref) <= TOL[precision] * abs(ref)
    kinds = ["height", "classes", "clock_cats", "clock_dist", "model", "height+classes"]
    for step in range(48):
        kind = kinds[step % len(kinds)]
        old_model = None
        if "height" in kind:
            _height_move(rng, tree)
        if "classes" in kind:
            for i in rng.integers(0, len(cats), size=5):
                cats.set(int(i), int(rng.integers(0, 3)))
        if kind == "clock_cats":
            rc = params[step % 2][0]
            for e in rng.integers(0, tree.num_nodes, size=3):
                rc.set(int(e), int(rng.integers(0, 4)))
        if kind == "clock_dist":
            params[step % 2][2].set(float(rng.uniform(0.2, 0.6)))
        if kind == "model":
            s = subs[step % 3]
            old_model = (s, s.rates.copy())
            s.set_rates(rng.dirichlet([3] * 6))
        got = lik.likelihood()
        ref, _ = _oracle_total(tree, subs, clocks, codes, w, cats.values)
        assert abs(got - ref) <= TOL[precision] * abs(ref), (step, kind, got, ref)
        accept = bool(rng.random() < 0.5)
        for obj in [tree, cats] + [p for trio in params for p in trio]:
            obj.accept() if accept else obj.reject()
        if not accept and old_model is not None:
            old_model[0].set_rates(old_model[1])
        lik.accept() if accept else lik.reject()
        # the next call sees no change (or only the restored model)
        again = lik.likelihood()
        ref, _ = _oracle_total(tree, subs, clocks, codes, w, cats.values)
        assert abs(again - ref) <= TOL[precision] * abs(ref), (step, kind, "after", accept)
        lik.accept()
        tree.accept()


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_gamma_likelihood_with_relaxed_clock(hip_device, precision):
    rng, tree, codes, w, subs, cats, clocks, params = _state(L=30, P=1500, K=2, seed=5)
    clock = clocks[1]
    rc, loc, scale = params[0]
    lik = GammaLikelihood(codes, w, tree, subs[0], num_categories=4, alpha=0.6,
                          calculator=_calc(precision), clock=clock)
    r = gamma_rates(0.6, 4)
    for step in range(12):
        if step % 3 == 0:
            _height_move(rng, tree)
        elif step % 3 == 1:
            for e in rng.integers(0, tree.num_nodes, size=4):
                rc.set(int(e), int(rng.integers(0, 4)))
        else:
            loc.set(float(rng.normal(0, 0.3)))
        got = lik.likelihood()
        ref, _ = _oracle_total(tree, [subs[0]] * 4, [clock], codes, w, cat_rates=r)
        assert abs(got - ref) <= TOL[precision] * abs(ref), (step, got, ref)
        accept = step % 2 == 0
        for obj in (tree, rc, loc, scale):
            obj.accept() if accept else obj.reject()
        lik.accept() if accept else lik.reject()


def _raw(precision, L=16, P=500, K=3, seed=9):
    rng = np.random.default_rng(seed)
    par, ch, h = random_tree(L, rng, mean_branch=0.1)
    codes = random_codes(L, P, rng)
    calc = _calc(precision).make(codes, P, K)
    t = O.Tree(L, par, ch, h)
    t.mark_all_edges_updated()
    internals, children = t.partials_lists()
    subs = [GTR(rng.dirichlet([8] * 4), rng.dirichlet([3] * 6)) for _ in range(K)]
    for c, s in enumerate(subs):
        calc.set_category_model(c, s.q_matrix(), s.frequencies())
    return rng, par, ch, h, codes, calc, t, internals, children, subs


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_classes_after_reject_use_accepted_sites(hip_device, precision):
    """A rejected propose leaves the accepted site values current: changing
    only the classes afterwards totals the accepted state (hetero.rs:117-120,
    149-160 with SkSliceBuf::reject)."""
    rng, par, ch, h, codes, calc, t, internals, children, subs = _raw(precision)
    K, P, N = 3, codes.shape[1], 2 * 16 - 1
    er = rng.uniform(0.5, 2.0, size=(K, N))
    w = np.ones(P, np.uint32)
    models = [O.Model.gtr(s.rates, s.freqs) for s in subs]

    def lengths(tt, ch_list):
        flat = ch_list.reshape(-1).astype(np.int64)
        tl = np.array([tt.edge_length(int(e)) for e in flat])
        return np.stack([tl * er[c][flat] for c in range(K)])

    cls1 = rng.integers(0, K, size=P).astype(np.uint32)
    calc.set_classes(cls1)
    calc.propose_lengths(internals, children, lengths(t, children))
    ref1, site1 = O.components_eval(t, models, er, codes, w, classes=cls1)
    assert abs(calc.total() - ref1) <= TOL[precision] * abs(ref1)
    for c in range(K):
        np.testing.assert_allclose(calc.likelihoods(c), site1[c], rtol=1e-5 if precision == "f32"
                                   else 1e-12)
    calc.accept()
    # a proposal on a different tree, rejected
    h2 = h.copy()
    par64 = par.astype(np.int64)
    n = next(i for i in range(16, N) if par[i] != 0xFFFFFFFF)
    lo = max(h[c] for c in ch[n - 16])
    h2[n] = 0.5 * (lo + h[par64[n]])
    t2 = O.Tree(16, par, ch, h2)
    t2.mark_all_edges_updated()
    i2, c2 = t2.partials_lists()
    calc.propose_lengths(i2, c2, lengths(t2, c2))
    ref2, _ = O.components_eval(t2, models, er, codes, w, classes=cls1)
    assert abs(calc.total() - ref2) <= TOL[precision] * abs(ref2)
    calc.reject()
    assert abs(calc.total() - ref1) <= TOL[precision] * abs(ref1)
    cls2 = rng.integers(0, K, size=P).astype(np.uint32)
    calc.set_classes(cls2)
    ref3, _ = O.components_eval(t, models, er, codes, w, classes=cls2)
    assert abs(calc.total() - ref3) <= TOL[precision] * abs(ref3)
    sel = calc.site_mixed()
    np.testing.assert_allclose(sel, site1[cls2, np.arange(P)],
                               rtol=1e-5 if precision == "f32" else 1e-12)
    # back to the category mean
    calc.set_classes(None)
    mean = np.log(np.exp(site1).mean(0))
    np.testing.assert_allclose(calc.site_mixed(), mean, rtol=1e-5 if precision == "f32" else 1e-12)


def test_hetero_error_paths(hip_device):
    rng, par, ch, h, codes, calc, t, internals, children, subs = _raw("f32")
    P = codes.shape[1]
    with pytest.raises(RuntimeError):
        calc.set_classes(np.full(P, 3, np.uint32))       # class >= K
    with pytest.raises(ValueError):
        calc.set_classes(np.zeros(P - 1, np.uint32))
    with pytest.raises(RuntimeError):
        calc.set_category_model(3, subs[0].q_matrix(), subs[0].frequencies())
    with pytest.raises(ValueError):
        calc.propose_lengths(internals, children, np.ones(5))
    fresh = Calculator.HIP().make(codes, P, 2)
    with pytest.raises(RuntimeError):  # no model yet
        fresh.propose_lengths(internals, children, np.ones(2 * 2 * internals.size))
    with pytest.raises(RuntimeError):  # no result yet
        fresh.total()
    codon = Calculator.HIPCodon(states=61).make(np.zeros((16, 40), np.uint8), 40, 1)
    q = np.full((61, 61), 1.0 / 60)
    np.fill_diagonal(q, -1.0)
    with pytest.raises(RuntimeError):
        codon.set_category_model(0, q[:4, :4], np.full(4, 0.25))
    lib = _lib.lib()
    assert lib.b3h_set_classes(None, None) != 0
    z1, z2, zd = np.zeros(1, np.uint32), np.zeros(2, np.uint32), np.zeros(2)
    assert lib.b3h_propose_lengths(None, 0, _lib.addr(z1), _lib.addr(z2), _lib.addr(zd)) != 0
    assert C.c_int(lib.b3h_set_category_model(None, 0, np.zeros(16), np.ones(4), 0)).value != 0
