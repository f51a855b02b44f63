"""Propagation storage on the GPU (b3's default GPU calculator,
calculator/cuda_propagations) against the oracle.

* FP64 propagation handle with host matrices against the oracle's Prop4
  restatement under height moves and nephew swaps with random accept /
  reject, scale_ln 3 and 30: per-site values agree to 1e-14 (the operation
  order and the scaling flags are b3's; only the device and host logarithms
  may differ in the last bit), so a wrong flag (a 3- or 30-nat jump) fails.
* FP32 and FP64 propagation handles with P(t) built on the device, K = 4,
  against a full Cpu4 recompute: FP32 per-site |d lnL| <= 2e-6 max(1, |lnL|)
  and total 1e-6; FP64 1e-11.
* GammaLikelihood over a propagation handle through its own lists.
* Protocol errors between the two list types.
"""
import numpy as np
import pytest

from aspartik_amd.b3 import HKY, Calculator, DatedTree, GammaLikelihood, gamma_rates
from oracle import oracle as O
from tests.helpers import model_matrices, random_codes, random_tree

pytestmark = pytest.mark.gpu

F32_SITE = 2e-6
F32_TOTAL = 1e-6


def _setup(L, P, seed, mean_branch=0.2):
    rng = np.random.default_rng(seed)
    par, ch, h = random_tree(L, rng, mean_branch)
    f = rng.dirichlet([10] * 4)
    kappa = float(rng.lognormal(1.0, 0.5))
    codes = random_codes(L, P, rng, unknown_frac=0.05, ambiguous=True)
    w = rng.integers(1, 6, size=P).astype(np.uint32)
    return rng, (par, ch, h), O.Model.hky(kappa, f), HKY(f, kappa), codes, w


def _move(tree, otree, rng):
    """A height move or a nephew swap, applied to the mirror and the oracle."""
    L = tree.L
    if rng.random() < 0.6:
        cand = [n for n in range(L, 2 * L - 1) if n != tree.root]
        n = int(rng.choice(cand))
        lo = max(tree.heights[c] for c in tree.children_of(n))
        hi = tree.heights[tree.parent_of(n)]
        nh = float(lo + (hi - lo) * rng.uniform(0.05, 0.95))
        tree.set_height(n, nh)
        otree.set_height(n, nh)
        return
    for _ in range(100):
        node = int(rng.integers(0, 2 * L - 1))
        par = tree.parent_of(node)
        if par is None or tree.parent_of(par) is None:
            continue
        gp = tree.parent_of(par)
        a, b = tree.children_of(gp)
        uncle = a if b == par else b
        if tree.heights[uncle] >= tree.heights[par]:
            continue
        for t in (tree, otree):
            t.replace_child(par, node, uncle)
            t.replace_child(gp, uncle, node)
        return


def _node_mats(om, tree, nodes, rates):
    nm = len(nodes) - 1
    out = np.zeros((len(rates), nm, 4, 4))
    for c, r in enumerate(rates):
        for m in range(nm):
            out[c, m] = om.transition(tree.edge_length(int(nodes[m])) * r)
    return out


def _cpu4_sites(otree, om, codes, rates):
    p, c, h = otree.get()
    t = O.Tree(otree.L, p, c, h)
    t.mark_all_edges_updated()
    internals, children = t.partials_lists()
    mats = model_matrices(om, t, children, rates)
    out = []
    for k in range(len(rates)):
        calc = O.Calc4(codes, codes.shape[1], 30, 0)
        calc.propose(internals, children, mats[k], om.frequencies())
        out.append(calc.likelihoods())
    return np.array(out)


def _mix(sites, w):
    K = sites.shape[0]
    mx = sites.max(0)
    mixed = mx + np.log(np.exp(sites - mx).sum(0)) + np.log(1.0 / K)
    return float(np.sum(mixed * w))


@pytest.mark.parametrize("scale_ln", [3, 30])
def test_f64_propagation_matches_prop4_oracle(hip_device, scale_ln):
    L, P = 45, 1500
    rng, arrays, om, pm, codes, w = _setup(L, P, seed=scale_ln, mean_branch=0.4)
    rates = [0.6, 1.7]
    calc = Calculator.HIPPropagationsF64().with_scale(scale_ln).make(codes, P, 2)
    oracles = [O.Prop4(codes, P, scale_ln) for _ in rates]
    tree = DatedTree(L, *arrays)
    otree = O.Tree(L, *arrays)
    tree.mark_all_edges_updated()
    otree.mark_all_edges_updated()
    f = om.frequencies()
    accepted = None
    for step in range(36):
        if step:
            _move(tree, otree, rng)
        nodes, children, nl = tree.propagation_lists()
        onodes, ochildren, onl = otree.propagation_lists()
        np.testing.assert_array_equal(nodes, onodes)
        mats = _node_mats(om, otree, onodes, rates)
        calc.propose_propagations(nodes, children, nl, mats, f)
        for k, oc in enumerate(oracles):
            oc.propose(onodes, ochildren, onl, mats[k], f)
            np.testing.assert_allclose(calc.likelihoods(k), oc.likelihoods(), rtol=1e-14,
                                       atol=1e-14, err_msg=f"step {step} category {k}")
        if step == 0 or rng.random() < 0.5:
            calc.accept()
            [oc.accept() for oc in oracles]
            tree.accept()
            otree.accept()
            accepted = [calc.likelihoods(k) for k in range(2)]
        else:
            calc.reject()
            [oc.reject() for oc in oracles]
            tree.reject()
            otree.reject()
            # the accepted results are current again
            for k in range(2):
                np.testing.assert_array_equal(calc.likelihoods(k), accepted[k])


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_propagation_device_matrices_against_cpu4(hip_device, precision):
    L, P, K = 60, 4000, 4
    rng, arrays, om, pm, codes, w = _setup(L, P, seed=7)
    rates = gamma_rates(0.5, K)
    cfg = Calculator.HIPPropagations() if precision == "f32" else Calculator.HIPPropagationsF64()
    calc = cfg.make(codes, P, K)
    calc.set_weights(w)
    calc.set_model(pm.q_matrix(), pm.frequencies(), True)
    tree = DatedTree(L, *arrays)
    otree = O.Tree(L, *arrays)
    tree.mark_all_edges_updated()
    for step in range(24):
        if step:
            _move(tree, otree, rng)
        nodes, children, nl = tree.propagation_lists()
        flat = nodes[:-1].astype(np.int64)
        times = tree.heights[tree.parents[flat]] - tree.heights[flat]
        calc.propose_propagations_times(nodes, children, nl, times, rates)
        got = np.array([calc.likelihoods(k) for k in range(K)])
        ref = _cpu4_sites(otree, om, codes, rates)
        if precision == "f32":
            assert np.all(np.abs(got - ref) <= F32_SITE * np.maximum(1.0, np.abs(ref))), step
            tol = F32_TOTAL
        else:
            np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-11)
            tol = 1e-12
        total_ref = _mix(ref, w)
        assert abs(calc.total() - total_ref) <= tol * abs(total_ref), step
        if step == 0 or rng.random() < 0.5:
            calc.accept()
            tree.accept()
            otree.accept()
        else:
            calc.reject()
            tree.reject()
            otree.reject()


def test_gamma_likelihood_over_propagations(hip_device):
    L, P = 40, 2500
    rng, arrays, om, pm, codes, w = _setup(L, P, seed=3)
    tree = DatedTree(L, *arrays)
    lik = GammaLikelihood(codes, w, tree, pm, num_categories=4, alpha=0.8,
                          calculator=Calculator.HIPPropagations())
    otree = O.Tree(L, *arrays)
    olik = O.GammaLikelihood(otree, om, codes, P, w, num_categories=4, alpha=0.8)
    for step in range(20):
        _move(tree, otree, rng)
        got, ref = lik.likelihood(), olik.likelihood()
        assert abs(got - ref) <= F32_TOTAL * abs(ref), step
        if rng.random() < 0.5:
            lik.accept(), olik.accept(), tree.accept(), otree.accept()
        else:
            lik.reject(), olik.reject(), tree.reject(), otree.reject()


def test_propagation_protocol_errors(hip_device):
    L, P = 12, 300
    rng, arrays, om, pm, codes, w = _setup(L, P, seed=1)
    prop = Calculator.HIPPropagations().make(codes, P, 1)
    part = Calculator.HIP().make(codes, P, 1)
    tree = DatedTree(L, *arrays)
    tree.mark_all_edges_updated()
    internals, children = tree.partials_lists()
    nodes, pch, nl = tree.propagation_lists()
    eye = np.tile(np.eye(4), (2 * internals.size, 1, 1))
    with pytest.raises(RuntimeError):
        prop.propose(internals, children, eye, pm.frequencies())
    with pytest.raises(RuntimeError):
        part.propose_propagations(nodes, pch, nl, np.tile(np.eye(4), (nodes.size - 1, 1, 1)),
                                  pm.frequencies())
    bad = nodes.copy()
    bad[0] = 2 * L  # out of range
    with pytest.raises(RuntimeError):
        prop.propose_propagations(bad, pch, nl, np.tile(np.eye(4), (nodes.size - 1, 1, 1)),
                                  pm.frequencies())
    with pytest.raises(RuntimeError):  # the last node must be internal
        prop.propose_propagations(nodes[:nl], pch[:0], nl, np.tile(np.eye(4), (nl, 1, 1)),
                                  pm.frequencies())
    with pytest.raises(RuntimeError):  # no device model yet
        prop.propose_propagations_times(nodes, pch, nl, np.ones(nodes.size - 1), [1.0])
    with pytest.raises(RuntimeError):
        prop.set_category_model(0, pm.q_matrix(), pm.frequencies())


def test_autotune_candidates_agree(hip_device):
    """Every candidate the autotuner times computes the same total (the
    choice changes the schedule and the storage, not the value)."""
    from aspartik_amd.b3 import autotune
    from aspartik_amd.b3.tuning import default_candidates
    L, P = 30, 3000
    rng, arrays, om, pm, codes, w = _setup(L, P, seed=21)
    rates = gamma_rates(0.5, 4)
    tree = DatedTree(L, *arrays)
    best, trials = autotune(codes, w, tree, pm, rates, evals=4)
    assert {t.name for t in trials} == {n for n, _ in default_candidates(4)}
    assert trials[0].config is best and all(t.full_ms > 0 and t.partial_ms > 0 for t in trials)
    totals = []
    for name, cfg in default_candidates(4):
        lik = GammaLikelihood(codes, w, tree.copy(), pm, num_categories=4, alpha=0.5,
                              calculator=cfg)
        totals.append(lik.likelihood())
    assert max(totals) - min(totals) <= 2e-6 * abs(totals[0])
