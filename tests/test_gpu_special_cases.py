"""Device P(t) build: JC and K80 (closed forms in b3, substitution.rs:131-148,
208-232) and the special cases of HKY/GTR (t > 20 -> every row pi,
substitution.rs:408-411,576-579; t < 1e-10 -> I + Q t with the diagonal set
to exactly 1, :259-268) against the FP64 oracle, whose transition function
restates those branches.

Trees are seeded Yule trees with edges forced into each branch: a zero-length
edge (a node at its parent's height), an edge of 1e-12, and the two root
edges lengthened so that t * r_c > 20 in every rate category.
"""
import numpy as np
import pytest

from aspartik_amd.b3 import GTR, HKY, JC, K80, Calculator, DatedTree, gamma_rates
from oracle import oracle as O
from tests.helpers import model_matrices, random_codes, random_tree

pytestmark = pytest.mark.gpu

F32_SITE = 2e-6
F32_TOTAL = 1e-6


def _edge_tree(L, seed, long_root=True, zero=True, tiny=True):
    rng = np.random.default_rng(seed)
    parents, children, heights = random_tree(L, rng, 0.1)
    heights = heights.copy()
    N = 2 * L - 1
    root = int(np.nonzero(parents == 0xFFFFFFFF)[0][0])
    internals = [n for n in range(L, N) if n != root]
    if zero:
        # an internal node moved up to its parent's height: a zero-length edge
        n = internals[0]
        heights[n] = heights[parents[n]]
    if tiny:
        # a leaf edge of 1e-12 (the leaf raised to just below its parent)
        leaf = int(np.argmin([heights[parents[x]] - heights[x] for x in range(L)]))
        heights[leaf] = heights[parents[leaf]] - 1e-12
    if long_root:
        # the root's two edges longer than 20 / (smallest category rate)
        heights[root] += 3000.0
    return rng, (parents, children, heights)


def _oracle(tree_arrays, om, codes, rates, weights):
    parents, ch, heights = tree_arrays
    L, P = codes.shape
    t = O.Tree(L, parents, ch, heights)
    t.mark_all_edges_updated()
    internals, children = t.partials_lists()
    mats = model_matrices(om, t, children, rates)
    sites = []
    for c in range(len(rates)):
        calc = O.Calc4(codes, P, 30, 0)
        calc.propose(internals, children, mats[c], om.frequencies())
        sites.append(calc.likelihoods())
    sites = np.array(sites)
    K = len(rates)
    mx = sites.max(0)
    mixed = mx + np.log(np.exp(sites - mx).sum(0)) + np.log(1.0 / K)
    return sites, float(O.weighted_sum(mixed, weights))


def _device(precision, tree_arrays, pm, codes, rates, weights):
    L, P = codes.shape
    cfg = Calculator.HIP() if precision == "f32" else Calculator.HIPPlainF64()
    calc = cfg.make(codes, P, len(rates))
    calc.set_weights(weights)
    calc.set_model(pm.q_matrix(), pm.frequencies(), pm.special_cases)
    t = DatedTree(L, *tree_arrays)
    t.mark_all_edges_updated()
    internals, children = t.partials_lists()
    flat = children.reshape(-1).astype(np.int64)
    times = t.heights[t.parents[flat]] - t.heights[flat]
    calc.propose_times(internals, children, times, rates)
    sites = np.array([calc.likelihoods(c) for c in range(len(rates))])
    total = calc.total()
    calc.accept()
    calc.close()
    return sites, total, times


def _check(precision, got, ref, total, total_ref):
    if precision == "f32":
        err = np.abs(got - ref) / np.maximum(1.0, np.abs(ref))
        assert err.max() <= F32_SITE, f"max per-site error {err.max():.3e}"
        assert abs(total - total_ref) <= F32_TOTAL * abs(total_ref)
    else:
        np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-11)
        assert abs(total - total_ref) <= 1e-12 * abs(total_ref)


def _models(name, rng):
    f = rng.dirichlet([10] * 4)
    if name == "jc":
        return JC(), O.Model.jc()
    if name == "k80":
        return K80(2.7), O.Model.k80(2.7)
    if name == "hky":
        return HKY(f, 3.1), O.Model.hky(3.1, f)
    r = rng.dirichlet([1] * 6)
    return GTR(f, r), O.Model.gtr(r, f)


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("name", ["jc", "k80", "hky", "gtr"])
@pytest.mark.parametrize("K", [1, 4])
def test_special_edges_match_oracle(hip_device, precision, name, K):
    L, P = 12, 1500
    rng, tree = _edge_tree(L, seed=31 + K)
    pm, om = _models(name, rng)
    codes = random_codes(L, P, rng, unknown_frac=0.05, ambiguous=True)
    w = rng.integers(1, 6, size=P).astype(np.uint32)
    rates = gamma_rates(0.5, K) if K > 1 else np.ones(1)
    ref, total_ref = _oracle(tree, om, codes, rates, w)
    got, total, times = _device(precision, tree, pm, codes, rates, w)
    # the tree really exercises every branch of the P(t) build
    assert (times == 0.0).any()
    assert ((times > 0) & (times < 1e-10)).any()
    assert (times * rates.min() > 20.0).any()
    if precision == "f64":
        # Two correct FP64 P(t) evaluations (b3's JC/K80 closed forms, an
        # eigen path such as the device's, the oracle's HKY analytic
        # eigensystem) form the off-diagonals of a short edge as differences
        # of O(1) terms, with ~4e-16 absolute error each: relative
        # ~4e-16 / (t r).  This tree has t r = 1.1e-6 (and a 1e-14 leaf edge
        # for JC/K80, which have no t < 1e-10 special case), so on the sites
        # that hinge on those entries any two paths differ by ~1e-10 (the
        # oracle's own closed form and eigen path differ by up to 2.8e-10
        # here, on the CPU).  Bound: 1e-9 per site, 1e-12 on the total.
        np.testing.assert_allclose(got, ref, rtol=1e-9, atol=1e-11)
        assert abs(total - total_ref) <= 1e-12 * abs(total_ref)
    else:
        _check(precision, got, ref, total, total_ref)


@pytest.mark.parametrize("t", [20.5, 6e-11])
def test_special_case_branches_are_taken_on_the_device(hip_device, t):
    """Two leaves, both edges of length t, FP64 device P(t).  At t = 20.5
    the special case makes every row pi (the eigen path differs by about
    e^-20.5); at t = 6e-11 it gives I + Q t with a unit diagonal (the eigen
    path differs in the diagonal by about |q_ii| t).  Both differences show
    in the site values, and the device must follow the special cases: the
    oracle's value with them, not the one without them (the same Q through
    the general reversible model, which has no special cases)."""
    rng = np.random.default_rng(12)
    f = rng.dirichlet([10] * 4)
    r = rng.dirichlet([1] * 6)
    om = O.Model.gtr(r, f)
    R = np.zeros((4, 4))
    iu = np.triu_indices(4, 1)
    R[iu] = r
    R = R + R.T
    plain = O.Model.reversible(R, f)
    np.testing.assert_allclose(plain.q, om.q, rtol=1e-13, atol=1e-15)
    parents = np.array([2, 2, 0xFFFFFFFF], np.uint32)
    children = np.array([[0, 1]], np.uint32)
    heights = np.array([0.0, 0.0, t])
    tree = (parents, children, heights)
    codes = np.array([[1, 2, 4, 8] * 4, [1] * 4 + [2] * 4 + [4] * 4 + [8] * 4], np.uint8)
    P = codes.shape[1]
    w = np.ones(P, np.uint32)
    rates = np.ones(1)
    ref, total_ref = _oracle(tree, om, codes, rates, w)
    ref_plain, _ = _oracle(tree, plain, codes, rates, w)
    # the two branches really differ by more than the tolerance
    assert (np.abs(ref - ref_plain) / (1e-11 + 1e-11 * np.abs(ref))).max() > 20
    got, total, _ = _device("f64", tree, GTR(f, r), codes, rates, w)
    _check("f64", got, ref, total, total_ref)
