"""GPU parity: the HIP calculators against the FP64 oracle on the same seeded
trees, alignments and models.

Tolerances: FP32-scaled path, per-site |d lnL| <= 2e-6 * max(1, |lnL|) and
total relative error <= 1e-6 (the stated target); FP64 plain path, per-site
relative 2e-15 with host matrices and 1e-11 with the device-built P(t).
"""
import numpy as np
import pytest

from aspartik_amd import B3hError, sim
from aspartik_amd.b3 import GTR, HKY, Calculator, DatedTree, GammaLikelihood, gamma_rates
from oracle import oracle as O
from tests.helpers import model_matrices, random_codes, random_tree

pytestmark = pytest.mark.gpu

F32_SITE = 2e-6
F32_TOTAL = 1e-6


def _setup(L, P, K, seed, mean_branch=0.1, unknown=0.05, model="gtr"):
    rng = np.random.default_rng(seed)
    parents, children, heights = random_tree(L, rng, mean_branch)
    f = rng.dirichlet([10] * 4)
    if model == "gtr":
        r = rng.dirichlet([1] * 6)
        om = O.Model.gtr(r, f)
        pm = GTR(f, r)
    else:
        kappa = float(rng.lognormal(1.0, 0.5))
        om = O.Model.hky(kappa, f)
        pm = HKY(f, kappa)
    codes = random_codes(L, P, rng, unknown_frac=unknown, ambiguous=True)
    weights = rng.integers(1, 6, size=P).astype(np.uint32)
    return rng, (parents, children, heights), om, pm, codes, weights


def _oracle_sites(tree_arrays, om, codes, rates, internals=None, children=None):
    parents, ch, heights = tree_arrays
    L, P = codes.shape
    t = O.Tree(L, parents, ch, heights)
    t.mark_all_edges_updated()
    internals, children = t.partials_lists()
    mats = model_matrices(om, t, children, rates)
    out = []
    for c in range(len(rates)):
        calc = O.Calc4(codes, P, 30, 0)
        calc.propose(internals, children, mats[c], om.frequencies())
        out.append(calc.likelihoods())
    return np.array(out)


def _mix(sites, weights):
    K = sites.shape[0]
    mx = sites.max(0)
    mixed = mx + np.log(np.exp(sites - mx).sum(0)) + np.log(1.0 / K)
    return mixed, float(np.sum(mixed * weights))


def _assert_f32(got, ref):
    err = np.abs(got - ref) / np.maximum(1.0, np.abs(ref))
    assert err.max() <= F32_SITE, f"max per-site error {err.max():.3e}"


@pytest.mark.parametrize("L,P,K", [(2, 1, 1), (3, 7, 1), (5, 64, 4), (8, 65, 4), (16, 1000, 1),
                                   (16, 1000, 4), (40, 3001, 4), (7, 200, 3)])
def test_f32_matches_oracle_host_matrices(hip_device, L, P, K):
    rng, tree, om, pm, codes, w = _setup(L, P, K, seed=L * 100 + P)
    rates = O.gamma_categories(0.5, K) if K > 1 else np.ones(1)
    ref = _oracle_sites(tree, om, codes, rates)
    t = O.Tree(L, *tree)
    t.mark_all_edges_updated()
    internals, children = t.partials_lists()
    mats = model_matrices(om, t, children, rates)
    calc = Calculator.HIP().make(codes, P, K)
    calc.set_weights(w)
    calc.propose(internals, children, mats, om.frequencies())
    for c in range(K):
        _assert_f32(calc.likelihoods(c), ref[c])
    mixed_ref, total_ref = _mix(ref, w)
    _assert_f32(calc.site_mixed(), mixed_ref)
    total = calc.total()
    assert abs(total - total_ref) <= F32_TOTAL * abs(total_ref)


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("model", ["gtr", "hky"])
def test_device_p_build_matches_oracle(hip_device, precision, model):
    L, P, K = 30, 2500, 4
    rng, tree, om, pm, codes, w = _setup(L, P, K, seed=7, model=model)
    rates_o = O.gamma_categories(0.5, K)
    rates_p = gamma_rates(0.5, K)
    np.testing.assert_allclose(rates_p, rates_o, rtol=1e-13)
    ref = _oracle_sites(tree, om, codes, rates_o)
    cfg = Calculator.HIP() if precision == "f32" else Calculator.HIPPlainF64()
    calc = cfg.make(codes, P, K)
    calc.set_weights(w)
    calc.set_model(pm.q_matrix(), pm.frequencies(), pm.special_cases)
    t = DatedTree(L, *tree)
    t.mark_all_edges_updated()
    internals, children = t.partials_lists()
    flat = children.reshape(-1).astype(np.int64)
    times = t.heights[t.parents[flat]] - t.heights[flat]
    calc.propose_times(internals, children, times, rates_p)
    for c in range(K):
        got = calc.likelihoods(c)
        if precision == "f32":
            _assert_f32(got, ref[c])
        else:  # device eigensystem vs host eigensystem: a few ulps in P(t)
            np.testing.assert_allclose(got, ref[c], rtol=1e-11, atol=1e-11)
    _, total_ref = _mix(ref, w)
    tol = F32_TOTAL if precision == "f32" else 1e-12
    assert abs(calc.total() - total_ref) <= tol * abs(total_ref)


@pytest.mark.parametrize("scale_ln", [3, 30, 300])
def test_f64_plain_bitwise_semantics(hip_device, scale_ln):
    """The FP64 plain kernel follows cpu4.rs's arithmetic order; with host
    matrices it agrees with the oracle to the last bits of the final log."""
    L, P, K = 60, 4000, 2
    rng, tree, om, pm, codes, w = _setup(L, P, K, seed=scale_ln, mean_branch=0.4)
    rates = O.gamma_categories(0.5, K)
    t = O.Tree(L, *tree)
    t.mark_all_edges_updated()
    internals, children = t.partials_lists()
    mats = model_matrices(om, t, children, rates)
    calc = Calculator.HIPPlainF64().with_scale(scale_ln).make(codes, P, K)
    calc.propose(internals, children, mats, om.frequencies())
    for c in range(K):
        oc = O.Calc4(codes, P, scale_ln, 0)
        oc.propose(internals, children, mats[c], om.frequencies())
        np.testing.assert_allclose(calc.likelihoods(c), oc.likelihoods(), rtol=2e-15, atol=2e-15)


def _move(tree: DatedTree, otree: O.Tree, rng):
    L = tree.L
    if rng.random() < 0.7:
        node = int(rng.integers(L, 2 * L - 1))
        kids = tree.children_of(node)
        lo = max(tree.heights[kids[0]], tree.heights[kids[1]])
        par = tree.parent_of(node)
        hi = tree.heights[par] if par is not None else lo + 0.5
        nh = lo + (hi - lo) * rng.uniform(0.05, 0.95)
        tree.set_height(node, nh)
        otree.set_height(node, nh)
        return
    # narrow exchange: swap a node with its uncle when heights allow
    for _ in range(20):
        node = int(rng.integers(0, 2 * L - 1))
        par = tree.parent_of(node)
        if par is None:
            continue
        gp = tree.parent_of(par)
        if gp is None:
            continue
        a, b = tree.children_of(gp)
        uncle = a if b == par else b
        if tree.heights[uncle] >= tree.heights[par]:
            continue
        tree.replace_child(par, node, uncle)
        tree.replace_child(gp, uncle, node)
        otree.replace_child(par, node, uncle)
        otree.replace_child(gp, uncle, node)
        return


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_partial_updates_accept_reject_sequence(hip_device, precision):
    """An MCMC-like sequence of height moves and nephew swaps, each evaluated
    through the partial update list and accepted or rejected; every step is
    compared with a full oracle recompute."""
    L, P, K = 50, 3000, 4
    rng, tree_arrays, om, pm, codes, w = _setup(L, P, K, seed=11, mean_branch=0.2)
    rates = gamma_rates(0.5, K)
    cfg = Calculator.HIP() if precision == "f32" else Calculator.HIPPlainF64()
    calc = cfg.make(codes, P, K)
    calc.set_weights(w)
    calc.set_model(pm.q_matrix(), pm.frequencies(), True)
    tree = DatedTree(L, *tree_arrays)
    otree = O.Tree(L, *tree_arrays)
    tree.mark_all_edges_updated()

    def evaluate():
        internals, children = tree.partials_lists()
        flat = children.reshape(-1).astype(np.int64)
        times = tree.heights[tree.parents[flat]] - tree.heights[flat]
        calc.propose_times(internals, children, times, rates)
        return internals.size

    evaluate()
    calc.accept()
    tree.accept()
    otree.accept()
    sizes = []
    for step in range(40):
        _move(tree, otree, rng)
        if not tree.is_changed():
            continue
        sizes.append(evaluate())
        got = np.array([calc.likelihoods(c) for c in range(K)])
        p_, c_, h_ = otree.get()
        ref = _oracle_sites((p_, c_, h_), om, codes, rates)
        if precision == "f32":
            for c in range(K):
                _assert_f32(got[c], ref[c])
        else:
            np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-11)
        _, total_ref = _mix(ref, w)
        tol = F32_TOTAL if precision == "f32" else 1e-12
        assert abs(calc.total() - total_ref) <= tol * abs(total_ref)
        if rng.random() < 0.5:
            calc.accept()
            tree.accept()
            otree.accept()
        else:
            calc.reject()
            tree.reject()
            otree.reject()
    assert min(sizes) < L - 1  # partial updates were exercised


def test_deep_tree_underflow_scaling(hip_device):
    """500 taxa with long branches: site lnL far below FP32's range, kept
    finite and exact by the power-of-two exponents."""
    L, P, K = 500, 2000, 4
    rng, tree, om, pm, codes, w = _setup(L, P, K, seed=3, mean_branch=0.6, unknown=0.0)
    rates = O.gamma_categories(0.5, K)
    ref = _oracle_sites(tree, om, codes, rates)
    assert ref.min() < -1500
    t = O.Tree(L, *tree)
    t.mark_all_edges_updated()
    internals, children = t.partials_lists()
    calc = Calculator.HIP().make(codes, P, K)
    calc.set_weights(w)
    calc.propose(internals, children, model_matrices(om, t, children, rates), om.frequencies())
    for c in range(K):
        _assert_f32(calc.likelihoods(c), ref[c])
    _, total_ref = _mix(ref, w)
    assert abs(calc.total() - total_ref) <= F32_TOTAL * abs(total_ref)


@pytest.mark.parametrize("tile,lanes", [(64, 1), (128, 1), (256, 1), (64, 2), (256, 2), (512, 2)])
def test_tile_shapes_are_bit_identical(hip_device, tile, lanes):
    L, P, K = 33, 5000, 2
    rng, tree, om, pm, codes, w = _setup(L, P, K, seed=5)
    t = O.Tree(L, *tree)
    t.mark_all_edges_updated()
    internals, children = t.partials_lists()
    rates = O.gamma_categories(0.5, K)
    mats = model_matrices(om, t, children, rates)
    base = Calculator.HIP().make(codes, P, K)
    base.propose(internals, children, mats, om.frequencies())
    other = Calculator.HIP(tile_patterns=tile, lane_patterns=lanes).make(codes, P, K)
    other.propose(internals, children, mats, om.frequencies())
    for c in range(K):
        np.testing.assert_array_equal(other.likelihoods(c), base.likelihoods(c))
    # repeated proposes give the bit-identical total
    t1 = base.total()
    base.propose(internals, children, mats, om.frequencies())
    assert base.total() == t1


@pytest.mark.parametrize("L,P,K,tile", [(33, 5003, 2, 64), (33, 5003, 2, 512), (33, 4001, 4, 256),
                                         (33, 777, 1, 64), (400, 3000, 4, 256)])
def test_one_state_per_lane_is_bit_identical(hip_device, L, P, K, tile):
    """lane_states = 1 (four lanes per pattern, the quad exchanging states by
    DPP) computes the same site values as four states per lane, bit for bit:
    in full evaluations, in the partial updates that follow (which read the
    partials it stored, through the four-states kernels), with per-pattern
    classes and per-category models and branch lengths, and with the in-place reads of stored children that a deep
    update list needs once the LDS stack is full (400 taxa).  Totals agree to
    the rounding of the weighted sum's grouping (16 patterns per wavefront
    instead of 64 or 128)."""
    rng, tree_arrays, om, pm, codes, w = _setup(L, P, K, seed=L + P)
    rates = gamma_rates(0.5, K) if K > 1 else np.ones(1)

    def make(**kw):
        c = Calculator.HIP(**kw).make(codes, P, K)
        c.set_weights(w)
        c.set_model(pm.q_matrix(), pm.frequencies(), True)
        return c

    base, one = make(), make(tile_patterns=tile, lane_states=1)
    tree = DatedTree(L, *tree_arrays)
    tree.mark_all_edges_updated()

    def evaluate(c, t):
        internals, children = t.partials_lists()
        flat = children.reshape(-1).astype(np.int64)
        c.propose_times(internals, children, t.heights[t.parents[flat]] - t.heights[flat], rates)

    def same():
        for c in range(K):
            np.testing.assert_array_equal(one.likelihoods(c), base.likelihoods(c))
        np.testing.assert_array_equal(one.site_mixed(), base.site_mixed())
        assert abs(one.total() - base.total()) <= 1e-13 * abs(base.total())

    for c in (base, one):
        evaluate(c, tree.copy())
        c.accept()
    tree.accept()
    same()
    # height moves: partial updates of the stored state, then a rejected one
    for step in range(4):
        cand = [nd for nd in range(L, 2 * L - 1) if nd != tree.root]
        nd = int(rng.choice(cand))
        lo = max(tree.heights[ch] for ch in tree.children_of(nd))
        tree.set_height(nd, float(lo + (tree.heights[tree.parent_of(nd)] - lo) * 0.3))
        for c in (base, one):
            evaluate(c, tree.copy())
        same()
        for c in (base, one):
            c.accept() if step < 3 else c.reject()
        tree.accept() if step < 3 else tree.reject()
    # per-pattern classes (HeteroLikelihood) over a fresh full evaluation
    classes = rng.integers(0, K, size=P).astype(np.uint32)
    for c in (base, one):
        c.set_classes(classes)
        t2 = tree.copy()
        t2.mark_all_edges_updated()
        evaluate(c, t2)
    same()
    # per-category models and branch lengths (each category its own root
    # frequencies)
    subs = [GTR(rng.dirichlet([20.0] * 4), rng.dirichlet([4.0] * 6)) for _ in range(K)]
    t2 = tree.copy()
    t2.mark_all_edges_updated()
    internals, children = t2.partials_lists()
    flat = children.reshape(-1).astype(np.int64)
    times = t2.heights[t2.parents[flat]] - t2.heights[flat]
    lengths = np.stack([times * rng.uniform(0.5, 2.0, size=times.size) for _ in range(K)])
    for c in (base, one):
        for k, sub in enumerate(subs):
            c.set_category_model(k, sub.q_matrix(), sub.frequencies())
        c.propose_lengths(internals, children, lengths)
    same()
    base.close()
    one.close()


def test_protocol_errors(hip_device):
    L, P, K = 6, 10, 1
    rng, tree, om, pm, codes, w = _setup(L, P, K, seed=1)
    calc = Calculator.HIP().make(codes, P, K)
    with pytest.raises(B3hError) as e:
        calc.likelihoods(0)
    assert e.value.status == 5
    t = DatedTree(L, *tree)
    t.mark_all_edges_updated()
    internals, children = t.partials_lists()
    with pytest.raises(B3hError) as e:
        calc.propose_times(internals, children, np.ones(children.size), np.ones(1))
    assert e.value.status == 5
    mats = np.tile(np.eye(4), (1, children.size, 1, 1))
    bad = internals[::-1].copy()  # root first: children after parents
    bad_children = children[::-1].copy()
    with pytest.raises(B3hError) as e:
        calc.propose(bad, bad_children, mats, np.full(4, 0.25))
    assert e.value.status == 1
    with pytest.raises(B3hError):
        calc.propose(np.array([L + 100], np.uint32), children[:1], mats[:, :2], np.full(4, 0.25))
    with pytest.raises(B3hError):
        calc.likelihoods(3)


@pytest.mark.parametrize("K", [1, 4])
def test_likelihood_layer_against_oracle(hip_device, K):
    """GammaLikelihood / DNALikelihood mirror vs the oracle's likelihood
    layer (gamma.rs:97-203) through parameter changes and rejects."""
    L, P = 24, 1500
    rng, tree_arrays, om, pm, codes, w = _setup(L, P, K, seed=40 + K, model="hky")
    tree = DatedTree(L, *tree_arrays)
    otree = O.Tree(L, *tree_arrays)
    lik = GammaLikelihood(codes, w, tree, pm, num_categories=K, alpha=0.5,
                          calculator=Calculator.HIP())
    olik = O.GammaLikelihood(otree, om, codes, P, w, num_categories=K, alpha=0.5)
    a, b = lik.likelihood(), olik.likelihood()
    assert abs(a - b) <= F32_TOTAL * abs(b)
    for step in range(12):
        _move(tree, otree, rng)
        if K > 1 and step % 4 == 1:
            alpha = float(rng.uniform(0.2, 2.0))
            lik.set_alpha(alpha)
            olik.set_alpha(alpha)
        a, b = lik.likelihood(), olik.likelihood()
        assert abs(a - b) <= F32_TOTAL * abs(b), (step, a, b)
        if rng.random() < 0.5:
            lik.accept(); olik.accept(); tree.accept(); otree.accept()
        else:
            lik.reject(); olik.reject(); tree.reject(); otree.reject()


@pytest.mark.parametrize("K", [1, 3, 4, 16])
def test_path_updates_match_full_evaluation_bitwise(hip_device, K):
    """One-edge updates run in the path kernel (every in-pass child is the
    previous entry); their site values equal a fresh handle's full
    evaluation of the same state bit for bit, since both round each node
    the same way.  K sets the path kernel's workgroup (16 / K sub-tiles of
    64 patterns: 1024, 960, 1024 and 1024 threads)."""
    L, P = 60, 4000
    rng, tree_arrays, om, pm, codes, w = _setup(L, P, K, seed=23, mean_branch=0.2)
    rates = gamma_rates(0.5, K) if K > 1 else np.ones(1)
    calc = Calculator.HIP().make(codes, P, K)
    calc.set_model(pm.q_matrix(), pm.frequencies(), True)
    tree = DatedTree(L, *tree_arrays)
    tree.mark_all_edges_updated()

    def evaluate(c, t):
        internals, children = t.partials_lists()
        flat = children.reshape(-1).astype(np.int64)
        c.propose_times(internals, children, t.heights[t.parents[flat]] - t.heights[flat], rates)

    evaluate(calc, tree)
    calc.accept()
    tree.accept()
    for step in range(12):
        cand = [nd for nd in range(L, 2 * L - 1) if nd != tree.root]
        nd = int(rng.choice(cand))
        lo = max(tree.heights[ch] for ch in tree.children_of(nd))
        tree.set_height(nd, float(lo + (tree.heights[tree.parent_of(nd)] - lo) * 0.4))
        evaluate(calc, tree)
        calc.accept()
        fresh = Calculator.HIP().make(codes, P, K)
        fresh.set_model(pm.q_matrix(), pm.frequencies(), True)
        t2 = tree.copy()
        t2.mark_all_edges_updated()
        evaluate(fresh, t2)
        for c in range(K):
            np.testing.assert_array_equal(calc.likelihoods(c), fresh.likelihoods(c))
        assert abs(calc.total() - fresh.total()) <= 1e-13 * abs(fresh.total())
        tree.accept()


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_total_over_many_tiles_equals_site_sum(hip_device, precision):
    """The device total comes from a last-tile reduction whose ticket orders
    each tile's partial sum with agent-scope atomics and no cache
    write-back.  At 200,000 patterns a full evaluation has about 400 tiles,
    a one-edge update in the path kernel 782, and the FP64 mix kernel 782:
    over repeated evaluations, accepted and rejected, every total equals the
    FP64 weighted sum of the per-pattern values the same evaluation wrote
    (to summation-order rounding), so no tile's partial was missed or
    stale."""
    L, P, K = 24, 200_000, 4
    rng, tree_arrays, om, pm, codes, w = _setup(L, P, K, seed=31, unknown=0.02)
    rates = gamma_rates(0.5, K)
    calc = (Calculator.HIP() if precision == "f32" else Calculator.HIPPlainF64()).make(codes, P, K)
    calc.set_weights(w)
    calc.set_model(pm.q_matrix(), pm.frequencies(), True)
    tree = DatedTree(L, *tree_arrays)

    def evaluate():
        internals, children = tree.partials_lists()
        flat = children.reshape(-1).astype(np.int64)
        calc.propose_times(internals, children,
                           tree.heights[tree.parents[flat]] - tree.heights[flat], rates)
        total = calc.total()
        expect = float(np.dot(calc.site_mixed(), w.astype(np.float64)))
        assert abs(total - expect) <= 1e-12 * abs(expect), (total, expect)
        return total

    tree.mark_all_edges_updated()
    evaluate()
    calc.accept()
    tree.accept()
    for step in range(8):
        cand = [nd for nd in range(L, 2 * L - 1) if nd != tree.root]
        nd = int(rng.choice(cand))
        lo = max(tree.heights[ch] for ch in tree.children_of(nd))
        tree.set_height(nd, float(lo + (tree.heights[tree.parent_of(nd)] - lo) * 0.6))
        evaluate()
        if step % 3 == 2:
            calc.reject()
            tree.reject()
        else:
            calc.accept()
            tree.accept()
