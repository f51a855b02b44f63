"""Batched evaluation (b3h_batch_*): every instance equals its own
single-instance calculator bit for bit per site, and the FP64 oracle to the
FP32 tolerance; partial updates, accept/reject and graph replays included."""
import numpy as np
import pytest

from aspartik_amd import B3hError
from aspartik_amd.b3 import HKY, Batch, Calculator, DatedTree, gamma_rates
from oracle import oracle as O
from tests.helpers import model_matrices, random_codes, random_tree

pytestmark = pytest.mark.gpu

F32_SITE = 2e-6


def _instances(seed, sizes):
    rng = np.random.default_rng(seed)
    out = []
    for L, P in sizes:
        par, ch, h = random_tree(L, rng, mean_branch=0.1)
        kappa = float(rng.lognormal(1.0, 0.5))
        f = rng.dirichlet([10] * 4)
        codes = random_codes(L, P, rng, unknown_frac=0.03, ambiguous=True)
        w = rng.integers(1, 4, size=P).astype(np.uint32)
        out.append(dict(tree=DatedTree(L, par, ch, h), arrays=(par, ch, h), model=HKY(f, kappa),
                        omodel=O.Model.hky(kappa, f), codes=codes, w=w))
    return out


def _full_update(inst, rates):
    t = inst["tree"]
    t.mark_all_edges_updated()
    internals, children = t.partials_lists()
    t.accept()
    flat = children.reshape(-1).astype(np.int64)
    times = t.heights[t.parents[flat]] - t.heights[flat]
    return internals, children, times, rates


def _oracle_sites(inst, rates):
    par, ch, h = inst["tree"].parents_u32(), inst["tree"].children, inst["tree"].heights
    L, P = inst["codes"].shape
    t = O.Tree(L, par, ch.astype(np.uint32), h)
    t.mark_all_edges_updated()
    internals, children = t.partials_lists()
    mats = model_matrices(inst["omodel"], t, children, rates)
    out = []
    for c in range(len(rates)):
        calc = O.Calc4(inst["codes"], P, 30, 0)
        calc.propose(internals, children, mats[c], inst["omodel"].frequencies())
        out.append(calc.likelihoods())
    return np.array(out)


SIZES = [(3, 1), (5, 63), (12, 700), (40, 2000), (120, 1500), (64, 513), (7, 4000), (90, 64)]


@pytest.mark.parametrize("single", [False, True])
def test_batch_equals_single_instances(hip_device, single):
    K = 4
    rates = gamma_rates(0.5, K)
    insts = _instances(1, SIZES)
    b = Batch([(x["codes"], x["w"]) for x in insts], num_categories=K, single_slot=single)
    for i, x in enumerate(insts):
        b.set_model(i, x["model"].q_matrix(), x["model"].frequencies())
    b.propose_times([_full_update(x, rates) for x in insts])
    totals = b.totals()
    for i, x in enumerate(insts):
        L, P = x["codes"].shape
        c = Calculator.HIP().make(x["codes"], P, K)
        c.set_weights(x["w"])
        c.set_model(x["model"].q_matrix(), x["model"].frequencies())
        c.propose_times(*_full_update(x, rates))
        ref = _oracle_sites(x, rates)
        for k in range(K):
            got = b.likelihoods(i, k)
            np.testing.assert_array_equal(got, c.likelihoods(k))
            err = np.abs(got - ref[k]) / np.maximum(1.0, np.abs(ref[k]))
            assert err.max() <= F32_SITE
        assert abs(totals[i] - c.total()) <= 1e-12 * abs(c.total())
    if single:
        with pytest.raises(B3hError):
            b.reject()


def test_batch_partial_updates_accept_reject(hip_device):
    K = 4
    rates = gamma_rates(0.5, K)
    insts = _instances(2, [(30, 900), (50, 1200), (9, 300), (77, 2000)])
    b = Batch([(x["codes"], x["w"]) for x in insts], num_categories=K)
    for i, x in enumerate(insts):
        b.set_model(i, x["model"].q_matrix(), x["model"].frequencies())
    b.propose_times([_full_update(x, rates) for x in insts])
    b.accept()
    rng = np.random.default_rng(3)
    for step in range(12):
        ups = []
        for x in insts:
            t = x["tree"]
            if rng.random() < 0.4:
                ups.append(None)
                continue
            node = int(rng.integers(t.L, 2 * t.L - 1))
            a, c = t.children_of(node)
            lo = max(t.heights[a], t.heights[c])
            par = t.parent_of(node)
            hi = t.heights[par] if par is not None else lo + 0.3
            t.set_height(node, lo + (hi - lo) * rng.uniform(0.1, 0.9))
            internals, children = t.partials_lists()
            flat = children.reshape(-1).astype(np.int64)
            ups.append((internals, children, t.heights[t.parents[flat]] - t.heights[flat], rates))
        if all(u is None for u in ups):
            continue
        b.propose_times(ups)
        for i, x in enumerate(insts):
            if ups[i] is None:
                continue
            ref = _oracle_sites(x, rates)
            for k in range(K):
                got = b.likelihoods(i, k)
                err = np.abs(got - ref[k]) / np.maximum(1.0, np.abs(ref[k]))
                assert err.max() <= F32_SITE, (step, i, k, err.max())
        if rng.random() < 0.5:
            b.accept()
            for x in insts:
                x["tree"].accept()
        else:
            b.reject()
            for x in insts:
                x["tree"].reject()


def test_batch_graph_replay_matches_eager(hip_device):
    """Captured-graph replays, with eager work and syncs between them, give
    the eager results every time, also after the edge times change."""
    K = 4
    rates = gamma_rates(0.5, K)
    insts = _instances(4, [(25, 1000), (60, 3000), (8, 200)])
    b = Batch([(x["codes"], x["w"]) for x in insts], num_categories=K)
    for i, x in enumerate(insts):
        b.set_model(i, x["model"].q_matrix(), x["model"].frequencies())
    ups = [_full_update(x, rates) for x in insts]
    b.propose_times(ups)
    eager = b.totals()
    for rep in range(4):
        b.propose_times(ups, graph=True)
        np.testing.assert_array_equal(b.totals(), eager)
        b.accept()
        b.propose_times(ups)  # eager work between replays
        np.testing.assert_array_equal(b.totals(), eager)
        b.accept()
    # same list shapes, new branch lengths: the replay must read the new data
    scaled = [(u[0], u[1], u[2] * 1.3, u[3]) for u in ups]
    b.propose_times(scaled)
    eager2 = b.totals()
    assert not np.array_equal(eager2, eager)
    b.accept()
    b.propose_times(scaled, graph=True)
    np.testing.assert_array_equal(b.totals(), eager2)


def test_batch_autotune_picks_a_candidate_with_identical_sites(hip_device):
    """The batch tile autotuner times every candidate on the caller's own
    instances and returns the fastest; whichever it picks, per-site values
    equal the default shape's bit for bit (only the schedule differs)."""
    from aspartik_amd.b3 import autotune_batch
    from aspartik_amd.b3.tuning import batch_candidates
    K = 4
    rates = gamma_rates(0.5, K)
    insts = _instances(7, [(30, 1500), (80, 700), (12, 3000)])
    data = [(x["codes"], x["w"]) for x in insts]
    models = [(x["model"].q_matrix(), x["model"].frequencies(), True) for x in insts]
    ups = [_full_update(x, rates) for x in insts]
    (tile, lanes, states), trials = autotune_batch(data, models, ups, K, evals=2)
    assert (tile, lanes, states) in batch_candidates(K)
    assert len(trials) == len(batch_candidates(K)) and trials[0].ms <= trials[-1].ms
    results = []
    # the default shape, the tuner's pick and one state per lane
    for tp, lp, ls in [(0, 0, 0), (tile, lanes, states), (128, 1, 1)]:
        b = Batch(data, num_categories=K, single_slot=True, tile_patterns=tp, lane_patterns=lp,
                  lane_states=ls)
        for i, (q, f, sc) in enumerate(models):
            b.set_model(i, q, f, sc)
        b.propose_times(ups)
        results.append([[b.likelihoods(i, k) for k in range(K)] for i in range(len(insts))])
        totals = b.totals()
        b.close()
        assert np.all(np.isfinite(totals))
    for other in results[1:]:
        for a, c in zip(results[0], other):
            for x, y in zip(a, c):
                np.testing.assert_array_equal(x, y)
