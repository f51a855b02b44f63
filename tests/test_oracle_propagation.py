"""Pins the oracle's restatement of b3's propagation calculator
(crates/b3/src/calculator/cuda_propagations/kernels.cu:73-187, mod.rs:77-164):

* full evaluations against brute-force enumeration (<= 6 taxa) and against
  the Cpu4 restatement (1e-12) at scale_ln 3, 30 and 300 (the propagation
  kernel also scales cherries, so its flags differ from Cpu4's; its values
  must not);
* update lists from DatedTree::propagation_lists under height and topology
  moves with random accept/reject equal a fresh full recompute, and a
  rejected proposal restores the accepted values bit-exactly;
* the lists themselves: updated leaves first, each changed edge's child
  node, all its ancestors, root last.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import brute_force_site_ll, model_matrices, random_codes, random_tree


def node_matrices(model, tree: O.Tree, nodes):
    """One transition per listed node but the root, P(edge length of node)."""
    return np.array([model.transition(tree.edge_length(int(n))) for n in nodes[:-1]])


def prop_full(tree, model, codes, scale_ln=30):
    tree.mark_all_edges_updated()
    nodes, ch, nl = tree.propagation_lists()
    c = O.Prop4(codes, codes.shape[1], scale_ln)
    c.propose(nodes, ch, nl, node_matrices(model, tree, nodes), model.frequencies())
    return c


def calc4_full(tree, model, codes, scale_ln=30):
    tree.mark_all_edges_updated()
    internals, children = tree.partials_lists()
    c = O.Calc4(codes, codes.shape[1], scale_ln, 1)
    c.propose(internals, children, model_matrices(model, tree, children, [1.0])[0],
              model.frequencies())
    return c.likelihoods()


@pytest.mark.parametrize("L", [2, 3, 4, 6])
def test_prop4_matches_brute_force(L):
    rng = np.random.default_rng(L)
    par, ch, h = random_tree(L, rng, mean_branch=0.3)
    tree = O.Tree(L, par, ch, h)
    f = rng.dirichlet([5] * 4)
    model = O.Model.gtr(rng.dirichlet([1] * 6), f)
    codes = random_codes(L, 20, rng, unknown_frac=0.2, ambiguous=True)
    site = prop_full(tree, model, codes).likelihoods()
    for i in range(codes.shape[1]):
        bf = brute_force_site_ll(par, h, codes[:, i], model.transition, f)
        assert abs(site[i] - bf) <= 1e-12 * max(1.0, abs(bf))


@pytest.mark.parametrize("scale_ln", [3, 30, 300])
def test_prop4_equals_calc4(scale_ln):
    rng = np.random.default_rng(scale_ln)
    L = 90
    par, ch, h = random_tree(L, rng, mean_branch=0.5)
    tree = O.Tree(L, par, ch, h)
    model = O.Model.hky(3.0, rng.dirichlet([10] * 4))
    codes = random_codes(L, 300, rng, unknown_frac=0.05)
    a = prop_full(tree, model, codes, scale_ln).likelihoods()
    b = calc4_full(tree, model, codes, scale_ln)
    assert b.min() < -100  # deep enough that scaling fires
    np.testing.assert_allclose(a, b, rtol=1e-12)


def _height_move(tree, rng):
    p, c, h = tree.get()
    L = tree.L
    node = int(rng.integers(L, 2 * L - 1))
    kids = c[node - L]
    lo = max(h[kids[0]], h[kids[1]])
    hi = h[p[node]] if p[node] != O.TOMBSTONE else lo + 1.0
    tree.set_height(node, lo + (hi - lo) * rng.uniform(0.05, 0.95))


def _narrow_exchange(tree, rng):
    """Swap a node's child with its uncle when the heights allow (b3's
    NarrowExchange, operators/_tree_exchange.py:23-57, via replace_child)."""
    p, c, h = tree.get()
    L = tree.L
    for _ in range(50):
        node = int(rng.integers(L, 2 * L - 1))
        parent = p[node]
        if parent == O.TOMBSTONE:
            continue
        kids = c[parent - L]
        uncle = int(kids[0] if kids[1] == node else kids[1])
        child = int(c[node - L][rng.integers(0, 2)])
        if h[uncle] >= h[node]:
            continue
        tree.replace_child(int(parent), uncle, child)
        tree.replace_child(node, child, uncle)
        return True
    return False


def test_propagation_lists_shape():
    rng = np.random.default_rng(5)
    L = 30
    par, ch, h = random_tree(L, rng)
    tree = O.Tree(L, par, ch, h)
    tree.accept()
    leaf = 7
    tree.mark_edge_updated(leaf)
    nodes, children, nl = tree.propagation_lists()
    assert nl == 1 and nodes[0] == leaf
    anc, n = [], leaf
    while par[n] != O.TOMBSTONE:
        n = int(par[n])
        anc.append(n)
    assert set(nodes[1:].tolist()) == set(anc) and nodes[-1] == tree.root
    tree.accept()
    inner = int(next(i for i in range(L, 2 * L - 1) if par[i] != O.TOMBSTONE))
    tree.mark_edge_updated(inner)
    nodes, children, nl = tree.propagation_lists()
    assert nl == 0 and inner in nodes.tolist()   # the changed edge's child node itself
    assert children.shape == (nodes.size, 2)


@pytest.mark.parametrize("seed", range(3))
def test_prop4_partial_updates_and_reject(seed):
    rng = np.random.default_rng(100 + seed)
    L, P = 40, 400
    par, ch, h = random_tree(L, rng, mean_branch=0.4)
    tree = O.Tree(L, par, ch, h)
    model = O.Model.hky(2.5, rng.dirichlet([10] * 4))
    codes = random_codes(L, P, rng, unknown_frac=0.05)
    f = model.frequencies()
    calc = prop_full(tree, model, codes, scale_ln=3)
    calc.accept()
    tree.accept()
    accepted = calc.likelihoods()
    for step in range(40):
        if step % 3 == 2:
            _narrow_exchange(tree, rng)
        else:
            _height_move(tree, rng)
        nodes, children, nl = tree.propagation_lists()
        calc.propose(nodes, children, nl, node_matrices(model, tree, nodes), f)
        got = calc.likelihoods()
        p2, c2, h2 = tree.get()
        full = calc4_full(O.Tree(L, p2, c2, h2), model, codes, 3)
        np.testing.assert_allclose(got, full, rtol=1e-12, atol=1e-12)
        if rng.random() < 0.5:
            calc.accept()
            tree.accept()
            accepted = got
        else:
            calc.reject()
            tree.reject()
            tree.mark_all_edges_updated()
            nodes, children, nl = tree.propagation_lists()
            calc.propose(nodes, children, nl, node_matrices(model, tree, nodes), f)
            np.testing.assert_array_equal(calc.likelihoods(), accepted)
            calc.accept()
            tree.accept()
