"""Unstored cherries of the FP32 path on trees with long one-tip chains.

A full pass does not write the partials of cherries (both children
leaves); a later pass that reads one recomputes it from its two leaves'
codes and the matrices recorded for its slot.  These tests build trees of
two caterpillars joined at the root (one cherry at the bottom of each, long
chains of one-tip nodes above) and a Yule tree, and check, after every
accepted move, that the calculator's per-site values equal a fresh handle's
full evaluation of the same state bit for bit (the total to the reduction
order, 1e-13), and that its total matches the FP64 oracle.  Moves: one node
height (a one-chain update, the path kernel), two node heights in different
subtrees (a list of two chains, the HBM-prefetch variant of the fused
kernel), nephew swaps (topology changes, which move the cherries), with
about a third of the proposals rejected.
"""
import numpy as np
import pytest

from aspartik_amd import sim
from aspartik_amd.b3 import GTR, Calculator, DatedTree, gamma_rates
from oracle import oracle as O
from tests.helpers import random_codes

pytestmark = pytest.mark.gpu

F32_TOTAL = 1e-6


def caterpillar_pair(a, b, step=0.07):
    """Two caterpillar subtrees of a and b leaves joined at the root:
    (((l0, l1), l2), ...) on each side.  Returns parents, children, heights
    in the simulator's format (root 2L - 2)."""
    L = a + b
    N = 2 * L - 1
    parents = np.full(N, sim.TOMBSTONE, np.uint32)
    children = np.zeros((L - 1, 2), np.uint32)
    heights = np.zeros(N, np.float64)
    nxt = [L]

    def ladder(leaves):
        cur = leaves[0]
        h = 0.0
        for leaf in leaves[1:]:
            node = nxt[0]
            nxt[0] += 1
            h += step
            children[node - L] = (cur, leaf)
            parents[cur] = node
            parents[leaf] = node
            heights[node] = h
            cur = node
        return cur

    left = ladder(list(range(a)))
    right = ladder(list(range(a, L)))
    root = nxt[0]
    assert root == 2 * L - 2
    children[root - L] = (left, right)
    parents[left] = root
    parents[right] = root
    heights[root] = max(heights[left], heights[right]) + step
    return parents, children, heights


def _times(tree, children):
    flat = children.reshape(-1).astype(np.int64)
    return tree.heights[tree.parents[flat]] - tree.heights[flat]


def _full(calc_args, tree, codes, P, K, rates, pm):
    fresh = Calculator.HIP().make(codes, P, K)
    fresh.set_model(pm.q_matrix(), pm.frequencies(), True)
    t2 = tree.copy()
    t2.mark_all_edges_updated()
    internals, children = t2.partials_lists()
    fresh.propose_times(internals, children, _times(t2, children), rates)
    sites = [fresh.likelihoods(c) for c in range(K)]
    total = fresh.total()
    fresh.close()
    return sites, total


def _oracle_total(tree, om, codes, P, K, alpha):
    ot = O.Tree(tree.L, tree.parents_u32(), tree.children.astype(np.uint32), tree.heights)
    lik = O.GammaLikelihood(ot, om, codes, P, np.ones(P, np.uint32), num_categories=K,
                            alpha=alpha)
    return lik.likelihood()


@pytest.mark.parametrize("shape", [(7, 6), (9, 8), "yule40"])
def test_cherry_trees_recomputed_bitwise(hip_device, shape):
    K, P, alpha = 4, 3000, 0.5
    rng = np.random.default_rng(61)
    if shape == "yule40":
        parents, children, heights = sim.yule_tree(40, rng, 1.0, 0.15)
    else:
        parents, children, heights = caterpillar_pair(*shape)
    L = len(children) + 1
    f = rng.dirichlet([10] * 4)
    r = rng.dirichlet([1] * 6)
    pm, om = GTR(f, r), O.Model.gtr(r, f)
    codes = random_codes(L, P, rng, unknown_frac=0.05, ambiguous=True)
    rates = gamma_rates(alpha, K)
    tree = DatedTree(L, parents, children, heights)
    calc = Calculator.HIP().make(codes, P, K)
    calc.set_model(pm.q_matrix(), pm.frequencies(), True)
    tree.mark_all_edges_updated()
    internals, ch = tree.partials_lists()
    calc.propose_times(internals, ch, _times(tree, ch), rates)
    calc.accept()
    tree.accept()

    def move_height(node):
        kids = tree.children_of(node)
        lo = max(tree.heights[kids[0]], tree.heights[kids[1]])
        par = tree.parent_of(node)
        hi = tree.heights[par] if par is not None else lo + 0.3
        tree.set_height(node, lo + (hi - lo) * float(rng.uniform(0.1, 0.9)))

    def subtree(node):
        out, stack = set(), [node]
        while stack:
            v = stack.pop()
            out.add(v)
            if v >= L:
                stack.extend(tree.children_of(v))
        return out

    kinds = {"one": 0, "two": 0, "swap": 0}
    for step in range(40):
        kind = ["one", "two", "swap"][step % 3]
        root = tree.root
        inner = [v for v in range(L, 2 * L - 1) if v != root]
        if kind == "one":
            move_height(int(rng.choice(inner)))
        elif kind == "two":
            a = int(rng.choice(inner))
            rest = [v for v in inner if v not in subtree(a) and a not in subtree(v)]
            move_height(a)
            if rest:
                move_height(int(rng.choice(rest)))
        else:
            done = False
            for _ in range(50):
                node = int(rng.integers(0, 2 * L - 1))
                par = tree.parent_of(node)
                if par is None:
                    continue
                gp = tree.parent_of(par)
                if gp is None:
                    continue
                g0, g1 = tree.children_of(gp)
                uncle = g0 if g1 == par else g1
                if tree.heights[uncle] >= tree.heights[par]:
                    continue
                tree.replace_child(par, node, uncle)
                tree.replace_child(gp, uncle, node)
                done = True
                break
            if not done:
                move_height(int(rng.choice(inner)))
        internals, ch = tree.partials_lists()
        calc.propose_times(internals, ch, _times(tree, ch), rates)
        if rng.uniform() < 0.33:
            calc.reject()
            tree.reject()
            continue
        calc.accept()
        tree.accept()
        kinds[kind] += 1
        sites, total = _full(None, tree, codes, P, K, rates, pm)
        for c in range(K):
            np.testing.assert_array_equal(calc.likelihoods(c), sites[c], err_msg=f"step {step}")
        # (the total's reduction order follows the kernel's tiles)
        assert abs(calc.total() - total) <= 1e-13 * abs(total), (step, calc.total(), total)
        if step % 8 == 0:
            ref = _oracle_total(tree, om, codes, P, K, alpha)
            assert abs(total - ref) <= F32_TOTAL * abs(ref), (step, total, ref)
    print(f"[cherry trees] {shape}: accepted moves {kinds}", flush=True)
    calc.close()
