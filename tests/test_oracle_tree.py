"""Pins the oracle's update schedule (crates/b3/src/parameters/tree.rs) and
EditBuf (crates/sk/src/editbuf.rs), and checks the product's Python
DatedTree mirror against it."""
import numpy as np
import pytest

from aspartik_amd.b3.tree import DatedTree
from oracle import oracle as O
from tests.helpers import random_tree


def test_editbuf_transition_tables():
    """The bit tables in editbuf.rs's own comments (set_edited 191-201,
    reject 175-185, accept 168-173)."""
    for v, expect in [(0, 3), (1, 2), (2, 2), (3, 3)]:
        m = np.array([v], np.uint8)
        O.editbuf_set_edited(m, 0)
        assert m[0] == expect
    for v, expect in [(0, 0), (1, 1), (2, 1), (3, 0)]:
        m = np.array([v], np.uint8)
        O.editbuf_reject(m)
        assert m[0] == expect
    for v, expect in [(0, 0), (1, 1), (2, 0), (3, 1)]:
        m = np.array([v], np.uint8)
        O.editbuf_accept(m)
        assert m[0] == expect


def test_editbuf_two_slot_model():
    """Model: active slot, slot written this epoch; accept keeps the written
    slot, reject returns to the slot active at the epoch start."""
    rng = np.random.default_rng(0)
    n = 16
    m = np.zeros(n, np.uint8)
    active = np.zeros(n, int)
    start = active.copy()
    edited = np.zeros(n, bool)
    for _ in range(2000):
        op = rng.integers(0, 3)
        if op == 0:
            i = int(rng.integers(n))
            O.editbuf_set_edited(m, i)
            if not edited[i]:
                active[i] ^= 1
                edited[i] = True
        elif op == 1:
            O.editbuf_accept(m)
            edited[:] = False
            start = active.copy()
        else:
            O.editbuf_reject(m)
            active = start.copy()
            edited[:] = False
        np.testing.assert_array_equal(m & 1, active)
        np.testing.assert_array_equal((m >> 1) & 1, edited.astype(np.uint8))


def _ancestors(parents, node):
    out = []
    while parents[node] != O.TOMBSTONE:
        node = int(parents[node])
        out.append(node)
    return out


@pytest.mark.parametrize("seed", range(6))
def test_partials_lists_is_ancestor_closure(seed):
    rng = np.random.default_rng(seed)
    L = int(rng.integers(3, 60))
    parents, children, heights = random_tree(L, rng)
    t = O.Tree(L, parents, children, heights)
    edges = rng.choice(2 * L - 2, size=int(rng.integers(1, 4)), replace=False)
    for e in edges:
        t.mark_edge_updated(int(e))
    internals, ch = t.partials_lists()
    expect = set()
    for e in edges:
        expect |= set(_ancestors(parents, int(e)))
    expect.add(int(np.nonzero(parents == O.TOMBSTONE)[0][0]))
    assert set(internals.tolist()) == expect
    assert internals[-1] == t.root
    pos = {int(v): i for i, v in enumerate(internals)}
    for i, v in enumerate(internals):
        a, b = ch[i]
        assert a < b
        assert sorted(children[v - L].tolist()) == [a, b]
        for c in (a, b):
            if c in pos:
                assert pos[c] < i
    assert sorted(t.edges_to_update().tolist()) == sorted(set(int(e) for e in edges))


@pytest.mark.parametrize("seed", range(6))
def test_python_mirror_matches_oracle_schedule(seed):
    """The product's DatedTree gives the same update lists as the oracle
    under random height moves, subtree swaps, accept and reject."""
    rng = np.random.default_rng(seed)
    L = int(rng.integers(4, 40))
    parents, children, heights = random_tree(L, rng)
    o = O.Tree(L, parents, children, heights)
    m = DatedTree(L, parents, children, heights)
    for _ in range(60):
        r = rng.random()
        if r < 0.5:
            p, c, h = o.get()
            node = int(rng.integers(L, 2 * L - 1))
            kids = c[node - L]
            lo = max(h[kids[0]], h[kids[1]])
            hi = h[p[node]] if p[node] != O.TOMBSTONE else lo + 1
            nh = lo + (hi - lo) * rng.uniform(0.1, 0.9)
            o.set_height(node, nh)
            m.set_height(node, nh)
        elif r < 0.7:
            # exchange two nephews: replace_child on both parents (tree.rs:535-550 style)
            p, c, h = o.get()
            node = int(rng.integers(0, 2 * L - 1))
            if p[node] == O.TOMBSTONE or p[p[node]] == O.TOMBSTONE:
                continue
            par = int(p[node])
            gp = int(p[par])
            kids = c[gp - L]
            uncle = int(kids[0] if kids[1] == par else kids[1])
            if h[uncle] >= h[par] or h[node] >= h[gp]:
                continue
            o.replace_child(par, node, uncle)
            o.replace_child(gp, uncle, node)
            m.replace_child(par, node, uncle)
            m.replace_child(gp, uncle, node)
        elif r < 0.85:
            o.accept()
            m.accept()
        else:
            o.reject()
            m.reject()
        oi, oc = o.partials_lists()
        mi, mc = m.partials_lists()
        np.testing.assert_array_equal(oi, mi)
        np.testing.assert_array_equal(oc, mc)
        on, ocp, onl = o.propagation_lists()
        mn, mcp, mnl = m.propagation_lists()
        np.testing.assert_array_equal(on, mn)
        np.testing.assert_array_equal(ocp, mcp)
        assert onl == mnl
        assert o.is_changed() == m.is_changed()
