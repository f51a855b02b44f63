"""Dated tree with b3's dirty-edge bookkeeping and update lists.

Mirrors the parts of ``DatedTree`` (crates/b3/src/parameters/tree.rs) that
feed the likelihood: node numbering (leaves 0..L-1, internals L..2L-2, edge
index = child id, tree.rs:728-743), ``mark_edge_updated`` (357-375),
``mark_all_edges_updated`` (382-386), ``partials_lists`` and
``propagation_lists`` (398-461),
``set_height`` (516-528), ``replace_child`` (471-495) and accept/reject
(975-995).
"""
from __future__ import annotations

from collections import deque

import numpy as np

TOMBSTONE = 0xFFFFFFFF


class DatedTree:
    def __init__(self, num_leaves, parents, children, heights):
        self.L = int(num_leaves)
        self.N = 2 * self.L - 1
        self.parents = np.array(parents, np.int64)
        self.children = np.array(children, np.int64).reshape(self.L - 1, 2)
        self.heights = np.array(heights, np.float64)
        self.parents[self.parents == TOMBSTONE] = -1
        roots = np.nonzero(self.parents < 0)[0]
        if roots.size != 1:
            raise ValueError("tree must have exactly one root")
        self.root = int(roots[0])
        self._upd_edges = np.zeros(self.N, bool)
        self._upd_partials = np.zeros(self.N, bool)
        self._upd_props = np.zeros(self.N, bool)
        self._save()

    # ---- structure -------------------------------------------------------
    @property
    def num_leaves(self):
        return self.L

    @property
    def num_nodes(self):
        return self.N

    def children_of(self, node):
        return int(self.children[node - self.L, 0]), int(self.children[node - self.L, 1])

    def parent_of(self, node):
        p = int(self.parents[node])
        return None if p < 0 else p

    def edge_length(self, edge):
        """tree.rs:732-736"""
        return self.heights[self.parents[edge]] - self.heights[edge]

    def edges(self):
        """Every node but the root (tree.rs:833-836)."""
        return [e for e in range(self.N) if e != self.root]

    # ---- dirty tracking ---------------------------------------------------
    def mark_edge_updated(self, edge):
        """tree.rs:357-375: the edge and its child node's projection, then
        ancestors up to the first marked."""
        self._upd_edges[edge] = True
        self._upd_props[edge] = True
        curr = int(edge)
        while self.parents[curr] >= 0:
            curr = int(self.parents[curr])
            if self._upd_partials[curr]:
                break
            self._upd_props[curr] = True
            self._upd_partials[curr] = True

    def mark_all_edges_updated(self):
        self._upd_edges[:] = True
        self._upd_partials[:] = True
        self._upd_props[:] = True

    def is_changed(self):
        return bool(self._upd_edges.any())

    def edges_to_update(self):
        return [e for e in self.edges() if self._upd_edges[e]]

    def _updated_internals(self, bitmap):
        """tree.rs:398-421: BFS from the root through marked internal
        children, reversed (children first, root last)."""
        out = [self.root]
        q = deque([self.root])
        while q:
            node = q.popleft()
            left, right = self.children_of(node)
            for ch in (left, right):
                if bitmap[ch] and ch >= self.L:
                    out.append(ch)
                    q.append(ch)
        out.reverse()
        return np.asarray(out, np.uint32)

    def _children_list(self, internals):
        """tree.rs:423-434: each pair sorted ascending."""
        ch = self.children[internals.astype(np.int64) - self.L]
        return np.sort(ch, axis=1).astype(np.uint32)

    def partials_lists(self):
        """tree.rs:456-461"""
        internals = self._updated_internals(self._upd_partials)
        return internals, self._children_list(internals)

    def propagation_lists(self):
        """tree.rs:436-454: (nodes, children, number of updated leaves); the
        updated leaves in order, then the internal nodes whose projections
        change, children first, root last."""
        leaves = np.nonzero(self._upd_props[:self.L])[0].astype(np.uint32)
        internals = self._updated_internals(self._upd_props)
        nodes = np.concatenate([leaves, internals]).astype(np.uint32)
        return nodes, self._children_list(internals), int(leaves.size)

    # ---- edits -----------------------------------------------------------
    def set_height(self, node, height):
        """tree.rs:516-528"""
        self.heights[node] = height
        if self.parents[node] >= 0:
            self.mark_edge_updated(node)
        if node >= self.L:
            left, right = self.children_of(node)
            self.mark_edge_updated(left)
            self.mark_edge_updated(right)

    def replace_child(self, parent, old_child, new_child):
        """tree.rs:471-495"""
        self.parents[new_child] = parent
        left, right = self.children_of(parent)
        if old_child == left:
            self.children[parent - self.L, 0] = new_child
        elif old_child == right:
            self.children[parent - self.L, 1] = new_child
        else:
            raise ValueError(f"replace_child: {old_child} is not a child of {parent}")
        self.mark_edge_updated(new_child)

    def load_parts(self, parents, heights):
        """DatedTree::load_parts (tree.rs:1135-1151, load_parents 906-930):
        parents [N] (0xFFFFFFFF for the root) and heights [N]; children are
        rebuilt in node order.  b3 leaves the dirty marks alone here; this
        mirror marks every edge so that the next evaluation sees the new
        tree (as the verbatim load path does, tree.rs:997-1009)."""
        parents = np.asarray(parents).astype(np.int64)
        heights = np.asarray(heights, np.float64)
        if parents.size != self.N or heights.size != self.N:
            raise ValueError("parents and heights must have one entry per node")
        parents[parents == TOMBSTONE] = -1
        children = np.full((self.L - 1, 2), -1, np.int64)
        root = None
        for node in range(self.N):
            par = int(parents[node])
            if par < 0:
                root = node
                continue
            row = children[par - self.L]
            row[0 if row[0] < 0 else 1] = node
        if root is None or (children < 0).any():
            raise ValueError("parents do not describe a binary tree")
        self.parents = parents
        self.children = children
        self.heights = heights.copy()
        self.root = int(root)
        self.mark_all_edges_updated()

    def total_length(self) -> float:
        """tree.rs:580-582: edge lengths summed in edge (node) order."""
        total = 0.0
        for e in range(self.N):
            if e != self.root:
                total += float(self.heights[self.parents[e]] - self.heights[e])
        return total

    def set_root(self, node):
        self.root = int(node)
        self.parents[node] = -1

    # ---- Parameter protocol (tree.rs:975-995) ----------------------------
    def _save(self):
        self._bak = (self.parents.copy(), self.children.copy(), self.heights.copy(), self.root)

    def _clear(self):
        self._upd_edges[:] = False
        self._upd_partials[:] = False
        self._upd_props[:] = False

    def accept(self):
        self._save()
        self._clear()

    def reject(self):
        p, c, h, r = self._bak
        self.parents, self.children, self.heights, self.root = p.copy(), c.copy(), h.copy(), r
        self._clear()

    def copy(self):
        t = DatedTree.__new__(DatedTree)
        t.L, t.N = self.L, self.N
        t.parents = self.parents.copy()
        t.children = self.children.copy()
        t.heights = self.heights.copy()
        t.root = self.root
        t._upd_edges = self._upd_edges.copy()
        t._upd_partials = self._upd_partials.copy()
        t._upd_props = self._upd_props.copy()
        t._save()
        return t

    def parents_u32(self):
        p = self.parents.copy()
        p[p < 0] = TOMBSTONE
        return p.astype(np.uint32)
