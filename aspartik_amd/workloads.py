"""Seeded synthetic workloads of the measured configurations.

Each builder returns a dict with the tree arrays (b3 numbering), the
substitution model object, IUPAC codes [L, P], pattern weights, and the
discrete-gamma settings.  Unspecified choices are stated here: Yule birth
rate 1 with branch lengths rescaled to mean 0.1; seed 4 by default.
"""
from __future__ import annotations

import numpy as np

from . import sim
from .b3.calculator import gamma_rates
from .b3.substitution import GTR, GY94, HKY


def gtr_gamma4(num_leaves=128, num_patterns=100_000, alpha=0.5, seed=4, weights=(1, 5)):
    """GTR+Gamma4: GTR rates ~ Dirichlet(1), freqs ~ Dirichlet(10), alpha 0.5,
    ``num_patterns`` unique patterns simulated on a Yule tree, integer
    weights uniform on [1, 5]."""
    rng = np.random.default_rng(seed)
    parents, children, heights = sim.yule_tree(num_leaves, rng, 1.0, 0.1)
    freqs = rng.dirichlet([10.0] * 4)
    rates6 = rng.dirichlet([1.0] * 6)
    model = GTR(freqs, rates6)
    K = 4
    cat = gamma_rates(alpha, K)
    states = sim.simulate_unique_patterns(parents, children, heights,
                                          sim.expm_transition(model.q_matrix()), freqs, cat,
                                          num_patterns, rng)
    codes = sim.dna_codes(states)
    lo, hi = weights
    w = rng.integers(lo, hi + 1, size=num_patterns).astype(np.uint32)
    return dict(name=f"gtr_gamma4_{num_leaves}taxa_{num_patterns}pat", parents=parents,
                children=children, heights=heights, model=model, codes=codes, weights=w,
                K=K, alpha=alpha, rates=cat, seed=seed)


def hky_gamma4(num_leaves, num_patterns, seed, alpha=0.5):
    """HKY+Gamma4 instance: kappa ~ LogNormal(1, 0.5), freqs 0.25, weights 1."""
    rng = np.random.default_rng(seed)
    parents, children, heights = sim.yule_tree(num_leaves, rng, 1.0, 0.1)
    kappa = float(rng.lognormal(1.0, 0.5))
    freqs = np.full(4, 0.25)
    model = HKY(freqs, kappa)
    K = 4
    cat = gamma_rates(alpha, K)
    states = sim.simulate_unique_patterns(parents, children, heights,
                                          sim.expm_transition(model.q_matrix()), freqs, cat,
                                          num_patterns, rng)
    return dict(name=f"hky_gamma4_{num_leaves}taxa_{num_patterns}pat", parents=parents,
                children=children, heights=heights, model=model, codes=sim.dna_codes(states),
                weights=np.ones(num_patterns, np.uint32), K=K, alpha=alpha, rates=cat,
                seed=seed)


def full_lists(tree):
    """Update list of a full evaluation (every edge marked)."""
    tree.mark_all_edges_updated()
    internals, children = tree.partials_lists()
    tree.accept()
    return internals, children


def path_lists(tree, edge):
    """Update list after changing one edge: the edge's ancestors to the root."""
    tree.mark_edge_updated(edge)
    internals, children = tree.partials_lists()
    tree.accept()
    return internals, children


def edge_times(tree, children):
    flat = np.asarray(children).reshape(-1).astype(np.int64)
    return tree.heights[tree.parents[flat]] - tree.heights[flat]


def algorithmic_bytes(internals, children, num_leaves, root, K, P, states=4, elem=4):
    """Algorithmic HBM bytes of one evaluation (each child partial read, each
    non-root result written; tip codes counted once across K; plus weights
    and the 8-byte result):
      B = sum_n [sum_child (leaf ? P : K P S b) + (n != root ? K P S b : 0)]
          + 4 P + 8
    """
    part = K * P * states * elem
    B = 0
    for i, n in enumerate(np.asarray(internals)):
        for c in np.asarray(children)[i]:
            B += P if c < num_leaves else part
        if n != root:
            B += part
    return B + 4 * P + 8


def store_bytes(internals, children, num_leaves, K, P, tree_children=None, unit=18,
                reread_in_pass=False, epilogue=True):
    """HBM bytes one evaluation of this update list must move with this
    library's storage (the roofline basis of the pruning kernels):
      - every listed node's partial written once: ``unit`` bytes per pattern
        and category (FP32: a float4 and an int16 exponent, 18 B; codon FP16:
        64 padded states and an exponent, 130 B).  With ``tree_children``
        (the tree's [L-1, 2] child array; the FP32 path) a cherry, both of
        whose children are leaves, is not written;
      - every stored child that the pass does not compute read once (``unit``
        bytes per pattern and category; a leaf: its code byte per pattern; a
        cherry, with ``tree_children``: its two leaves' code bytes, from which
        it is recomputed);
      - with ``reread_in_pass`` (the codon kernel, which has no LDS stack), the
        first-evaluated internal child of a node whose children are both
        internal is read back as well;
      - with ``epilogue``, the pattern weights read (4 B) and the site values
        written (8 B per category plus 8 B mixed) per pattern; without it,
        only the per-category site values.
    A child computed in the same pass is otherwise consumed from registers or
    LDS and not read back.  The few hundred bytes of transition matrices and
    update list are not counted."""
    internals = np.asarray(internals).reshape(-1)
    children = np.asarray(children).reshape(-1, 2)
    L = num_leaves
    listed = set(int(n) for n in internals)

    def cherry(n):
        if tree_children is None:
            return False
        a, b = tree_children[n - L]
        return a < L and b < L

    part = unit * K * P
    B = 0
    leaves_read = set()
    for n, ch in zip(internals, children):
        if not (tree_children is not None and ch[0] < L and ch[1] < L):
            B += part
        both_internal = True
        for c in ch:
            c = int(c)
            if c < L:
                leaves_read.add(c)
                both_internal = False
            elif c not in listed:
                if cherry(c):
                    leaves_read.update(int(x) for x in tree_children[c - L])
                else:
                    B += part
                both_internal = False
        if reread_in_pass and both_internal:
            B += part
    B += len(leaves_read) * P
    if epilogue:
        B += 4 * P + 8 * K * P + 8 * P
    else:
        B += 8 * K * P
    return B


def algorithmic_flops(internals, children, num_leaves, K, P, states=4):
    """Per (node, pattern, category): sum over internal children of 2 S^2,
    plus S for the product; at the root 2S + 1 more."""
    f = 0
    S = states
    for i in range(len(internals)):
        per = S
        for c in np.asarray(children)[i]:
            if c >= num_leaves:
                per += 2 * S * S
        f += per
    f += 2 * S + 1
    return f * K * P


def hky_gamma4_batch(count=256, taxa=(50, 500), patterns=(2_000, 50_000), seed=4, alpha=0.5):
    """Batched HKY+Gamma4 workload: ``count`` instances, each with its own
    seeded Yule tree of U{taxa} leaves and U{patterns} sites simulated on it
    (multithreaded host simulator), kappa ~ LogNormal(1, 0.5), equal base
    frequencies, alpha 0.5, weights 1.  With at least 50 taxa, duplicate
    simulated columns are vanishingly rare, so the sites are used as the
    instance's patterns without compression."""
    rng = np.random.default_rng(seed)
    cat = gamma_rates(alpha, 4)
    out = []
    for i in range(count):
        L = int(rng.integers(taxa[0], taxa[1] + 1))
        P = int(rng.integers(patterns[0], patterns[1] + 1))
        parents, children, heights = sim.yule_tree(L, rng, 1.0, 0.1)
        kappa = float(rng.lognormal(1.0, 0.5))
        freqs = np.full(4, 0.25)
        model = HKY(freqs, kappa)
        states = sim.simulate_states_fast(parents, children, heights, model.q_matrix(), freqs,
                                          cat, P, int(rng.integers(0, 2**62)))
        out.append(dict(parents=parents, children=children, heights=heights, kappa=kappa,
                        model=model, codes=sim.dna_codes(states), L=L, P=P))
    return dict(name=f"hky_gamma4_batch{count}", instances=out, K=4, alpha=alpha, rates=cat,
                seed=seed)


def gy94_codon(num_leaves=64, num_patterns=20_000, omega=0.3, kappa=2.0, alpha=0.5, seed=4):
    """GY94 codon workload: omega 0.3, kappa 2, F3x4 frequencies from three
    Dirichlet(10) draws, Gamma4 alpha 0.5, ``num_patterns`` unique codon
    patterns simulated on a seeded Yule tree (multithreaded host simulator,
    first-occurrence order), weights 1.  States are the 61 sense codons in
    GY94's order."""
    rng = np.random.default_rng(seed)
    parents, children, heights = sim.yule_tree(num_leaves, rng, 1.0, 0.1)
    model = GY94(omega, kappa, np.stack([rng.dirichlet([10.0] * 4) for _ in range(3)]))
    K = 4
    cat = gamma_rates(alpha, K)
    pi = model.frequencies()
    q = model.q_matrix()
    seen, parts, have = set(), [], 0
    batch = int(num_patterns * 1.05) + 64
    for _ in range(100):
        st = sim.simulate_states_fast(parents, children, heights, q, pi, cat, batch,
                                      int(rng.integers(0, 2**62)))
        keep = sim.unique_columns(st, seen)[: num_patterns - have]
        parts.append(st[:, keep])
        have += keep.size
        if have >= num_patterns:
            break
    codes = np.ascontiguousarray(np.concatenate(parts, axis=1)[:, :num_patterns])
    return dict(name=f"gy94_{num_leaves}taxa_{num_patterns}pat", parents=parents,
                children=children, heights=heights, model=model, codes=codes,
                weights=np.ones(num_patterns, np.uint32), K=K, alpha=alpha, rates=cat,
                seed=seed)


def hky_cpu_mcmc(num_leaves=16, num_sites=1_000, kappa=2.0, seed=4):
    """The reference's CPU-runnable MCMC case: a seeded Yule tree (birth rate
    1, branch lengths rescaled to mean 0.1), HKY with kappa 2 and equal base
    frequencies, ``num_sites`` columns simulated on the tree and compressed to
    unique patterns with their counts as weights (first-occurrence order)."""
    from .data import deduplicate
    rng = np.random.default_rng(seed)
    parents, children, heights = sim.yule_tree(num_leaves, rng, 1.0, 0.1)
    freqs = np.full(4, 0.25)
    model = HKY(freqs, kappa)
    states, _ = sim.simulate_states(parents, children, heights,
                                    sim.expm_transition(model.q_matrix()), freqs, np.ones(1),
                                    num_sites, rng)
    codes, weights, _ = deduplicate(sim.dna_codes(states))
    return dict(name=f"hky_{num_leaves}taxa_{num_sites}sites_mcmc", parents=parents,
                children=children, heights=heights, model=model, codes=codes,
                weights=weights, K=1, kappa=kappa, freqs=freqs,
                gtr_rates=np.array([1.0, kappa, 1.0, 1.0, kappa, 1.0]), seed=seed)


def gtr_gamma4_mcmc(num_leaves=200, num_sites=200_000, alpha=0.5, seed=4):
    """MCMC workload: GTR rates ~ Dirichlet(1), frequencies ~ Dirichlet(10),
    Gamma4 alpha 0.5, ``num_sites`` sites simulated on a seeded Yule tree
    (multithreaded host simulator).  With 200 taxa duplicate columns are
    vanishingly rare, so the sites are the patterns (weights 1)."""
    rng = np.random.default_rng(seed)
    parents, children, heights = sim.yule_tree(num_leaves, rng, 1.0, 0.1)
    freqs = rng.dirichlet([10.0] * 4)
    rates6 = rng.dirichlet([1.0] * 6)
    model = GTR(freqs, rates6)
    K = 4
    cat = gamma_rates(alpha, K)
    states = sim.simulate_states_fast(parents, children, heights, model.q_matrix(), freqs, cat,
                                      num_sites, int(rng.integers(0, 2**62)))
    return dict(name=f"gtr_gamma4_{num_leaves}taxa_{num_sites}pat_mcmc", parents=parents,
                children=children, heights=heights, model=model, codes=sim.dna_codes(states),
                weights=np.ones(num_sites, np.uint32), K=K, alpha=alpha, rates=cat,
                gtr_rates=rates6, freqs=freqs, seed=seed)
