"""Parity at the full sizes of the measured workloads, against the FP64 oracle.

- GTR+Gamma4, 128 taxa, 100,000 unique simulated patterns with weights
  U{1..5}, seed 4 (the benchmark workload), and the delayed-acceptance
  benchmark's 200 taxa x 200,000 sites: a full evaluation and one-edge
  partial evaluations after node-height moves, per site and in total.
- Batched HKY+Gamma4: 256 instances of 50-500 taxa and 2,000-50,000
  patterns, evaluated in one launch (single slot, as benchmarked); every
  instance's total against the plain FP64 GPU path, and a sample of
  instances (totals and per-site values) against the oracle itself.
- GY94 codons on the matrix cores: 64 taxa, 20,000 unique simulated codon
  patterns, Gamma4 (the codon benchmark workload), matrices built on the
  device; per site, per category and in total against the oracle, then
  one-edge partial evaluations after node-height moves, bitwise against a
  fresh full evaluation of the moved tree.

Tolerance (FP32 partials, FP64 root and sums): total relative <= 1e-6, per
site |d lnL| <= 2e-6 max(1, |lnL|).  Codons (FP16 matrices and partials):
per site and category |d lnL| <= 3 (L - 1) 2^-11, the bound derived in
test_gpu_codon.py, and total relative <= 1e-6.
"""
import os

import numpy as np
import pytest

from aspartik_amd import workloads
from aspartik_amd.b3 import Batch, Calculator, DatedTree
from oracle import oracle as O

pytestmark = pytest.mark.gpu

F32_SITE = 2e-6
F32_TOTAL = 1e-6


def _threads(K, P):
    ncores = len(os.sched_getaffinity(0))
    return max(1, min(ncores // K, (P + 999) // 1000))


def _oracle_mixed(olik, K):
    site = np.array([olik.site(c) for c in range(K)])
    mx = site.max(0)
    return mx + np.log(np.exp(site - mx).sum(0)) + np.log(1.0 / K)


def _check_sites(got, ref):
    err = np.abs(got - ref) / np.maximum(1.0, np.abs(ref))
    assert err.max() <= F32_SITE, f"max per-site error {err.max():.3e}"


def test_gtr_gamma4_128taxa_100000pat(hip_device):
    _gtr_gamma4_full_shape(workloads.gtr_gamma4(128, 100_000, seed=4), 128, 100_000)


def test_gtr_gamma4_200taxa_200000sites_mcmc_workload(hip_device):
    """The delayed-acceptance benchmark's likelihood: 200 taxa, 200,000
    simulated sites (weights 1)."""
    _gtr_gamma4_full_shape(workloads.gtr_gamma4_mcmc(200, 200_000), 200, 200_000)


def _gtr_gamma4_full_shape(wl, L, P):
    K = wl["K"]
    print(f"[full-shape] GTR+Gamma4 {L} taxa {P} patterns", flush=True)
    tree = DatedTree(L, wl["parents"], wl["children"], wl["heights"])
    model = wl["model"]
    calc = Calculator.HIP().make(wl["codes"], P, K)
    calc.set_weights(wl["weights"])
    calc.set_model(model.q_matrix(), model.frequencies(), True)
    full = workloads.full_lists(tree)
    calc.propose_times(*full, workloads.edge_times(tree, full[1]), wl["rates"])
    total = calc.total()
    mixed = calc.site_mixed()
    calc.accept()

    otree = O.Tree(L, wl["parents"], wl["children"], wl["heights"])
    om = O.Model.gtr(model.rates, model.freqs)
    olik = O.GammaLikelihood(otree, om, wl["codes"], P, wl["weights"], num_categories=K,
                             alpha=wl["alpha"], threads_per_category=_threads(K, P),
                             parallel_categories=True)
    ref = olik.likelihood()
    olik.accept()
    otree.accept()
    assert abs(total - ref) <= F32_TOTAL * abs(ref), (total, ref)
    _check_sites(mixed, _oracle_mixed(olik, K))

    # one-edge partial evaluations: node-height moves, each accepted, against
    # the oracle moved the same way (its own partial update)
    rng = np.random.default_rng(8)
    for step in range(4):
        node = int(rng.integers(L, 2 * L - 2))  # not the root
        kids = tree.children_of(node)
        lo = max(tree.heights[kids[0]], tree.heights[kids[1]])
        hi = tree.heights[tree.parent_of(node)]
        h = lo + (hi - lo) * float(rng.uniform(0.05, 0.95))
        tree.set_height(node, h)
        otree.set_height(node, h)
        internals, children = tree.partials_lists()
        calc.propose_times(internals, children, workloads.edge_times(tree, children),
                           wl["rates"])
        total = calc.total()
        ref = olik.likelihood()
        assert abs(total - ref) <= F32_TOTAL * abs(ref), (step, total, ref)
        _check_sites(calc.site_mixed(), _oracle_mixed(olik, K))
        calc.accept()
        tree.accept()
        olik.accept()
        otree.accept()
    calc.close()


def _f64_plain_total(x, K, rates, up):
    """One instance on the plain FP64 GPU path (oracle-pinned to 1e-12 with
    device-built P(t) by the parity tests)."""
    calc = Calculator.HIPPlainF64().make(x["codes"], x["P"], K)
    try:
        calc.set_weights(np.ones(x["P"], np.uint32))
        calc.set_model(x["model"].q_matrix(), x["model"].frequencies(), True)
        calc.propose_times(*up)
        return calc.total(), [calc.likelihoods(k) for k in range(K)]
    finally:
        calc.close()


def test_batched_hky_gamma4_256_instances(hip_device):
    """Every instance's total against the plain FP64 GPU evaluation of the
    same instance, and a sample of instances, per site, against the oracle
    itself (the oracle alone would take minutes over all 7e9 node-pattern-
    category units of the batch)."""
    wl = workloads.hky_gamma4_batch(count=256)
    insts = wl["instances"]
    K, rates = wl["K"], wl["rates"]
    print(f"[full-shape] batched HKY+Gamma4, {len(insts)} instances", flush=True)
    ups = []
    for x in insts:
        t = DatedTree(x["L"], x["parents"], x["children"], x["heights"])
        internals, children = workloads.full_lists(t)
        ups.append((internals, children, workloads.edge_times(t, children), rates))
    b = Batch([(x["codes"], None) for x in insts], num_categories=K, single_slot=True)
    for i, x in enumerate(insts):
        b.set_model(i, x["model"].q_matrix(), x["model"].frequencies(), True)
    b.propose_times(ups, graph=True)
    totals = b.totals()
    b.accept()
    sample = set(np.random.default_rng(3).choice(len(insts), size=4, replace=False).tolist())
    site_got = {i: np.array([b.likelihoods(i, k) for k in range(K)]) for i in sample}
    b.close()
    worst = 0.0
    for i, x in enumerate(insts):
        ref, ref_sites = _f64_plain_total(x, K, rates, ups[i])
        rel = abs(totals[i] - ref) / abs(ref)
        worst = max(worst, rel)
        assert rel <= F32_TOTAL, (i, totals[i], ref)
    print(f"[full-shape] worst instance total rel err vs FP64 GPU {worst:.2e}", flush=True)
    for i in sorted(sample):
        x = insts[i]
        ot = O.Tree(x["L"], x["parents"], x["children"], x["heights"])
        om = O.Model.hky(x["kappa"], np.full(4, 0.25))
        lik = O.GammaLikelihood(ot, om, x["codes"], x["P"], np.ones(x["P"], np.uint32),
                                num_categories=K, alpha=wl["alpha"],
                                threads_per_category=_threads(K, x["P"]),
                                parallel_categories=True)
        ref = lik.likelihood()
        assert abs(totals[i] - ref) <= F32_TOTAL * abs(ref), (i, totals[i], ref)
        f64_total, _ = _f64_plain_total(x, K, rates, ups[i])
        assert abs(f64_total - ref) <= 1e-12 * abs(ref), (i, f64_total, ref)
        for k in range(K):
            _check_sites(site_got[i][k], lik.site(k))
        print(f"[full-shape] instance {i} ({x['L']} taxa, {x['P']} patterns) vs oracle: "
              f"{abs(totals[i] - ref) / abs(ref):.2e}", flush=True)
        del lik, ot


def test_gy94_codon_64taxa_20000pat(hip_device):
    wl = workloads.gy94_codon(64, 20_000)
    L, P, K = 64, 20_000, wl["K"]
    print(f"[full-shape] GY94 codons {L} taxa {P} patterns", flush=True)
    model, rates = wl["model"], wl["rates"]
    pi = model.frequencies()
    tree = DatedTree(L, wl["parents"], wl["children"], wl["heights"])
    calc = Calculator.HIPCodon(61).make(wl["codes"], P, K)
    calc.set_model(model.q_matrix(), pi)
    full = workloads.full_lists(tree)
    calc.propose_times(*full, workloads.edge_times(tree, full[1]), rates)
    calc.accept()
    got = np.array([calc.likelihoods(k) for k in range(K)])
    total, mixed = calc.total(), calc.site_mixed()

    otree = O.Tree(L, wl["parents"], wl["children"], wl["heights"])
    om = O.Model.reversible(model.exchangeability(), pi)
    flat = np.asarray(full[1]).reshape(-1)
    ncores = len(os.sched_getaffinity(0))
    ref = []
    for c in range(K):
        mats = np.stack([om.transition(otree.edge_length(int(m)) * rates[c]) for m in flat])
        cn = O.CalcN(wl["codes"], P, 61, 30, ncores)
        cn.propose(full[0], full[1], mats, pi)
        ref.append(cn.likelihoods())
    ref = np.array(ref)
    mx = ref.max(0)
    ref_mixed = mx + np.log(np.exp(ref - mx).sum(0)) + np.log(1.0 / K)
    bound = 3 * (L - 1) * 2.0 ** -11
    err = np.abs(got - ref)
    print(f"  per-category site max {err.max():.3e} (bound {bound:.3e}); total rel "
          f"{abs(total - ref_mixed.sum()) / abs(ref_mixed.sum()):.2e}", flush=True)
    assert err.max() <= bound
    assert np.abs(mixed - ref_mixed).max() <= bound
    assert abs(total - ref_mixed.sum()) <= 1e-6 * abs(ref_mixed.sum())

    # one-edge partial evaluations, each accepted: bitwise equal to a fresh
    # full evaluation of the moved tree (every node computes the same FP16
    # values whether its children come from registers or from memory)
    fresh = Calculator.HIPCodon(61).make(wl["codes"], P, K)
    fresh.set_model(model.q_matrix(), pi)
    rng = np.random.default_rng(9)
    try:
        for step in range(3):
            node = int(rng.integers(L, 2 * L - 2))  # not the root
            kids = tree.children_of(node)
            lo = max(tree.heights[kids[0]], tree.heights[kids[1]])
            hi = tree.heights[tree.parent_of(node)]
            tree.set_height(node, lo + (hi - lo) * float(rng.uniform(0.05, 0.95)))
            internals, children = tree.partials_lists()
            calc.propose_times(internals, children, workloads.edge_times(tree, children),
                               rates)
            calc.accept()
            tree.accept()
            whole = workloads.full_lists(tree)
            fresh.propose_times(*whole, workloads.edge_times(tree, whole[1]), rates)
            fresh.accept()
            for k in range(K):
                np.testing.assert_array_equal(calc.likelihoods(k), fresh.likelihoods(k),
                                              err_msg=f"step {step} category {k}")
            assert calc.total() == fresh.total()
    finally:
        fresh.close()
        calc.close()
