"""GPU deduplication against the host path and the oracle, and the whole
ingest pipeline (FASTA -> MSA -> patterns -> GPU likelihood).

* b3d_deduplicate on the device returns the host path's outputs exactly:
  pattern bytes, weights and each site's pattern, from 1 site up to
  2,000,000 sites, all-identical and all-distinct alignments included.
* The likelihood of the deduplicated alignment (patterns x weights) equals
  the oracle's likelihood of every site with weight 1 (a sum-preserving
  invariant of the compression) to the FP32 path's 1e-6.
"""
import numpy as np
import pytest

from aspartik_amd import sim
from aspartik_amd.b3 import GTR, DatedTree, GammaLikelihood, gamma_rates
from aspartik_amd.data import MSA, deduplicate
from oracle import oracle as O
from tests.helpers import random_tree

pytestmark = pytest.mark.gpu

ASCII = "-ACMGRSVTWYHKDBN"


@pytest.mark.parametrize("S,N,kind", [(1, 1, "random"), (7, 3, "random"), (16, 5000, "dups"),
                                      (64, 20000, "identical"), (8, 30000, "distinct"),
                                      (32, 2_000_000, "dups")])
def test_device_dedup_equals_host(hip_device, S, N, kind):
    rng = np.random.default_rng(S + N)
    if kind == "identical":
        a = np.repeat(rng.choice([0, 1, 2, 4, 8, 15], size=(S, 1)), N, axis=1).astype(np.uint8)
    elif kind == "distinct":
        a = rng.integers(0, 16, size=(S, N)).astype(np.uint8)
    else:
        a = rng.choice([1, 2, 4, 8, 0, 15], size=(S, N), p=[0.3, 0.2, 0.2, 0.2, 0.05, 0.05])
        a = a.astype(np.uint8)
        if kind == "dups" and N > 10:
            src = rng.integers(0, N // 10, size=N // 2)
            a[:, rng.integers(0, N, size=N // 2)] = a[:, src]
    host = deduplicate(a)
    dev = deduplicate(a, device=0)
    for h, d in zip(host, dev):
        np.testing.assert_array_equal(h, d)
    assert host[1].sum() == N
    if N <= 5000:
        ol, ow, os2p = O.deduplicate(a)
        np.testing.assert_array_equal(dev[0], ol)
        np.testing.assert_array_equal(dev[1], ow)


def test_fasta_to_likelihood_pipeline(hip_device, tmp_path):
    rng = np.random.default_rng(11)
    L, sites = 24, 6000
    par, ch, h = random_tree(L, rng, mean_branch=0.05)
    freqs = rng.dirichlet([10] * 4)
    rates = rng.dirichlet([2] * 6)
    model = GTR(freqs, rates)
    states = sim.simulate_states_fast(par, ch, h, model.q_matrix(), freqs, gamma_rates(0.5, 4),
                                      sites, 5)
    codes = sim.dna_codes(states)
    codes[rng.random(codes.shape) < 0.02] = 0      # gaps
    text = "\n".join(f">taxon_{i}\n" + "".join(ASCII[c] for c in row)
                     for i, row in enumerate(codes))
    path = tmp_path / "aln.fasta"
    path.write_text(text + "\n")
    msa = MSA.from_fasta_file(path)
    assert msa.num_sequences == L and msa.num_sites == sites
    leaves, w, s2p = deduplicate(msa, device=0)
    assert leaves.shape[1] < sites and w.sum() == sites
    tree = DatedTree(L, par, ch, h)
    lik = GammaLikelihood(leaves, w, tree, model, num_categories=4, alpha=0.5)
    got = lik.likelihood()
    # every site with weight 1, no compression, in the oracle
    full = np.where(codes == 0, 15, codes).astype(np.uint8)
    otree = O.Tree(L, par, ch, h)
    olik = O.GammaLikelihood(otree, O.Model.gtr(rates, freqs), full, sites,
                             np.ones(sites, np.uint32), num_categories=4, alpha=0.5)
    ref = olik.likelihood()
    assert abs(got - ref) <= 1e-6 * abs(ref)
