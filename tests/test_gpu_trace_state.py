"""Chain state files and traces on the GPU.

* b3m_chain_dump_state / load_state: a chain resumed from a state file in a
  second chain (another seed, another starting tree) continues with exactly
  the same steps -- parameters, generator, operator tuning and the
  likelihood evaluated from scratch all agree -- and the Python reader of
  the verbatim format decodes the native writer's file.
* A traced chain run (Arrow trace of tree, rates, alpha and the chain's ln L)
  replays through replicate_b3: with a fresh FP32 handle the likelihoods
  agree to 1e-12 (same arithmetic), with the FP64 oracle to 1e-7 (the FP32
  path's stated error is 1e-6).
"""
import numpy as np
import pytest

from aspartik_amd import sim
from aspartik_amd.b3 import (GTR, Chain, DatedTree, GammaLikelihood, Real, RealVector,
                             gamma_rates)
from aspartik_amd.b3.trace import TraceWriter, read_trace, replicate_b3
from aspartik_amd.b3.verbatim import load_state
from oracle import oracle as O
from tests.helpers import random_tree

pytestmark = pytest.mark.gpu


def _data(L=14, sites=800, seed=3):
    rng = np.random.default_rng(seed)
    par, ch, h = random_tree(L, rng, mean_branch=0.1)
    freqs = rng.dirichlet([10] * 4)
    rates = rng.dirichlet([2] * 6)
    codes = sim.dna_codes(sim.simulate_states_fast(par, ch, h, GTR(freqs, rates).q_matrix(),
                                                   freqs, gamma_rates(0.5, 4), sites, 17))
    return dict(codes=codes, par=par, ch=ch, h=h, freqs=freqs, rates=rates, L=L)


class _Op:
    def __init__(self, tuning):
        self.accepts, self.rejects, self.tuning = 0, 0, tuning


def test_chain_resume_from_state_is_exact(hip_device):
    d = _data()
    kw = dict(birth_rate=2.0, tune_every=25, tune_cutoff=100000, w_narrow=1.0)
    a = Chain(d["codes"], d["par"], d["ch"], d["h"], d["rates"], d["freqs"], 0.5, seed=7, **kw)
    a.run(400)
    state = a.dump_state()
    par_a, ch_a, h_a, r_a, al_a = a.params()
    trace_a, stats_a, _ = a.run(300)
    # a second chain from another start and seed, then the state
    rng = np.random.default_rng(99)
    par2, ch2, h2 = random_tree(d["L"], rng, mean_branch=0.2)
    b = Chain(d["codes"], par2, ch2, h2, [1, 2, 3, 4, 5, 6], d["freqs"], 1.3, seed=12345, **kw)
    b.load_state(state)
    trace_b, stats_b, _ = b.run(300)
    np.testing.assert_array_equal(trace_a[:, 1:], trace_b[:, 1:])
    assert stats_a["op_accepted"] == stats_b["op_accepted"]
    assert stats_a["tuning"] == stats_b["tuning"]
    pa_, ca, ha, ra, aa = a.params()
    pb, cb, hb, rb, ab = b.params()
    np.testing.assert_array_equal(pa_, pb)
    np.testing.assert_array_equal(ha, hb)
    np.testing.assert_array_equal(ra, rb)
    assert aa == ab
    # the loaded tree's children were rebuilt in node order (tree.rs:913-927)
    np.testing.assert_array_equal(np.sort(ca, axis=1), np.sort(cb, axis=1))
    # the Python reader of the verbatim format decodes the native file
    tree = DatedTree(d["L"], d["par"], d["ch"], d["h"])
    rates, alpha = RealVector.repeat(0.0, 6), Real(0.0)
    ops = [_Op(None), _Op(0.0), _Op(0.0), _Op(0.0), _Op(None)]
    step, post, rng_words = load_state(state, [tree, rates, alpha], ops)
    assert step == 400
    np.testing.assert_array_equal(tree.parents_u32(), par_a)
    np.testing.assert_array_equal(tree.heights, h_a)
    np.testing.assert_array_equal(rates.values, r_a)
    assert alpha.value == al_a
    assert all(0.01 <= o.tuning <= 0.99 for o in ops[1:4])
    assert rng_words[0] != 0 or rng_words[1] != 0


def test_chain_trace_replays(hip_device, tmp_path):
    d = _data(L=16, sites=1200, seed=5)
    c = Chain(d["codes"], d["par"], d["ch"], d["h"], d["rates"], d["freqs"], 0.5, seed=3,
              birth_rate=2.0)
    tree = DatedTree(d["L"], d["par"], d["ch"], d["h"])
    rates, alpha = RealVector(*d["rates"]), Real(0.5)
    path = tmp_path / "chain.trace"
    w = TraceWriter({"tree": tree, "rates": rates, "alpha": alpha}, path, every=25)
    c.run_traced(300, 25, w, tree=tree, rates=rates, alpha=alpha)
    w.finish()
    tab = read_trace(path)
    assert tab.num_rows == 12 and tab.column("step").to_pylist()[-1] == 300
    # replay with a fresh FP32 handle: same arithmetic
    rt = DatedTree(d["L"], d["par"], d["ch"], d["h"])
    sub = GTR(d["freqs"], d["rates"])
    lik = GammaLikelihood(d["codes"], np.ones(d["codes"].shape[1], np.uint32), rt, sub,
                          num_categories=4, alpha=0.5)
    worst = replicate_b3(path, {"tree": rt, "rates": sub.set_rates, "alpha": lik.set_alpha},
                         likelihoods=[lik], tolerance=1e-12)
    assert worst <= 1e-12

    # replay with the FP64 oracle
    class OracleLik:
        def __init__(self):
            self.rates, self.alpha = d["rates"], 0.5

        def set_rates(self, r):
            self.rates = np.asarray(r)

        def set_alpha(self, a):
            self.alpha = float(a)

        def likelihood(self):
            ot = O.Tree(rt.L, rt.parents_u32(), rt.children.astype(np.uint32), rt.heights)
            g = O.GammaLikelihood(ot, O.Model.gtr(self.rates, d["freqs"]), d["codes"],
                                  d["codes"].shape[1], np.ones(d["codes"].shape[1], np.uint32),
                                  num_categories=4, alpha=self.alpha)
            return g.likelihood()

        def reject(self):
            pass

    ol = OracleLik()
    worst = replicate_b3(path, {"tree": rt, "rates": ol.set_rates, "alpha": ol.set_alpha},
                         likelihoods=[ol], tolerance=1e-7)
    assert worst < 1e-7
