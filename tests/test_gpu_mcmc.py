"""MCMC driver, MLP surrogate and delayed acceptance on the GPU.

* The surrogate's forward pass and gradients (HIP kernels, BF16 weights,
  FP32 accumulation) against a float64 PyTorch-CPU model holding the same
  BF16-rounded weights; the tolerance (2e-3 relative, in norm) covers FP32
  accumulation over <= 404 terms.
* The chain's ln L after many accepted and rejected moves (partial updates,
  model restores, topology moves) against the FP64 oracle for the chain's
  current state, and its log prior against the formula.
* Delayed acceptance keeps the posterior: posterior means of a plain chain
  and of delayed-acceptance chains (trained and untrained surrogate;
  topology moves straight to the exact test or screened by the surrogate,
  first stage tempered)
  agree within 5 Monte Carlo standard errors (ESS-based).
"""
import math

import numpy as np
import pytest

from aspartik_amd import sim
from aspartik_amd.b3 import GTR, Chain, Surrogate, ess, gamma_rates
from oracle import oracle as O
from tests.helpers import random_tree

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _bf16(a):
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def _torch_ref(s, p, x, t):
    W1, b1, W2, b2, w3, b3 = s.unpack(p)
    T = lambda a: torch.tensor(np.asarray(a, np.float64), dtype=torch.float64,  # noqa: E731
                               requires_grad=True)
    W1t, b1t, W2t, b2t, w3t, b3t = (T(_bf16(W1)), T(b1), T(_bf16(W2)), T(b2), T(_bf16(w3)), T(b3))
    xt = torch.tensor(np.asarray(x, np.float64))
    gelu = lambda z: torch.nn.functional.gelu(z, approximate="tanh")  # noqa: E731
    y = gelu(gelu(xt @ W1t + b1t) @ W2t + b2t) @ w3t + b3t
    loss = ((y - torch.tensor(np.asarray(t, np.float64))) ** 2).mean()
    loss.backward()
    g = np.concatenate([v.grad.numpy().reshape(-1) for v in (W1t, b1t, W2t, b2t, w3t, b3t)])
    return y.detach().numpy(), float(loss), g


@pytest.mark.parametrize("D,H,B", [(404, 256, 64), (37, 64, 5), (404, 256, 1)])
def test_surrogate_forward_and_gradients(hip_device, D, H, B):
    s = Surrogate(D, H, seed=D + B)
    rng = np.random.default_rng(B)
    p = s.params()
    p[s.num_params - H - 1:] = rng.normal(size=H + 1).astype(np.float32) * 0.3  # nonzero output
    p[D * H:D * H + H] = rng.normal(size=H).astype(np.float32) * 0.1            # and biases
    s.set_params(p)
    x = rng.normal(size=(B, D)).astype(np.float32)
    t = rng.normal(size=B).astype(np.float32)
    y_ref, loss_ref, g_ref = _torch_ref(s, s.params(), x, t)
    y = s.forward(x)
    np.testing.assert_allclose(y, y_ref, rtol=2e-3, atol=2e-3 * np.abs(y_ref).max())
    loss, g = s.loss_grad(x, t)
    assert abs(loss - loss_ref) <= 2e-3 * loss_ref
    assert np.linalg.norm(g - g_ref) <= 2e-3 * np.linalg.norm(g_ref)
    # each block separately (a transposed or misplaced block fails here)
    for got, ref in zip(s.unpack(g), s.unpack(g_ref)):
        assert np.linalg.norm(got - ref) <= 5e-3 * np.linalg.norm(ref) + 1e-7


def test_single_sample_path_matches_batched(hip_device):
    """The fused one-sample forward used inside the chain equals the batched
    kernels' result to FP32 rounding (it sums in another order: sixteen
    partial sums per first-layer unit, sixteen per second-layer unit)."""
    s = Surrogate(404, 256, seed=17)
    rng = np.random.default_rng(17)
    p = s.params()
    p[s.num_params - 257:] = rng.normal(size=257).astype(np.float32) * 0.3
    s.set_params(p)
    s.set_scaling(rng.normal(size=404), rng.uniform(0.5, 2.0, size=404), -1234.5, 17.0)
    x = rng.normal(size=(9, 404)).astype(np.float32)
    batched = s.forward(x)
    single = np.array([s.forward(x[i:i + 1])[0] for i in range(9)])
    np.testing.assert_allclose(single, batched, rtol=1e-6, atol=1e-4)
    # after training steps the one-sample path reads the updated weights (its
    # transposed copy is rebuilt), and a width that is not a multiple of 8
    # exercises the padded tail of each row
    s.train(x, rng.normal(size=9).astype(np.float32), epochs=3, batch=9)
    batched = s.forward(x)
    single = np.array([s.forward(x[i:i + 1])[0] for i in range(9)])
    np.testing.assert_allclose(single, batched, rtol=1e-6, atol=1e-4)
    s.close()
    t = Surrogate(37, 45, seed=3)
    t.set_scaling(np.zeros(37), np.ones(37), 0.0, 1.0)
    z = rng.normal(size=(5, 37)).astype(np.float32)
    np.testing.assert_allclose([t.forward(z[i:i + 1])[0] for i in range(5)], t.forward(z),
                               rtol=1e-6, atol=1e-5)
    t.close()


def test_surrogate_gradient_is_a_derivative(hip_device):
    """Finite differences of the reference loss pin the reference gradient."""
    s = Surrogate(11, 16, seed=3)
    rng = np.random.default_rng(3)
    x = rng.normal(size=(7, 11)).astype(np.float32)
    t = rng.normal(size=7).astype(np.float32)
    p = s.params().astype(np.float64)
    _, _, g = _torch_ref(s, p, x, t)
    # biases are used unrounded, so the loss is differentiable in them
    b1, b2 = 11 * 16, 11 * 16 + 16 + 16 * 16
    for idx in [b1, b1 + 9, b2 + 2, b2 + 15, len(p) - 1]:
        e = np.zeros_like(p)
        e[idx] = 1e-4
        lp = _torch_ref(s, p + e, x, t)[1]
        lm = _torch_ref(s, p - e, x, t)[1]
        assert abs((lp - lm) / 2e-4 - g[idx]) < 1e-5 + 1e-3 * abs(g[idx])


def test_surrogate_learns_a_smooth_function(hip_device):
    rng = np.random.default_rng(5)
    D, n = 16, 12288
    a = rng.normal(size=D)
    X = rng.normal(size=(n + 2048, D)).astype(np.float32)
    y = (X @ a - 0.5 * (X ** 2).sum(1) + 3.0).astype(np.float32)
    s = Surrogate(D, 128, seed=9)
    s.fit_scaling(X[:n], y[:n])
    before = np.mean((s.forward(X[n:]) - y[n:]) ** 2)
    loss = s.train(X[:n], y[:n], epochs=60, batch=256, lr=2e-3)
    after = np.mean((s.forward(X[n:]) - y[n:]) ** 2)
    r2 = 1 - after / np.var(y[n:])
    assert loss < 0.05 and after < 0.05 * before and r2 > 0.95, (loss, before, after, r2)


def _data(L, sites, seed, alpha=0.5):
    rng = np.random.default_rng(seed)
    par, ch, h = random_tree(L, rng, mean_branch=0.1)
    freqs = rng.dirichlet([10] * 4)
    rates = rng.dirichlet([2] * 6)
    model = GTR(freqs, rates)
    codes = sim.dna_codes(sim.simulate_states_fast(par, ch, h, model.q_matrix(), freqs,
                                                   gamma_rates(alpha, 4), sites,
                                                   int(rng.integers(0, 2**62))))
    return dict(codes=codes, parents=par, children=ch, heights=h, rates=rates, freqs=freqs,
                alpha=alpha, L=L)


def _oracle_loglik(d, chain):
    par, ch, h, r, a = chain.params()
    t = O.Tree(d["L"], par, ch, h)
    lik = O.GammaLikelihood(t, O.Model.gtr(r, d["freqs"]), d["codes"], d["codes"].shape[1],
                            np.ones(d["codes"].shape[1], np.uint32), num_categories=4, alpha=a)
    return lik.likelihood(), par, h, a


def test_chain_full_evaluations_hky_one_category(hip_device):
    """The reference's CPU-runnable case in small: HKY (kappa 2, equal
    frequencies) as fixed exchangeabilities, one category, the substitution
    model never moved, every node recomputed at each step
    (full_evaluations); after a few hundred steps the chain's ln L equals the
    oracle's FP64 value of its state, with the pattern weights of a
    compressed alignment."""
    from aspartik_amd import workloads
    wl = workloads.hky_cpu_mcmc(num_leaves=10, num_sites=400, seed=5)
    L, P = wl["codes"].shape
    c = Chain(wl["codes"], wl["parents"], wl["children"], wl["heights"], wl["gtr_rates"],
              wl["freqs"], 1.0, weights=wl["weights"], categories=1, w_rates=0.0, w_alpha=0.0,
              full_evaluations=1, seed=9)
    for rep in range(3):
        _, stats, _ = c.run(200)
        assert stats["op_proposed"][2] == 0 and stats["op_proposed"][3] == 0
        ll = c.state()[0]
        par, ch, h, r, _ = c.params()
        np.testing.assert_allclose(r / r.sum(), wl["gtr_rates"] / wl["gtr_rates"].sum())
        t = O.Tree(L, par, ch, h)
        ref = O.GammaLikelihood(t, O.Model.hky(wl["kappa"], wl["freqs"]), wl["codes"], P,
                                wl["weights"], num_categories=1).likelihood()
        assert abs(ll - ref) <= 1e-6 * abs(ref), (rep, ll, ref)
    c.close()


def test_chain_loglik_matches_oracle_after_many_moves(hip_device):
    d = _data(12, 600, 1)
    c = Chain(d["codes"], d["parents"], d["children"], d["heights"], d["rates"], d["freqs"],
              d["alpha"], birth_rate=2.0, seed=11)
    for rep in range(4):
        trace, stats, _ = c.run(150)
        ll, lp, post = c.state()
        ref, par, h, a = _oracle_loglik(d, c)
        assert abs(ll - ref) <= 1e-6 * abs(ref), (rep, ll, ref)
        # log prior: Yule (n - 1) ln lambda - lambda * tree length + Exp(1) on alpha
        nonroot = par != 0xFFFFFFFF
        tl = float((h[par[nonroot].astype(np.int64)] - h[nonroot]).sum())
        assert abs(lp - ((d["L"] - 1) * math.log(2.0) - 2.0 * tl - a)) < 1e-9
        assert abs(post - ll - lp) < 1e-9
        assert stats["accepted"] > 0 and stats["exact_evals"] > 0
        assert trace[-1, 2] == ll
    assert all(n > 0 for n in stats["op_proposed"])


def test_narrow_exchange_topology_frequencies_with_no_data(hip_device):
    """NarrowExchange on 4 taxa with every tip unknown (ln L = 0): the prior
    gives every labelled ranked history the same mass (the tree length is
    the sum of the internal heights plus the root's whatever the topology),
    so a balanced tree has probability 6 / 18 = 1/3.  narrow_reversible=1
    (a grandparent has at least one internal child) gets it within 5
    standard errors; b3's rule (two internal children, tree.rs:767-770)
    turns the balanced start into a caterpillar that never returns, as the
    CPU chain of the oracle does."""
    parents = np.array([4, 4, 5, 5, 6, 6, 0xFFFFFFFF], np.uint32)
    children = np.array([[0, 1], [2, 3], [4, 5]], np.uint32)
    heights = np.array([0.0, 0.0, 0.0, 0.0, 0.2, 0.3, 0.5])
    codes = np.full((4, 2), 15, np.uint8)
    out = {}
    for loose in (0, 1):
        c = Chain(codes, parents, children, heights, np.ones(6), np.full(4, 0.25), 1.0,
                  categories=1, w_rates=0.0, w_alpha=0.0, birth_rate=2.0, seed=7,
                  narrow_reversible=loose)
        flags = []
        for _ in range(12_000):
            c.run(1)
            par, ch, _, _, _ = c.params()
            root = int(np.flatnonzero(par == 0xFFFFFFFF)[0])
            flags.append(float(all(int(v) >= 4 for v in ch[root - 4])))
        c.close()
        out[loose] = np.array(flags[500:])
    assert out[0][-5000:].sum() == 0.0
    m, se = _mean_and_se(out[1])
    assert abs(m - 1.0 / 3.0) < 5 * se, (m, se)


def _mean_and_se(x):
    e = max(ess(x), 2.0)
    return x.mean(), x.std() / math.sqrt(e)


def test_delayed_acceptance_keeps_the_posterior(hip_device):
    d = _data(8, 400, 2)
    kw = dict(birth_rate=2.0, w_narrow=1.0)
    burn = Chain(d["codes"], d["parents"], d["children"], d["heights"], d["rates"], d["freqs"],
                 d["alpha"], seed=5, **kw)
    _, _, (X, y) = burn.run(3000, collect=3000)
    par, ch, h, r, a = burn.params()
    s = Surrogate(burn.num_features, 128, seed=1)
    s.fit_scaling(X, y)
    s.train(X, y, epochs=200, batch=256, lr=1e-3)
    untrained = Surrogate(burn.num_features, 128, seed=2)
    untrained.fit_scaling(X, y)
    cols = {"tree_length": 5, "alpha": 6, "root_height": 4}
    runs = {}
    # "screened": topology moves through the first stage, tempered at beta 0.5
    for name, sur, mode, seed, screen, beta in [("plain", None, "plain", 21, 0, 1.0),
                                                ("delayed", s, "delayed", 22, 0, 1.0),
                                                ("untrained", untrained, "delayed", 23, 0, 1.0),
                                                ("screened", s, "delayed", 24, 1, 0.5)]:
        c = Chain(d["codes"], par, ch, h, r, d["freqs"], a, seed=seed, screen_topology=screen,
                  surrogate_beta=beta, **kw)
        if sur is not None:
            c.set_surrogate(sur)
        trace, stats, _ = c.run(20000, mode=mode)
        runs[name] = (trace, stats)
    assert runs["delayed"][1]["stage1_rejects"] > 0
    assert runs["delayed"][1]["exact_evals"] < runs["plain"][1]["exact_evals"]
    # with screen_topology every proposal that is neither aborted nor
    # prior-rejected is screened (one surrogate call each, plus one for the
    # starting state); without it NarrowExchange skips the surrogate
    st = runs["screened"][1]
    assert st["surrogate_evals"] == st["steps"] - st["aborts"] - st["prior_rejects"] + 1
    sd = runs["delayed"][1]
    assert sd["surrogate_evals"] < sd["steps"] - sd["aborts"] - sd["prior_rejects"]
    for col_name, col in cols.items():
        m0, s0 = _mean_and_se(runs["plain"][0][:, col])
        for other in ("delayed", "untrained", "screened"):
            m1, s1 = _mean_and_se(runs[other][0][:, col])
            z = abs(m0 - m1) / math.hypot(s0, s1)
            assert z < 5.0, (col_name, other, m0, m1, s0, s1)


def test_trained_surrogate_gains_over_untrained_and_plain(hip_device):
    """What the trained network gains.  Delayed-acceptance chains at the same
    seeds, every move screened (so every exact evaluation follows a passed
    first stage), one with the trained surrogate and one with an untrained
    network of the same shape and input scaling; plain MH beside them.  The
    trained first stage must pass moves the exact test then accepts far more
    often (stage-2 acceptance), so each accepted move costs fewer exact
    evaluations than with the untrained network and than plain MH.  The
    trained side pools three networks trained from different initial
    weights: on this 8-taxon chain one network's exact evaluations per
    accepted move range over 1.8-2.7, while the pooled figure is steady.
    b3 has no counterpart (it has no surrogate)."""
    d = _data(8, 400, 3)
    kw = dict(birth_rate=2.0, w_narrow=1.0)
    burn = Chain(d["codes"], d["parents"], d["children"], d["heights"], d["rates"], d["freqs"],
                 d["alpha"], seed=7, **kw)
    _, _, (X, y) = burn.run(4000, collect=4000)
    par, ch, h, r, a = burn.params()
    nf = burn.num_features
    burn.close()
    nets = []
    for seed in (1, 11, 21):
        net = Surrogate(nf, 128, seed=seed)
        net.fit_scaling(X, y)
        net.train(X, y, epochs=200, batch=256, lr=1e-3)
        nets.append(("trained", net))
    untrained = Surrogate(nf, 128, seed=2)
    untrained.fit_scaling(X, y)
    out = {}
    for name, sur in nets + [("untrained", untrained), ("plain", None)]:
        c = Chain(d["codes"], par, ch, h, r, d["freqs"], a, seed=31, screen_topology=1, **kw)
        if sur is not None:
            c.set_surrogate(sur)
        _, st, _ = c.run(10000, mode="plain" if sur is None else "delayed")
        c.close()
        if sur is not None:
            sur.close()
        acc = out.setdefault(name, {"accepted": 0, "exact_evals": 0, "runs": 0})
        acc["accepted"] += st["accepted"]
        acc["exact_evals"] += st["exact_evals"]
        acc["runs"] += 1
    # per chain of 10,000 steps
    for v in out.values():
        v["accepted"] /= v["runs"]
        v["exact_evals"] /= v["runs"]
    t, u, p = out["trained"], out["untrained"], out["plain"]
    stage2 = {k: out[k]["accepted"] / out[k]["exact_evals"] for k in ("trained", "untrained")}
    per_accept = {k: out[k]["exact_evals"] / out[k]["accepted"] for k in out}
    print(f"[gain] stage-2 acceptance {stage2}; exact evaluations per accepted move {per_accept}; "
          f"exact evaluations {dict((k, out[k]['exact_evals']) for k in out)}", flush=True)
    # measured at these seeds: stage 2 accepts 0.56 of the trained network's
    # moves and 0.13 of the untrained one's; an accepted move costs 1.79
    # exact evaluations with the trained network, 7.5 untrained, 2.55 plain
    assert t["accepted"] > 100 and u["accepted"] > 20 and p["accepted"] > 100
    assert stage2["trained"] > 2.0 * stage2["untrained"], stage2
    assert per_accept["trained"] < 0.5 * per_accept["untrained"], per_accept
    assert per_accept["trained"] < 0.9 * per_accept["plain"], per_accept
    assert t["exact_evals"] < 0.9 * p["exact_evals"], (t["exact_evals"], p["exact_evals"])
