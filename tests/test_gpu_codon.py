"""Codon pruning on the MFMA matrix cores (B3H_F16_MFMA) against the FP64
oracle (CalcN: cpu4.rs semantics for S states).

Error statement.  Transition matrices and stored partials are FP16 (unit
roundoff u = 2^-11, normal range down to 2^-14); the products accumulate in
FP32 on the matrix cores; per-node scaling is by exact powers of two and the
root sum, logarithm and mixing are FP64.  All quantities are non-negative,
so relative perturbations of at most u in every P entry and every stored
partial entry move a site likelihood by a factor within (1 +- u)^m, where
m = 2 (L - 1) edges + (L - 1) internal partials rounded to FP16:

    |ln L_gpu - ln L_oracle| <= m u          (normal-range rounding)

Entries below FP16's normal range (tiny multi-substitution probabilities on
short branches) keep only absolute precision 2^-25; their effect is bounded
by the same test on realistic data (sequences simulated on the tree), while
for uniformly random codes, whose likelihoods are dominated by such
entries, the test bounds the error of the rate-mixed site likelihood.
"""
import numpy as np
import pytest

from aspartik_amd import sim
from aspartik_amd.b3 import GY94, Calculator, DatedTree, gamma_rates
from oracle import oracle as O
from tests.helpers import model_matrices, random_tree

pytestmark = pytest.mark.gpu

U16 = 2.0 ** -11


def _bound(L):
    return 3 * (L - 1) * U16


def _gy94(seed, omega=0.3, kappa=2.0):
    rng = np.random.default_rng(seed)
    gy = GY94(omega, kappa, np.stack([rng.dirichlet([10] * 4) for _ in range(3)]))
    pi = gy.frequencies()
    return gy, pi, O.Model.reversible(gy.exchangeability(), pi)


def _full(tree):
    tree.mark_all_edges_updated()
    internals, children = tree.partials_lists()
    tree.accept()
    return internals, children


def _oracle(codes, P, S, internals, children, mats, pi):
    out = []
    for c in range(mats.shape[0]):
        cn = O.CalcN(codes, P, S, 30, 8)
        cn.propose(internals, children, mats[c], pi)
        out.append(cn.likelihoods())
    return np.array(out)


def _mix(site):
    K = site.shape[0]
    m = site.max(0)
    return m + np.log(np.exp(site - m).sum(0)) + np.log(1.0 / K)


def _simulated(L, P, seed, gy, pi, rates, mean_branch=0.1):
    rng = np.random.default_rng(seed)
    par, ch, h = random_tree(L, rng, mean_branch=mean_branch)
    codes = sim.simulate_states_fast(par, ch, h, gy.q_matrix(), pi, rates, P,
                                     int(rng.integers(0, 2**62)))
    return par, ch, h, codes


@pytest.mark.parametrize("L,P", [(12, 700), (40, 1000), (3, 1)])
def test_codon_simulated_data_host_matrices(hip_device, L, P):
    K = 4
    rates = gamma_rates(0.5, K)
    gy, pi, model = _gy94(L)
    par, ch, h, codes = _simulated(L, P, 100 + L, gy, pi, rates)
    tree = O.Tree(L, par, ch, h)
    internals, children = _full(tree)
    mats = model_matrices(model, tree, children, rates)
    ref = _oracle(codes, P, 61, internals, children, mats, pi)
    calc = Calculator.HIPCodon(61).make(codes, P, K)
    calc.propose(internals, children, mats, pi)
    got = np.array([calc.likelihoods(k) for k in range(K)])
    err = np.abs(got - ref)
    print(f"L={L} P={P}: max |dlnL| per category {err.max():.3e} "
          f"mean {err.mean():.3e}; bound {_bound(L):.3e}")
    assert err.max() <= _bound(L)
    mixed = calc.site_mixed()
    np.testing.assert_allclose(mixed, _mix(got), rtol=0, atol=1e-12 * max(1, np.abs(mixed).max()))
    assert abs(calc.total() - mixed.sum()) <= 1e-9 * abs(mixed.sum())


def test_codon_random_codes_with_unknowns(hip_device):
    """Uniform random codons plus unknown states (codes >= 61): every tip
    state combination, padding of the last pattern tile included."""
    K, L, P = 4, 10, 333
    rates = gamma_rates(0.5, K)
    gy, pi, model = _gy94(7)
    rng = np.random.default_rng(8)
    par, ch, h = random_tree(L, rng, mean_branch=0.3)
    codes = rng.integers(0, 61, size=(L, P)).astype(np.uint8)
    unk = rng.random((L, P)) < 0.1
    codes[unk] = rng.choice([61, 62, 200, 255], size=unk.sum())
    codes[:, 5] = 255  # an all-unknown column: ln L = 0 in every category
    tree = O.Tree(L, par, ch, h)
    internals, children = _full(tree)
    mats = model_matrices(model, tree, children, rates)
    ref = _oracle(codes, P, 61, internals, children, mats, pi)
    calc = Calculator.HIPCodon(61).make(codes, P, K)
    calc.propose(internals, children, mats, pi)
    got = np.array([calc.likelihoods(k) for k in range(K)])
    assert np.all(np.abs(got[:, 5]) <= 2 * 61 * U16)
    err_mix = np.abs(_mix(got) - _mix(ref))
    print(f"random codes: mixed max {err_mix.max():.3e}, per-category max "
          f"{np.abs(got - ref).max():.3e}")
    assert err_mix.max() <= _bound(L)


def test_codon_device_matrices_match_host_matrices(hip_device):
    """P(t) built on the device (FP64 eigen + V diag(e^{lambda t}) V^-1, then
    FP16) gives the same likelihoods as FP64 host matrices to within a few
    FP16 half-ulps."""
    K, L, P = 4, 24, 900
    rates = gamma_rates(0.5, K)
    gy, pi, model = _gy94(11)
    par, ch, h, codes = _simulated(L, P, 12, gy, pi, rates)
    tree = O.Tree(L, par, ch, h)
    internals, children = _full(tree)
    mats = model_matrices(model, tree, children, rates)
    host = Calculator.HIPCodon(61).make(codes, P, K)
    host.propose(internals, children, mats, pi)
    dev = Calculator.HIPCodon(61).make(codes, P, K)
    dev.set_model(gy.q_matrix(), pi)
    flat = children.reshape(-1).astype(np.int64)
    times = h[par[flat]] - h[flat]
    dev.propose_times(internals, children, times, rates)
    ref = _oracle(codes, P, 61, internals, children, mats, pi)
    for k in range(K):
        d = np.abs(dev.likelihoods(k) - host.likelihoods(k))
        assert d.max() <= _bound(L), (k, d.max())
        assert np.abs(dev.likelihoods(k) - ref[k]).max() <= _bound(L)


def test_codon_partial_updates_accept_reject(hip_device):
    """Height moves re-evaluate only the changed path; accept/reject follow
    the two-slot EditBuf like the FP32 path."""
    K, L, P = 4, 30, 500
    rates = gamma_rates(0.5, K)
    gy, pi, model = _gy94(21)
    par, ch, h, codes = _simulated(L, P, 22, gy, pi, rates)
    otree = O.Tree(L, par, ch, h)
    tree = DatedTree(L, par, ch, h)
    internals, children = _full(otree)
    _full(tree)
    calc = Calculator.HIPCodon(61).make(codes, P, K)
    oc = [O.CalcN(codes, P, 61, 30, 8) for _ in range(K)]
    mats = model_matrices(model, otree, children, rates)
    calc.propose(internals, children, mats, pi)
    for k in range(K):
        oc[k].propose(internals, children, mats[k], pi)
    calc.accept()
    for k in range(K):
        oc[k].accept()
    rng = np.random.default_rng(23)
    for step in range(10):
        node = int(rng.integers(L, 2 * L - 1))
        a, c = tree.children_of(node)
        lo = max(tree.heights[a], tree.heights[c])
        parent = tree.parent_of(node)
        hi = tree.heights[parent] if parent is not None else lo + 0.3
        new_h = lo + (hi - lo) * rng.uniform(0.1, 0.9)
        tree.set_height(node, new_h)
        otree.set_height(node, new_h)
        internals, children = otree.partials_lists()
        tree.partials_lists()
        mats = model_matrices(model, otree, children, rates)
        calc.propose(internals, children, mats, pi)
        for k in range(K):
            oc[k].propose(internals, children, mats[k], pi)
            err = np.abs(calc.likelihoods(k) - oc[k].likelihoods())
            assert err.max() <= _bound(L), (step, k, err.max())
        if rng.random() < 0.5:
            calc.accept()
            otree.accept()
            tree.accept()
            for k in range(K):
                oc[k].accept()
        else:
            calc.reject()
            otree.reject()
            tree.reject()
            for k in range(K):
                oc[k].reject()


def test_codon_deep_tree_scaling(hip_device):
    """A 300-taxon caterpillar: partial products underflow FP16 (and FP32)
    many times over without the per-node power-of-two exponents."""
    K, L, P = 2, 300, 128
    rates = gamma_rates(1.0, K)
    gy, pi, model = _gy94(31)
    rng = np.random.default_rng(32)
    N = 2 * L - 1
    parents = np.full(N, 0xFFFFFFFF, np.uint32)
    children = np.zeros((L - 1, 2), np.uint32)
    heights = np.zeros(N)
    prev = 0
    for i in range(1, L):
        node = L + i - 1
        children[i - 1] = (prev, i)
        parents[prev] = node
        parents[i] = node
        heights[node] = 0.05 * i
        prev = node
    codes = rng.integers(0, 61, size=(L, P)).astype(np.uint8)
    tree = O.Tree(L, parents, children, heights)
    internals, ch = _full(tree)
    mats = model_matrices(model, tree, ch, rates)
    ref = _oracle(codes, P, 61, internals, ch, mats, pi)
    assert ref.min() < -1500  # far below FP32's range as a plain product
    calc = Calculator.HIPCodon(61).make(codes, P, K)
    calc.propose(internals, ch, mats, pi)
    got = np.array([calc.likelihoods(k) for k in range(K)])
    rel = np.abs(got - ref) / np.abs(ref)
    print(f"caterpillar: max rel {rel.max():.3e}")
    assert rel.max() <= 2e-3


@pytest.mark.parametrize("S", [5, 20, 64])
def test_mfma_other_state_counts(hip_device, S):
    """The MFMA path is written for up to 64 states; a random reversible
    model on S states (amino acids: 20) against the oracle."""
    K, L, P = 2, 16, 257
    rng = np.random.default_rng(S)
    R = rng.gamma(1.0, 1.0, size=(S, S))
    R = np.triu(R, 1)
    R = R + R.T
    pi = rng.dirichlet([5] * S)
    model = O.Model.reversible(R, pi)
    par, ch, h = random_tree(L, rng, mean_branch=0.1)
    rates = gamma_rates(0.5, K)
    tree = O.Tree(L, par, ch, h)
    internals, children = _full(tree)
    mats = model_matrices(model, tree, children, rates)
    # simulate on the tree from the oracle's own matrices
    codes = np.zeros((L, P), np.uint8)
    Nn = 2 * L - 1
    st = np.zeros((Nn, P), np.int64)
    root = Nn - 1
    st[root] = rng.choice(S, size=P, p=pi)
    for node in sorted(range(L, Nn), key=lambda n: -h[n]):
        for c in ch[node - L]:
            Pm = np.clip(model.transition(h[node] - h[c]), 0, None)
            cum = np.cumsum(Pm, 1)
            u = rng.random(P) * cum[st[node], -1]
            st[c] = np.minimum((u[:, None] >= cum[st[node]]).sum(1), S - 1)
    codes[:] = st[:L]
    codes[rng.random((L, P)) < 0.05] = S  # unknown
    ref = _oracle(codes, P, S, internals, children, mats, pi)
    calc = Calculator.HIPCodon(S).make(codes, P, K)
    calc.propose(internals, children, mats, pi)
    got = np.array([calc.likelihoods(k) for k in range(K)])
    err = np.abs(got - ref)
    print(f"S={S}: max {err.max():.3e}")
    assert err.max() <= _bound(L)
