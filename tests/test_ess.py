"""Geyer initial-monotone-sequence ESS (b3m_ess, host code of the library)
against the analytic ESS of AR(1) series, n (1 - rho) / (1 + rho)."""
import numpy as np
import pytest

from aspartik_amd.b3 import ess


def _ar1(n, rho, seed):
    rng = np.random.default_rng(seed)
    e = rng.normal(size=n)
    x = np.empty(n)
    x[0] = e[0] / np.sqrt(1 - rho * rho)
    for t in range(1, n):
        x[t] = rho * x[t - 1] + e[t]
    return x


@pytest.mark.parametrize("rho", [0.0, 0.5, 0.9, -0.3])
def test_ar1_known_answer(rho):
    n = 200_000
    vals = [ess(_ar1(n, rho, s)) / n for s in range(3)]
    want = (1 - rho) / (1 + rho) if rho >= 0 else 1.0
    # negative rho: Geyer's estimator truncates at the first non-positive pair
    # sum and is capped at n by construction of the monotone sequence
    if rho < 0:
        assert all(v >= 0.99 for v in vals), vals
    else:
        assert abs(np.mean(vals) - want) < 0.04 * want + 0.01, (vals, want)


def test_degenerate_traces():
    assert ess(np.ones(100)) == 0.0
    assert ess(np.arange(3.0)) == 0.0
    x = np.random.default_rng(0).normal(size=5000)
    assert 4000 < ess(x) <= 5000 * 1.2


def test_planted_mistake_is_caught():
    """A trace that is the AR(1) series repeated twice (each value kept for
    two steps) has about half the ESS; an estimator that ignored lag-1
    correlation would not see it."""
    x = _ar1(100_000, 0.5, 7)
    base = ess(x)
    doubled = ess(np.repeat(x, 2))
    assert 0.8 < doubled / base < 1.25  # same information in twice the samples
