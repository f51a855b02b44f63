"""Distributions with the surface b3's clocks use: ``inverse_cdf`` and
``is_changed`` (crates/distributions/src/lib.rs:8-128).

b3 wraps the un-vendored ``computare-distributions`` crate; the quantile
functions here are the textbook closed forms of the same parametrisations
(LogNormal(location, scale) = exp(Normal(location, scale)), Exponential by
rate, Uniform(min, max), Normal(mean, std), Gamma(shape, scale)).  Each
parameter is a float or an object with ``__float__`` (a ``Real``); the
distribution reports ``is_changed`` when any parameter that has an
``is_changed`` method does (lib.rs:116-124).
"""
from __future__ import annotations

import math
from statistics import NormalDist

_STD_NORMAL = NormalDist()


def _is_changed(*params) -> bool:
    return any(getattr(p, "is_changed", lambda: False)() for p in params)


class _Dist:
    _fields: tuple = ()

    def is_changed(self) -> bool:
        return _is_changed(*(getattr(self, f) for f in self._fields))

    def __repr__(self):
        args = ", ".join(f"{f}={float(getattr(self, f))}" for f in self._fields)
        return f"{type(self).__name__}({args})"


def _check_p(p: float) -> float:
    p = float(p)
    if not 0.0 <= p <= 1.0:
        raise ValueError("probability out of [0, 1]")
    return p


class Normal(_Dist):
    _fields = ("mean", "std")

    def __init__(self, mean, std):
        self.mean, self.std = mean, std

    def inverse_cdf(self, p: float) -> float:
        p = _check_p(p)
        if p in (0.0, 1.0):
            return -math.inf if p == 0.0 else math.inf
        return float(self.mean) + float(self.std) * _STD_NORMAL.inv_cdf(p)


class LogNormal(_Dist):
    _fields = ("location", "scale")

    def __init__(self, location, scale):
        self.location, self.scale = location, scale

    def inverse_cdf(self, p: float) -> float:
        p = _check_p(p)
        if p in (0.0, 1.0):
            return 0.0 if p == 0.0 else math.inf
        return math.exp(float(self.location) + float(self.scale) * _STD_NORMAL.inv_cdf(p))


class Exponential(_Dist):
    _fields = ("rate",)

    def __init__(self, rate):
        self.rate = rate

    def inverse_cdf(self, p: float) -> float:
        p = _check_p(p)
        return -math.log1p(-p) / float(self.rate)


class Uniform(_Dist):
    _fields = ("min", "max")

    def __init__(self, min, max):  # noqa: A002 - b3's parameter names
        self.min, self.max = min, max

    def inverse_cdf(self, p: float) -> float:
        p = _check_p(p)
        lo, hi = float(self.min), float(self.max)
        return lo + p * (hi - lo)


class Gamma(_Dist):
    _fields = ("shape", "scale")

    def __init__(self, shape, scale):
        self.shape, self.scale = shape, scale

    def inverse_cdf(self, p: float) -> float:
        from scipy.special import gammaincinv
        p = _check_p(p)
        return float(gammaincinv(float(self.shape), p)) * float(self.scale)
