"""Substitution models: rate matrices handed to the device eigensolver.

b3's models (crates/b3/src/substitution.rs) build Q on the host and
decompose it there; here the host only forms the normalised Q and pi and
the eigensystem and P(t) are computed on the GPU (b3h_set_model /
b3h_propose_times).  ``special_cases`` reports whether the model uses the
t > 20 and t < 1e-10 shortcuts (HKY and GTR do, substitution.rs:408-414,
576-582; JC and K80 use closed forms without them).  The Goldman-Yang codon
model has no counterpart in b3.
"""
from __future__ import annotations

import numpy as np


class _Model:
    special_cases = False
    states = 4

    def __init__(self):
        self._version = 0

    def _touch(self):
        self._version += 1

    @property
    def version(self):
        return self._version

    def q_matrix(self) -> np.ndarray:
        raise NotImplementedError

    def frequencies(self) -> np.ndarray:
        raise NotImplementedError


class JC(_Model):
    """substitution.rs:102-156"""

    def q_matrix(self):
        q = np.full((4, 4), 1.0 / 3.0)
        np.fill_diagonal(q, -1.0)
        return q

    def frequencies(self):
        return np.full(4, 0.25)


class K80(_Model):
    """substitution.rs:158-243 (A C G T; transitions A<->G, C<->T)"""

    def __init__(self, kappa):
        super().__init__()
        self.kappa = float(kappa)

    def set_kappa(self, kappa):
        self.kappa = float(kappa)
        self._touch()

    def q_matrix(self):
        k = self.kappa
        q = np.array([[0, 1, k, 1], [1, 0, 1, k], [k, 1, 0, 1], [1, k, 1, 0]], np.float64)
        np.fill_diagonal(q, -q.sum(1))
        return q / (-(np.diag(q)).mean())

    def frequencies(self):
        return np.full(4, 0.25)


class HKY(_Model):
    """substitution.rs:270-431: Q normalised by div = 2(gt + ac + at + cg +
    kappa (ag + ct))."""

    special_cases = True

    def __init__(self, frequencies, kappa):
        super().__init__()
        self.freqs = np.asarray(frequencies, np.float64).copy()
        self.kappa = float(kappa)

    def set_kappa(self, kappa):
        self.kappa = float(kappa)
        self._touch()

    def set_frequencies(self, f):
        self.freqs = np.asarray(f, np.float64).copy()
        self._touch()

    def q_matrix(self):
        a, c, g, t = self.freqs
        k = self.kappa
        q = np.array([[0.0, c, k * g, t], [a, 0.0, g, k * t], [k * a, c, 0.0, t],
                      [a, k * c, g, 0.0]])
        div = 2.0 * (g * t + a * c + a * t + c * g + k * (a * g + c * t))
        np.fill_diagonal(q, -q.sum(1))
        return q / div

    def frequencies(self):
        return self.freqs.copy()


class GTR(_Model):
    """substitution.rs:433-604: rates (AC, AG, AT, CG, CT, GT), normalised by
    div = 2 sum r_ij pi_i pi_j."""

    special_cases = True

    def __init__(self, frequencies, rates):
        super().__init__()
        self.freqs = np.asarray(frequencies, np.float64).copy()
        self.rates = np.asarray(rates, np.float64).copy()

    def set_rates(self, r):
        self.rates = np.asarray(r, np.float64).copy()
        self._touch()

    def set_frequencies(self, f):
        self.freqs = np.asarray(f, np.float64).copy()
        self._touch()

    def q_matrix(self):
        pa, pc, pg, pt = self.freqs
        a, b, c, d, e, f = self.rates
        q = np.array([
            [-a * pc - b * pg - c * pt, a * pc, b * pg, c * pt],
            [a * pa, -a * pa - d * pg - e * pt, d * pg, e * pt],
            [b * pa, d * pc, -b * pa - d * pc - f * pt, f * pt],
            [c * pa, e * pc, f * pg, -c * pa - e * pc - f * pg],
        ])
        div = 2.0 * (a * pa * pc + b * pa * pg + c * pa * pt + d * pc * pg + e * pc * pt
                     + f * pg * pt)
        return q / div

    def frequencies(self):
        return self.freqs.copy()


# ---- codons ---------------------------------------------------------------
_NUC = "ACGT"
# standard genetic code in TCAG order (codon index 16 i1 + 4 i2 + i3)
_CODE_TCAG = "FFLLSSSSYY**CC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG"
_TCAG = {"T": 0, "C": 1, "A": 2, "G": 3}


def _amino(codon: str) -> str:
    i = 16 * _TCAG[codon[0]] + 4 * _TCAG[codon[1]] + _TCAG[codon[2]]
    return _CODE_TCAG[i]


SENSE_CODONS = [a + b + c for a in _NUC for b in _NUC for c in _NUC if _amino(a + b + c) != "*"]
assert len(SENSE_CODONS) == 61


def _is_transition(x, y):
    return {x, y} in ({"A", "G"}, {"C", "T"})


class GY94(_Model):
    """Goldman & Yang (1994) codon model over the 61 sense codons of the
    standard code: q_ij = pi_j * (kappa if transition) * (omega if
    nonsynonymous) for codons one nucleotide apart, 0 otherwise; F3x4
    frequencies pi_codon = prod_pos f_pos(nucleotide), renormalised over
    sense codons; Q scaled to mean rate 1."""

    states = 61

    def __init__(self, omega, kappa, f3x4):
        super().__init__()
        self.omega = float(omega)
        self.kappa = float(kappa)
        self.f3x4 = np.asarray(f3x4, np.float64).reshape(3, 4).copy()

    def frequencies(self):
        pi = np.array([np.prod([self.f3x4[p, _NUC.index(c[p])] for p in range(3)])
                       for c in SENSE_CODONS])
        return pi / pi.sum()

    def exchangeability(self):
        n = len(SENSE_CODONS)
        R = np.zeros((n, n))
        for i, ci in enumerate(SENSE_CODONS):
            for j, cj in enumerate(SENSE_CODONS):
                if i == j:
                    continue
                diff = [p for p in range(3) if ci[p] != cj[p]]
                if len(diff) != 1:
                    continue
                p = diff[0]
                r = self.kappa if _is_transition(ci[p], cj[p]) else 1.0
                if _amino(ci) != _amino(cj):
                    r *= self.omega
                R[i, j] = r
        return R

    def q_matrix(self):
        pi = self.frequencies()
        R = self.exchangeability()
        q = R * pi[None, :]
        np.fill_diagonal(q, 0.0)
        np.fill_diagonal(q, -q.sum(1))
        mean = -(pi * np.diag(q)).sum()
        return q / mean
