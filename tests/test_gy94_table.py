"""GY94 rate matrix against entries written out by hand from the model's
definition (Goldman & Yang 1994), not from the product's GY94 code.

For codons i != j one nucleotide apart, q_ij = pi_j * (kappa if the change is
a transition) * (omega if it changes the amino acid); codons two or three
nucleotides apart have rate 0; the three stop codons of the standard code
(TAA, TAG, TGA) are not states; Q is scaled to mean rate 1.  With F3x4,
pi_codon = f1(n1) f2(n2) f3(n3) / (sum over sense codons).

The expected values below use only the codon strings and their amino acids
as written here, so a kappa/omega swap, omega on a synonymous change, a
nonzero multi-nucleotide rate or a stop codon among the states each fails
one of these checks.  These checks pin the Q that the codon parity tests
then hand to the oracle.
"""
import itertools

import numpy as np
import pytest

from aspartik_amd.b3 import substitution as S

OMEGA, KAPPA = 0.3, 2.0
F3X4 = np.array([[0.1, 0.2, 0.3, 0.4],    # position 1: A C G T
                 [0.25, 0.15, 0.35, 0.25],
                 [0.4, 0.3, 0.2, 0.1]])
NUC = {"A": 0, "C": 1, "G": 2, "T": 3}
STOPS = {"TAA", "TAG", "TGA"}


def _pi_unnorm(codon):
    return np.prod([F3X4[p, NUC[codon[p]]] for p in range(3)])


def _pi(codon):
    norm = sum(_pi_unnorm("".join(c)) for c in itertools.product("ACGT", repeat=3)
               if "".join(c) not in STOPS)
    return _pi_unnorm(codon) / norm


@pytest.fixture(scope="module")
def gy():
    m = S.GY94(OMEGA, KAPPA, F3X4)
    q = m.q_matrix()
    idx = {c: i for i, c in enumerate(S.SENSE_CODONS)}
    return m, q, idx


def test_states_are_the_61_sense_codons(gy):
    _, q, idx = gy
    assert q.shape == (61, 61)
    assert len(idx) == 61
    for stop in STOPS:
        assert stop not in idx
    # every one of the other 61 codons is a state
    for c in itertools.product("ACGT", repeat=3):
        c = "".join(c)
        if c not in STOPS:
            assert c in idx


def test_frequencies_are_f3x4(gy):
    m, _, idx = gy
    pi = m.frequencies()
    for codon in ("TTT", "ATG", "GGG", "CTA", "TGG"):
        assert pi[idx[codon]] == pytest.approx(_pi(codon), rel=1e-13)


# (from, to, multiplier of pi_to): hand-classified single-nucleotide changes
#   TTT (Phe) -> TTC (Phe): T<->C transition, synonymous        -> kappa
#   TTT (Phe) -> TTA (Leu): T<->A transversion, nonsynonymous   -> omega
#   TTT (Phe) -> TCT (Ser): T<->C transition, nonsynonymous     -> kappa omega
#   CTT (Leu) -> CTA (Leu): T<->A transversion, synonymous      -> 1
#   ATG (Met) -> ATA (Ile): G<->A transition, nonsynonymous     -> kappa omega
#   ATG (Met) -> ATT (Ile): G<->T transversion, nonsynonymous   -> omega
#   GGG (Gly) -> GGA (Gly): G<->A transition, synonymous        -> kappa
#   AGA (Arg) -> CGA (Arg): A<->C transversion, synonymous      -> 1
#   TGG (Trp) -> TGT (Cys): G<->T transversion, nonsynonymous   -> omega
#   TAT (Tyr) -> TAC (Tyr): T<->C transition, synonymous        -> kappa
HAND = [
    ("TTT", "TTC", KAPPA),
    ("TTT", "TTA", OMEGA),
    ("TTT", "TCT", KAPPA * OMEGA),
    ("CTT", "CTA", 1.0),
    ("ATG", "ATA", KAPPA * OMEGA),
    ("ATG", "ATT", OMEGA),
    ("GGG", "GGA", KAPPA),
    ("AGA", "CGA", 1.0),
    ("TGG", "TGT", OMEGA),
    ("TAT", "TAC", KAPPA),
]


def test_hand_classified_entries_up_to_the_common_scale(gy):
    _, q, idx = gy
    # the mean-rate scale is common to every entry: take it from the first
    scale = q[idx["TTT"], idx["TTC"]] / (KAPPA * _pi("TTC"))
    for a, b, mult in HAND:
        want = mult * _pi(b) * scale
        assert q[idx[a], idx[b]] == pytest.approx(want, rel=1e-12), (a, b)
        # reverse direction: the same multiplier times pi of the other codon
        assert q[idx[b], idx[a]] == pytest.approx(mult * _pi(a) * scale, rel=1e-12), (b, a)


def test_mean_rate_one_fixes_the_scale(gy):
    """With the hand-derived rates, mean rate 1 gives the scale in closed
    form: 1 / sum_i pi_i sum_{j ~ i} m_ij pi_j, summed over the hand
    classification of every one-nucleotide change below."""
    _, q, idx = gy
    amino = {}
    code = ("TTT F TTC F TTA L TTG L CTT L CTC L CTA L CTG L ATT I ATC I ATA I ATG M "
            "GTT V GTC V GTA V GTG V TCT S TCC S TCA S TCG S CCT P CCC P CCA P CCG P "
            "ACT T ACC T ACA T ACG T GCT A GCC A GCA A GCG A TAT Y TAC Y CAT H CAC H "
            "CAA Q CAG Q AAT N AAC N AAA K AAG K GAT D GAC D GAA E GAG E TGT C TGC C "
            "TGG W CGT R CGC R CGA R CGG R AGT S AGC S AGA R AGG R GGT G GGC G GGA G "
            "GGG G").split()
    for c, a in zip(code[::2], code[1::2]):
        amino[c] = a
    assert len(amino) == 61
    purines = {"A", "G"}

    def mult(ci, cj):
        d = [p for p in range(3) if ci[p] != cj[p]]
        if len(d) != 1:
            return 0.0
        x, y = ci[d[0]], cj[d[0]]
        m = KAPPA if ((x in purines) == (y in purines)) else 1.0
        return m * (OMEGA if amino[ci] != amino[cj] else 1.0)

    mean = sum(_pi(ci) * mult(ci, cj) * _pi(cj) for ci in amino for cj in amino)
    for ci in amino:
        for cj in amino:
            m = mult(ci, cj)
            if m:
                want = m * _pi(cj) / mean
                assert q[idx[ci], idx[cj]] == pytest.approx(want, rel=1e-12), (ci, cj)


def test_multi_nucleotide_changes_have_rate_zero(gy):
    _, q, idx = gy
    for ci, i in idx.items():
        for cj, j in idx.items():
            d = sum(ci[p] != cj[p] for p in range(3))
            if d >= 2:
                assert q[i, j] == 0.0, (ci, cj)
            elif d == 1:
                assert q[i, j] > 0.0, (ci, cj)
    # rows sum to zero
    np.testing.assert_allclose(q.sum(1), 0.0, atol=1e-13)


def test_kappa_and_omega_are_not_interchangeable():
    """Swapping the two parameters changes the hand-classified entries: the
    check above distinguishes them (a model built with them swapped fails)."""
    m = S.GY94(KAPPA, OMEGA, F3X4)  # omega and kappa swapped
    q = m.q_matrix()
    idx = {c: i for i, c in enumerate(S.SENSE_CODONS)}
    r_syn_ts = q[idx["TTT"], idx["TTC"]] / _pi("TTC")     # should be kappa * scale
    r_nonsyn_tv = q[idx["TTT"], idx["TTA"]] / _pi("TTA")  # should be omega * scale
    assert r_syn_ts / r_nonsyn_tv == pytest.approx(OMEGA / KAPPA, rel=1e-12)
    assert r_syn_ts / r_nonsyn_tv != pytest.approx(KAPPA / OMEGA, rel=1e-3)
