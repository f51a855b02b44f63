"""FASTA ingest and host deduplication (include/b3data.h) on the CPU.

Parsing is pinned by the rules of FastaParser::parse_line
(crates/data/src/fasta/mod.rs:103-150) and Msa::from_fasta_reader
(crates/data/src/msa/mod.rs:77-119) case by case, by the round trip the
reference's own test checks (crates/data/tests/it/fasta.rs:30-41: a record
written with Record::to_string parses back equal) and by its MSA tests
(tests/it/msa.rs: two records, uneven lengths rejected).  The data/
alignments the reference's Python tests read are not in the repository, so
none of them is used.  Deduplication is pinned by the oracle's or_deduplicate
(first-occurrence order, counts, round trip).
"""
import numpy as np
import pytest

from aspartik_amd import B3hError
from aspartik_amd.data import MSA, deduplicate
from oracle import oracle as O

ASCII = "-ACMGRSVTWYHKDBN"   # DnaNucleotide byte -> character (nucleotides.rs:223-247)


def _seq(codes):
    return "".join(ASCII[c] for c in codes)


def test_two_records_and_names():
    m = MSA.from_fasta_str(">a\nAC\n>b\nGT\n")
    assert m.num_sequences == 2 and m.num_sites == 2
    assert m.sequence_names() == ["a", "b"]
    assert m.sequence(0) == "AC" and m.sequence(1) == "GT"
    np.testing.assert_array_equal(m.data, [[1, 2], [4, 8]])


def test_alphabet_case_and_multiline():
    text = ">x description with spaces\nacgtWSMKRY\nBDHVN-\n\n>y\nACGTWSMKRYBDHVN-\n"
    m = MSA.from_fasta_str(text)
    assert m.sequence_name(0) == "x description with spaces"
    np.testing.assert_array_equal(m.data[0], m.data[1])
    np.testing.assert_array_equal(m.data[1], [1, 2, 4, 8, 9, 6, 3, 12, 5, 10, 14, 13, 11, 7, 15, 0])


def test_comments_and_line_endings():
    text = ";top comment\r\n>a\r\nAC\r\n   ;indented comment\r\nGT\r\n>b\r\nACGT"
    m = MSA.from_fasta_str(text)
    assert m.sequence_names() == ["a", "b"]
    assert m.sequence(0) == "ACGT" and m.sequence(1) == "ACGT"


def test_errors():
    with pytest.raises(B3hError, match="before any record"):
        MSA.from_fasta_str("ACGT\n>a\nAC\n")
    with pytest.raises(B3hError, match="before any record"):
        MSA.from_fasta_str("\n>a\nAC\n")          # a blank line is sequence data
    with pytest.raises(B3hError, match="invalid character 'X'"):
        MSA.from_fasta_str(">a\nACXT\n")
    with pytest.raises(B3hError, match="invalid character"):
        MSA.from_fasta_str(">a\nAC GT\n")
    with pytest.raises(B3hError, match="invalid character"):
        MSA.from_fasta_str(">a\nACGT\r")          # no "\n": the "\r" is kept
    with pytest.raises(B3hError, match="sites"):
        MSA.from_fasta_str(">a\nAC\n>b\nG\n")     # tests/it/msa.rs:29-34
    with pytest.raises(B3hError):
        MSA.from_fasta_file("/nonexistent/file.fasta")


def test_empty_input_is_one_unnamed_empty_sequence():
    m = MSA.from_fasta_str("")
    assert m.num_sequences == 1 and m.num_sites == 0 and m.sequence_names() == [""]


def test_round_trip_of_random_records(tmp_path):
    """Record::to_string (fasta/mod.rs:66-86): '>' + description + newline +
    the sequence on one line; parsed back equal (tests/it/fasta.rs:30-41)."""
    rng = np.random.default_rng(0)
    for trial in range(20):
        S, N = int(rng.integers(1, 9)), int(rng.integers(1, 300))
        codes = rng.integers(0, 16, size=(S, N)).astype(np.uint8)
        names = [f"seq{i} some comment {trial}" for i in range(S)]
        text = "\n".join(f">{n}\n{_seq(c)}" for n, c in zip(names, codes))
        m = MSA.from_fasta_str(text)
        assert m.sequence_names() == names
        np.testing.assert_array_equal(m.data, codes)
        p = tmp_path / f"t{trial}.fasta"
        p.write_text(text)
        assert MSA.from_fasta_file(p) == m


def test_base_frequencies_sequential_sum():
    """msa/mod.rs:193-212: the per-character split added in sequence order."""
    rng = np.random.default_rng(3)
    codes = rng.integers(0, 16, size=(5, 333)).astype(np.uint8)
    m = MSA.from_fasta_str("\n".join(f">s{i}\n{_seq(c)}" for i, c in enumerate(codes)))
    third = 1.0 / 3.0
    split = {}
    for v in range(16):
        bits = [b for b in range(4) if v >> b & 1]
        if v in (0, 15):
            split[v] = [0.25] * 4
        else:
            share = {1: 1.0, 2: 0.5, 3: third}[len(bits)]
            split[v] = [share if b in bits else 0.0 for b in range(4)]
    acc = [0.0] * 4
    for v in codes.reshape(-1):
        for k in range(4):
            acc[k] += split[int(v)][k]
    ref = tuple(a / codes.size for a in acc)
    assert m.base_frequencies() == ref
    assert abs(sum(ref) - 1.0) < 1e-12


@pytest.mark.parametrize("S,N,dup", [(1, 1, 0.0), (1, 50, 0.5), (9, 2000, 0.6), (40, 5000, 0.0),
                                     (3, 4000, 0.95)])
def test_host_dedup_matches_oracle(S, N, dup):
    rng = np.random.default_rng(S * 7 + N)
    a = rng.choice([0, 1, 2, 4, 8, 15, 5, 10], size=(S, N)).astype(np.uint8)
    n_dup = int(dup * N)
    if n_dup and N > 1:
        src = rng.integers(0, N, size=n_dup)
        dst = rng.integers(0, N, size=n_dup)
        a[:, dst] = a[:, src]
    leaves, w, s2p = deduplicate(a)
    ol, ow, os2p = O.deduplicate(a)
    np.testing.assert_array_equal(leaves, ol)
    np.testing.assert_array_equal(w, ow)
    np.testing.assert_array_equal(s2p, os2p)
    assert w.sum() == N
    # round trip: the pattern of every site restores its column (gap -> 15)
    restored = leaves[:, s2p]
    np.testing.assert_array_equal(restored, np.where(a == 0, 15, a))


def test_gap_and_any_are_distinct_patterns():
    """Columns are compared before the gap -> 0b1111 rewrite (mod.rs:80-84,117-122)."""
    a = np.array([[0, 15, 0], [1, 1, 1]], np.uint8)
    leaves, w, s2p = deduplicate(a)
    assert leaves.shape[1] == 2 and list(w) == [2, 1] and list(s2p) == [0, 1, 0]
    np.testing.assert_array_equal(leaves, [[15, 15], [1, 1]])


def test_dedup_argument_errors():
    with pytest.raises(B3hError):
        deduplicate(np.full((2, 3), 16, np.uint8))   # not a DnaNucleotide byte
    leaves, w, s2p = deduplicate(np.zeros((2, 0), np.uint8))
    assert leaves.shape == (2, 0) and w.size == 0 and s2p.size == 0
