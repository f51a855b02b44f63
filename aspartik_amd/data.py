"""Alignments: FASTA ingest and pattern deduplication (include/b3data.h).

``MSA`` mirrors the parts of aspartik's ``aspartik.data.msa.MSA``
(crates/data/src/msa/python.rs:37-106) the likelihood needs:
``from_fasta_file``, ``num_sequences``, ``num_sites``, ``sequence_name``,
``sequence_names``, ``sequence`` and ``base_frequencies``.  Characters are
DnaNucleotide bytes (crates/data/src/nucleotides.rs:36-68).
``deduplicate`` is likelihood/mod.rs:74-132 on the host or on a GPU.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import check, lib

_TO_ASCII = np.frombuffer(b"-ACMGRSVTWYHKDBN", np.uint8)


class MSA:
    def __init__(self, handle):
        self._h = handle
        S = int(lib().b3d_msa_num_sequences(handle))
        N = int(lib().b3d_msa_num_sites(handle))
        ptr = lib().b3d_msa_data(handle)
        self._data = (np.ctypeslib.as_array(ptr, shape=(S * N,)).copy().reshape(S, N)
                      if S * N else np.zeros((S, N), np.uint8))
        self._names = [lib().b3d_msa_name(handle, i).decode() for i in range(S)]

    def __del__(self):
        if getattr(self, "_h", None):
            lib().b3d_msa_destroy(self._h)
            self._h = None

    @classmethod
    def from_fasta_file(cls, path) -> "MSA":
        """msa/python.rs:37-42 (any path-like)."""
        h = C.c_void_p()
        check(lib().b3d_msa_from_fasta_file(os.fsencode(os.fspath(path)), C.byref(h)))
        return cls(h)

    @classmethod
    def from_fasta_str(cls, text) -> "MSA":
        raw = text.encode() if isinstance(text, str) else bytes(text)
        h = C.c_void_p()
        check(lib().b3d_msa_from_fasta(raw, len(raw), C.byref(h)))
        return cls(h)

    @property
    def num_sequences(self) -> int:
        return self._data.shape[0]

    @property
    def num_sites(self) -> int:
        return self._data.shape[1]

    def sequence_name(self, index: int) -> str:
        return self._names[index]

    def sequence_names(self) -> list[str]:
        return list(self._names)

    def sequence(self, index: int) -> str:
        return _TO_ASCII[self._data[index]].tobytes().decode()

    @property
    def data(self) -> np.ndarray:
        """[num_sequences][num_sites] DnaNucleotide bytes (a copy)."""
        return self._data.copy()

    def base_frequencies(self):
        """msa/mod.rs:193-212"""
        out = np.zeros(4, np.float64)
        check(lib().b3d_msa_base_frequencies(self._h, out))
        return tuple(float(x) for x in out)

    def __eq__(self, other):
        return (isinstance(other, MSA) and self._names == other._names
                and np.array_equal(self._data, other._data))


def deduplicate(msa, device: int | None = None, no_alloc: bool = False):
    """likelihood/mod.rs:74-132: unique columns in first-occurrence order, their
    counts as weights and each site's pattern; gaps become 0b1111 in the
    pattern bytes.  ``msa`` is an ``MSA`` or a [num_sequences, num_sites] u8
    array of DnaNucleotide bytes.  device None: on the host; an int: on that
    GPU.  ``no_alloc``: on the GPU through b3d_deduplicate_noalloc, which uses
    the library's fixed device scratch (64 MiB) instead of allocating.
    Returns (leaves [num_sequences, num_patterns], weights, sites_to_patterns)."""
    a = msa._data if isinstance(msa, MSA) else np.ascontiguousarray(msa, np.uint8)
    S, N = a.shape
    leaves = np.zeros((S, max(N, 1)), np.uint8)
    weights = np.zeros(max(N, 1), np.uint32)
    s2p = np.zeros(max(N, 1), np.uint64)
    P = C.c_uint64()
    fn = lib().b3d_deduplicate_noalloc if no_alloc else lib().b3d_deduplicate
    check(fn(np.ascontiguousarray(a).reshape(-1), S, N, -1 if device is None else int(device),
             leaves.reshape(-1), leaves.shape[1], weights, s2p, C.byref(P)))
    P = P.value
    return leaves[:, :P].copy(), weights[:P].copy(), s2p[:N].astype(np.int64)


__all__ = ["MSA", "deduplicate", "_lib"]
