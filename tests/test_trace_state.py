"""Verbatim state format, Arrow traces and trace replay, on the CPU.

* The writer's padding rules (crates/verbatim/src/lib.rs:9-90) byte by byte
  on hand-built cases, and writer/reader round trips of random streams.
* Mcmc state files (mcmc.rs:183-268): header, parameter order, scheduler and
  generator words; a foreign tag and another protocol version are refused.
* The trace schema of trace_writer.rs:347-470 (names, types, non-null,
  metadata keys), batching, zstd, overwrite protection.
* replicate_b3 (python/utils/replicate.py) against the oracle: rows written
  with the oracle's likelihood replay at 1e-12, and a planted error in one
  row is caught.
"""
import numpy as np
import pyarrow as pa
import pytest

from aspartik_amd.b3 import ClassVector, DatedTree, IntVector, Real, RealVector
from aspartik_amd.b3 import trace as T
from aspartik_amd.b3.verbatim import (STATE_TAG, VerbatimReader, VerbatimWriter, dump_state,
                                      load_state)
from oracle import oracle as O
from tests.helpers import random_codes, random_tree


def test_padding_rules_byte_by_byte():
    w = VerbatimWriter()
    w.write("u8", 1)
    w.write("u32", 2)            # padded to 4
    w.write_tag(b"xy")           # tags are not padded
    w.write("u64", 3)            # padded to 8
    w.write_slice(np.array([7, 8], np.uint32))   # u32 length at 24, elements at 28
    w.write_slice(np.zeros(0, np.float64))       # length at 36; empty: no element padding
    w.write("u128", 5)           # 40 -> padded to 48
    expect = (b"\x01\0\0\0" + b"\x02\0\0\0" + b"xy" + b"\0" * 6 + (3).to_bytes(8, "little")
              + (2).to_bytes(4, "little") + (7).to_bytes(4, "little") + (8).to_bytes(4, "little")
              + (0).to_bytes(4, "little") + b"\0" * 8 + (5).to_bytes(16, "little"))
    assert w.getvalue() == expect
    w2 = VerbatimWriter()
    w2.write("u8", 9)
    w2.write_slice(np.array([1.5], np.float64))  # length at 4, element padded to 8
    assert w2.getvalue() == b"\x09\0\0\0" + (1).to_bytes(4, "little") + \
        np.array([1.5]).tobytes()


def test_random_round_trips():
    rng = np.random.default_rng(0)
    kinds = ["u8", "i8", "u16", "i16", "u32", "i32", "u64", "i64", "f32", "f64", "u128",
             "tag", "slice32", "slice64", "bool"]
    for trial in range(50):
        ops = [kinds[i] for i in rng.integers(0, len(kinds), size=40)]
        vals = []
        w = VerbatimWriter()
        for k in ops:
            if k == "tag":
                v = bytes(rng.integers(0, 256, size=int(rng.integers(1, 6))).astype(np.uint8))
                w.write_tag(v)
            elif k == "slice32":
                v = rng.integers(0, 2**32, size=int(rng.integers(0, 5))).astype(np.uint32)
                w.write_slice(v)
            elif k == "slice64":
                v = rng.normal(size=int(rng.integers(0, 5)))
                w.write_slice(v)
            elif k == "bool":
                v = bool(rng.integers(0, 2))
                w.write_bool(v)
            elif k[0] == "f":
                v = float(np.float32(rng.normal())) if k == "f32" else float(rng.normal())
                w.write(k, v)
            else:
                bits = int(k[1:])
                lo = -(1 << (bits - 1)) if k[0] == "i" else 0
                hi = (1 << (bits - 1)) if k[0] == "i" else (1 << bits)
                v = int(rng.integers(lo, hi)) if bits <= 32 else int(lo + (hi - lo) // 3)
                w.write(k, v)
            vals.append(v)
        r = VerbatimReader(w.getvalue())
        for k, v in zip(ops, vals):
            if k == "tag":
                r.read_tag(v)
            elif k.startswith("slice"):
                np.testing.assert_array_equal(r.read_slice(v.dtype), v)
            elif k == "bool":
                assert r.read_bool() == v
            else:
                assert r.read(k) == v
        assert r.pos == len(w.getvalue())


def _params(seed=1, L=9):
    rng = np.random.default_rng(seed)
    par, ch, h = random_tree(L, rng)
    return (DatedTree(L, par, ch, h), RealVector(*rng.dirichlet([1] * 6)), Real(0.7),
            IntVector(3, -4, 5), ClassVector(4, 7, rng.integers(0, 4, size=7)))


class _Op:
    def __init__(self, a, r, t):
        self.accepts, self.rejects, self.tuning = a, r, t


def test_state_round_trip_and_layout():
    params = _params()
    ops = [_Op(3, 4, None), _Op(10, 2, 0.6), _Op(0, 0, 0.25)]
    data = dump_state(12345, -678.5, params, ops, rng=(2**127 + 5, 99))
    assert data[:4] == STATE_TAG == bytes([0x62, 0x33, 0xB3, 0x44])
    assert int.from_bytes(data[4:8], "little") == 19
    assert int.from_bytes(data[8:16], "little") == 12345
    # fresh objects of the same shape receive the values
    fresh = _params(seed=2)
    fops = [_Op(0, 0, None), _Op(0, 0, 0.0), _Op(0, 0, 0.0)]
    step, post, rng = load_state(data, fresh, fops)
    assert (step, post, rng) == (12345, -678.5, (2**127 + 5, 99))
    np.testing.assert_array_equal(fresh[0].parents_u32(), params[0].parents_u32())
    np.testing.assert_array_equal(fresh[0].heights, params[0].heights)
    # children rebuilt from parents in node order (tree.rs:913-927)
    np.testing.assert_array_equal(fresh[0].children, np.sort(params[0].children, axis=1))
    for a, b in zip(fresh[1:], params[1:]):
        np.testing.assert_array_equal(np.asarray(a.values if hasattr(a, "values") else a.value),
                                      np.asarray(b.values if hasattr(b, "values") else b.value))
    assert [(o.accepts, o.rejects, o.tuning) for o in fops] == \
        [(3, 4, None), (10, 2, 0.6), (0, 0, 0.25)]
    assert fresh[0].is_changed()  # tree.rs:1007: every edge marked after a load
    with pytest.raises(ValueError, match="Not a b3 state file"):
        load_state(b"\0" + data[1:], fresh, fops)
    bad = bytearray(data)
    bad[4] = 18
    with pytest.raises(ValueError, match="v18"):
        load_state(bytes(bad), fresh, fops)
    with pytest.raises(EOFError):
        load_state(data[:-3], fresh, fops)


def test_trace_schema_batches_and_zstd(tmp_path, monkeypatch):
    tree, rates, alpha, iv, cv = _params()

    class Prior:
        def probability(self):
            return -1.25

    monkeypatch.setattr(T, "FLUSH_BYTES", 400)   # several record batches
    path = tmp_path / "t.trace"
    w = T.TraceWriter({"tree": tree, "rates": rates, "alpha": alpha, "iv": iv, "cv": cv,
                       "prior:p": Prior()}, path, every=10, zstd=True)
    for i in range(7):
        alpha.set(0.1 * i)
        w.call(i * 10, -10.0 - i, -1.0, -9.0 - i)
    w.finish()
    with pytest.raises(FileExistsError):
        T.TraceWriter({"alpha": alpha}, path, every=1)
    tab = T.read_trace(path)
    names = [f.name for f in tab.schema]
    assert names == ["step", "posterior", "prior", "likelihood", "tree:parents",
                     "tree:heights", "tree:length", "tree:height", "rates", "alpha", "iv", "cv",
                     "prior:p"]
    assert all(not f.nullable for f in tab.schema)
    assert set(tab.schema.metadata) == {b"aspartik", b"os", b"arch", b"time"}
    assert tab.schema.field("tree:parents").type == pa.list_(pa.uint32(), tree.num_nodes)
    assert tab.schema.field("iv").type == pa.list_(pa.int64(), 3)
    assert tab.schema.field("cv").type == pa.list_(pa.uint32(), 7)
    assert tab.num_rows == 7
    assert tab.column("step").to_pylist() == [0, 10, 20, 30, 40, 50, 60]
    assert tab.column("alpha").to_pylist() == [0.1 * i for i in range(7)]
    assert tab.column("tree:parents")[0].as_py()[tree.root] == 0xFFFFFFFF
    assert tab.column("tree:height")[3].as_py() == tree.heights[tree.root]
    assert tab.column("tree:length")[0].as_py() == tree.total_length()
    with pa.memory_map(str(path), "r") as src:
        assert pa.ipc.open_file(src).num_record_batches > 1


class _OracleLikelihood:
    """A likelihood over the oracle for replay tests (DNA, one category)."""

    def __init__(self, tree, model, codes):
        self.tree, self.model, self.codes = tree, model, codes

    def likelihood(self):
        t = self.tree
        ot = O.Tree(t.L, t.parents_u32(), t.children.astype(np.uint32), t.heights)
        lik = O.GammaLikelihood(ot, self.model, self.codes, self.codes.shape[1],
                                np.ones(self.codes.shape[1], np.uint32), num_categories=1)
        return lik.likelihood()

    def reject(self):
        pass


def test_replicate_against_the_oracle(tmp_path):
    rng = np.random.default_rng(4)
    L, P = 12, 300
    par, ch, h = random_tree(L, rng)
    tree = DatedTree(L, par, ch, h)
    codes = random_codes(L, P, rng)
    model = O.Model.hky(2.0, rng.dirichlet([10] * 4))
    lik = _OracleLikelihood(tree, model, codes)
    path = tmp_path / "r.trace"
    w = T.TraceWriter({"tree": tree}, path, every=1)
    values = []
    for step in range(6):
        for node in range(L, 2 * L - 1):
            if node == tree.root:
                continue
            lo = max(tree.heights[c] for c in tree.children_of(node))
            tree.set_height(node, lo + (tree.heights[tree.parent_of(node)] - lo) * 0.5)
        values.append(lik.likelihood())
        w.call(step, values[-1], 0.0, values[-1])
    w.finish()
    replay = DatedTree(L, par, ch, h)
    worst = T.replicate_b3(path, {"tree": replay},
                           likelihoods=[_OracleLikelihood(replay, model, codes)], tolerance=1e-12)
    assert worst < 1e-12
    # a wrong likelihood in one row is caught
    tab = T.read_trace(path)
    bad = tab.column("likelihood").to_pylist()
    bad[3] *= 1 + 1e-8
    tab = tab.set_column(tab.schema.get_field_index("likelihood"),
                         pa.field("likelihood", pa.float64(), nullable=False),
                         pa.array(bad, pa.float64()))
    bad_path = tmp_path / "bad.trace"
    with pa.OSFile(str(bad_path), "wb") as sink, pa.ipc.new_file(sink, tab.schema) as wr:
        wr.write_table(tab)
    again = DatedTree(L, par, ch, h)
    with pytest.raises(AssertionError, match="row 3"):
        T.replicate_b3(bad_path, {"tree": again},
                       likelihoods=[_OracleLikelihood(again, model, codes)])
