"""MCMC traces as Arrow IPC files, and replaying them.

``TraceWriter`` writes what b3's TraceWriter callback writes
(crates/b3/src/callback/trace_writer.rs:168-470): an Arrow IPC file whose
columns are ``step`` (u64), ``posterior``, ``prior`` and ``likelihood``
(f64), then one column per item -- a ``Real`` or a prior (anything with
``probability()``) as f64, a ``RealVector`` / ``IntVector`` / ``ClassVector``
as a fixed-size list of f64 / i64 / u32, and a ``DatedTree`` as four
columns ``{name}:parents`` (fixed-size list of u32, 0xFFFFFFFF for the
root), ``{name}:heights`` (fixed-size list of f64), ``{name}:length`` (the
tree length) and ``{name}:height`` (the root height).  Columns are not
nullable; the schema carries the metadata keys aspartik, os, arch and time;
rows are flushed as a record batch once the buffered arrays reach 100 MB,
and at ``finish``.  Compression is optional zstd.

``replicate_b3`` is python/utils/replicate.py:9-33: every row's parameters
are loaded into the given objects and each likelihood (and prior) is
compared with the row's value.
"""
from __future__ import annotations

import os
import platform
import time

import numpy as np
import pyarrow as pa
import pyarrow.ipc as ipc

from .parameters import ClassVector, IntVector, Real, RealVector
from .tree import TOMBSTONE, DatedTree

FLUSH_BYTES = 100_000_000


def _metadata():
    from .. import __version__
    return {"aspartik": f"aspartik_amd {__version__}", "os": platform.system().lower(),
            "arch": platform.machine(), "time": str(int(time.time()))}


class _Column:
    def __init__(self, name, typ, getter, width=1):
        self.field = pa.field(name, typ, nullable=False)
        self.getter = getter
        self.values = []
        self.width = width

    def array(self):
        typ = self.field.type
        if pa.types.is_fixed_size_list(typ):
            flat = np.concatenate(self.values) if self.values else np.zeros(0)
            flat = pa.array(flat.astype(typ.value_type.to_pandas_dtype()), typ.value_type)
            return pa.FixedSizeListArray.from_arrays(flat, typ.list_size)
        return pa.array(self.values, typ)


class TraceWriter:
    def __init__(self, items: dict, path, *, every: int, zstd: bool = False,
                 overwrite: bool = False):
        path = os.fspath(path)
        if os.path.isfile(path) and not overwrite:
            raise FileExistsError(f"File {path!r} already exists. Add `overwrite=True` to "
                                  "replace it")
        self.every = int(every)
        self._scalars = {"step": [], "posterior": [], "prior": [], "likelihood": []}
        cols = []
        for name, item in items.items():
            if isinstance(item, DatedTree):
                n = item.num_nodes
                cols += [
                    _Column(f"{name}:parents", pa.list_(pa.uint32(), n),
                            lambda t=item: t.parents_u32(), n),
                    _Column(f"{name}:heights", pa.list_(pa.float64(), n),
                            lambda t=item: np.asarray(t.heights, np.float64), n),
                    _Column(f"{name}:length", pa.float64(), lambda t=item: t.total_length()),
                    _Column(f"{name}:height", pa.float64(),
                            lambda t=item: float(t.heights[t.root])),
                ]
            elif isinstance(item, Real):
                cols.append(_Column(name, pa.float64(), lambda r=item: r.value))
            elif isinstance(item, RealVector):
                cols.append(_Column(name, pa.list_(pa.float64(), len(item)),
                                    lambda v=item: v.values, len(item)))
            elif isinstance(item, IntVector):
                cols.append(_Column(name, pa.list_(pa.int64(), len(item)),
                                    lambda v=item: v.values, len(item)))
            elif isinstance(item, ClassVector):
                cols.append(_Column(name, pa.list_(pa.uint32(), len(item)),
                                    lambda v=item: v.values.astype(np.uint32), len(item)))
            elif hasattr(item, "probability"):
                cols.append(_Column(name, pa.float64(), lambda p=item: float(p.probability())))
            else:
                raise TypeError(f"`TraceWriter` doesn't support type {type(item).__name__}")
        self._cols = cols
        fields = [pa.field("step", pa.uint64(), nullable=False)] + [
            pa.field(k, pa.float64(), nullable=False) for k in ("posterior", "prior", "likelihood")
        ] + [c.field for c in cols]
        self.schema = pa.schema(fields, metadata=_metadata())
        opts = ipc.IpcWriteOptions(compression="zstd" if zstd else None)
        self._sink = pa.OSFile(path, "wb")
        self._writer = ipc.new_file(self._sink, self.schema, options=opts)
        self._buffered = 0

    def call(self, step: int, posterior: float, prior: float, likelihood: float):
        """One row (trace_writer.rs:347-369)."""
        for k, v in (("step", int(step)), ("posterior", posterior), ("prior", prior),
                     ("likelihood", likelihood)):
            self._scalars[k].append(v)
        self._buffered += 32
        for c in self._cols:
            v = c.getter()
            c.values.append(np.array(v) if c.width > 1 or isinstance(v, np.ndarray) else v)
            self._buffered += 8 * c.width
        if self._buffered >= FLUSH_BYTES:
            self._write_batch()

    def _write_batch(self):
        if not self._scalars["step"]:
            return
        arrays = [pa.array(self._scalars["step"], pa.uint64())] + [
            pa.array(self._scalars[k], pa.float64()) for k in ("posterior", "prior", "likelihood")
        ] + [c.array() for c in self._cols]
        self._writer.write_batch(pa.record_batch(arrays, schema=self.schema))
        for k in self._scalars:
            self._scalars[k].clear()
        for c in self._cols:
            c.values.clear()
        self._buffered = 0

    def finish(self):
        self._write_batch()
        self._writer.close()
        self._sink.close()


def read_trace(path) -> pa.Table:
    with pa.memory_map(os.fspath(path), "r") as src:
        return ipc.open_file(src).read_all()


def _assign(param, value):
    if isinstance(param, Real):
        param.set(float(value))
    elif isinstance(param, (RealVector, IntVector)):
        for j, v in enumerate(value):
            param[j] = v
    elif isinstance(param, ClassVector):
        for j, v in enumerate(value):
            param.set(j, int(v))
    elif callable(param):
        param(value)
    else:
        raise TypeError(f"Parameter {type(param).__name__} isn't supported")


def replicate_b3(trace_path, parameters: dict, priors: dict | None = None,
                 likelihoods=(), tolerance: float = 1e-9) -> float:
    """replicate.py:9-33 and _check (replicate.py:108-125): load each row and
    compare; ``parameters`` values are DatedTree / Real / RealVector /
    IntVector / ClassVector mirrors or callables taking the row's value.
    Returns the largest relative likelihood difference seen."""
    table = read_trace(trace_path)
    cols = {name: table.column(name).to_pylist() for name in table.column_names}
    worst = 0.0
    for i in range(table.num_rows):
        for name, param in parameters.items():
            if isinstance(param, DatedTree):
                parents = np.asarray(cols[f"{name}:parents"][i], np.uint32)
                heights = np.asarray(cols[f"{name}:heights"][i], np.float64)
                param.load_parts(parents, heights)
            else:
                _assign(param, cols[name][i])
        for name, prior in (priors or {}).items():
            expected = float(cols[name][i])
            got = float(prior.probability())
            rel = abs(expected - got) / abs(expected)
            assert rel < 1e-12, f"got {got}, expected {expected}, prior: {prior}"
        for lik in likelihoods:
            expected = float(cols["likelihood"][i])
            got = lik.likelihood()
            rel = abs(got - expected) / abs(expected)
            worst = max(worst, rel)
            assert rel < tolerance, (
                f"{type(lik).__name__}: got {got}, expected {expected} ({rel}) at row {i}")
            lik.reject()
    return worst


__all__ = ["TraceWriter", "read_trace", "replicate_b3", "TOMBSTONE"]
