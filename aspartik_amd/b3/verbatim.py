"""b3's verbatim binary format and its state files.

``VerbatimWriter`` / ``VerbatimReader`` follow crates/verbatim/src/lib.rs:
little-endian primitives, each padded with zeros to a multiple of its own
size counted from the start of the stream; tags written as they are; a
slice is a u32 length, then (when non-empty) padding to the element
alignment, then the elements.

``dump_state`` / ``load_state`` write and read Mcmc's state (mcmc.rs:
183-268): the tag 62 33 b3 44, protocol version 19, the step (u64), the
posterior (f64), every parameter in order (Parameter::dump of DatedTree,
Real, RealVector, IntVector, ClassVector), the scheduler (per operator u32
accepts, u32 rejects and the f64 tuning of tunable operators) and the
generator (two u128 words).
"""
from __future__ import annotations

import struct

import numpy as np

STATE_TAG = bytes([0x62, 0x33, 0xB3, 0x44])
STATE_PROTOCOL_VERSION = 19

_PRIM = {"u8": ("<B", 1), "i8": ("<b", 1), "u16": ("<H", 2), "i16": ("<h", 2),
         "u32": ("<I", 4), "i32": ("<i", 4), "u64": ("<Q", 8), "i64": ("<q", 8),
         "f32": ("<f", 4), "f64": ("<d", 8)}


class VerbatimWriter:
    def __init__(self):
        self.buf = bytearray()

    def _pad(self, align):
        r = len(self.buf) % align
        if r:
            self.buf.extend(b"\0" * (align - r))

    def write(self, kind: str, value):
        if kind in ("u128", "i128"):
            self._pad(16)
            self.buf.extend(int(value).to_bytes(16, "little", signed=kind == "i128"))
            return
        fmt, size = _PRIM[kind]
        if size > 1:
            self._pad(size)
        self.buf.extend(struct.pack(fmt, value))

    def write_bool(self, value: bool):
        self.write("u8", 1 if value else 0)

    def write_tag(self, tag: bytes):
        self.buf.extend(tag)

    def write_slice(self, values: np.ndarray):
        values = np.ascontiguousarray(values)
        self.write("u32", values.size)
        if values.size:
            self._pad(values.dtype.alignment)
        self.buf.extend(values.astype(values.dtype.newbyteorder("<"), copy=False).tobytes())

    def getvalue(self) -> bytes:
        return bytes(self.buf)


class VerbatimReader:
    def __init__(self, data: bytes):
        self.data = memoryview(bytes(data))
        self.pos = 0

    def _need(self, n):
        if self.pos + n > len(self.data):
            raise EOFError("unexpected end while filling the buffer")

    def _pad(self, align):
        r = self.pos % align
        if r:
            self._need(align - r)
            self.pos += align - r

    def read(self, kind: str):
        if kind in ("u128", "i128"):
            self._pad(16)
            self._need(16)
            v = int.from_bytes(self.data[self.pos:self.pos + 16], "little",
                               signed=kind == "i128")
            self.pos += 16
            return v
        fmt, size = _PRIM[kind]
        if size > 1:
            self._pad(size)
        self._need(size)
        (v,) = struct.unpack_from(fmt, self.data, self.pos)
        self.pos += size
        return v

    def read_bool(self) -> bool:
        v = self.read("u8")
        if v not in (0, 1):
            raise ValueError(f"invalid boolean value: {v:x}")
        return bool(v)

    def read_tag(self, tag: bytes):
        self._need(len(tag))
        if bytes(self.data[self.pos:self.pos + len(tag)]) != tag:
            raise ValueError(f"expected tag: {list(tag)}")
        self.pos += len(tag)

    def read_slice(self, dtype) -> np.ndarray:
        dtype = np.dtype(dtype).newbyteorder("<")
        n = self.read("u32")
        if n:
            self._pad(dtype.alignment)
        self._need(n * dtype.itemsize)
        out = np.frombuffer(self.data, dtype, count=n, offset=self.pos).copy()
        self.pos += n * dtype.itemsize
        return out


# ---- parameters --------------------------------------------------------------

def dump_parameter(w: VerbatimWriter, param):
    """Parameter::dump of the mirrors in aspartik_amd.b3 (tree.rs:1011-1016,
    real.rs:59-62, real_vector.rs:77-82, int_vector.rs:77-82,
    class_vector.rs:82-88)."""
    from .parameters import ClassVector, IntVector, Real, RealVector
    from .tree import DatedTree
    if isinstance(param, DatedTree):
        w.write_slice(param.parents_u32())
        w.write_slice(np.asarray(param.heights, np.float64))
    elif isinstance(param, Real):
        w.write("f64", param.value)
    elif isinstance(param, RealVector):
        for v in param.values:
            w.write("f64", float(v))
    elif isinstance(param, IntVector):
        for v in param.values:
            w.write("i64", int(v))
    elif isinstance(param, ClassVector):
        for v in param.values:
            w.write("u32", int(v))
    else:
        raise TypeError(f"cannot dump {type(param).__name__}")


def load_parameter(r: VerbatimReader, param):
    """Parameter::load (tree.rs:997-1009 marks every edge; the others set
    every value so that the parameter reports a change)."""
    from .parameters import ClassVector, IntVector, Real, RealVector
    from .tree import DatedTree
    if isinstance(param, DatedTree):
        parents = r.read_slice(np.uint32)
        heights = r.read_slice(np.float64)
        param.load_parts(parents, heights)
    elif isinstance(param, Real):
        param.set(r.read("f64"))
    elif isinstance(param, RealVector):
        for i in range(len(param)):
            param[i] = r.read("f64")
    elif isinstance(param, IntVector):
        for i in range(len(param)):
            param[i] = r.read("i64")
    elif isinstance(param, ClassVector):
        for i in range(len(param)):
            param.set(i, r.read("u32"))
    else:
        raise TypeError(f"cannot load {type(param).__name__}")


def dump_state(step: int, posterior: float, parameters, operators=(), rng=(0, 0)) -> bytes:
    """Mcmc::dump (mcmc.rs:250-268).  operators: objects with ``accepts``,
    ``rejects`` and ``tuning`` (None when the operator has none); rng: the
    generator's two u128 words."""
    w = VerbatimWriter()
    w.write_tag(STATE_TAG)
    w.write("u32", STATE_PROTOCOL_VERSION)
    w.write("u64", int(step))
    w.write("f64", float(posterior))
    for p in parameters:
        dump_parameter(w, p)
    for op in operators:
        w.write("u32", int(op.accepts))
        w.write("u32", int(op.rejects))
        if op.tuning is not None:
            w.write("f64", float(op.tuning))
    w.write("u128", int(rng[0]))
    w.write("u128", int(rng[1]))
    return w.getvalue()


def load_state(data: bytes, parameters, operators=()):
    """Mcmc::load_state (mcmc.rs:183-213) into the given objects; returns
    (step, posterior, (rng word 0, rng word 1))."""
    r = VerbatimReader(data)
    try:
        r.read_tag(STATE_TAG)
    except (ValueError, EOFError) as e:
        raise ValueError("Not a b3 state file") from e
    version = r.read("u32")
    if version != STATE_PROTOCOL_VERSION:
        raise ValueError(f"Cannot load state: protocol version v{version} is incompatible "
                         f"with the currently supported version v{STATE_PROTOCOL_VERSION}.")
    step = r.read("u64")
    posterior = r.read("f64")
    for p in parameters:
        load_parameter(r, p)
    for op in operators:
        op.accepts = r.read("u32")
        op.rejects = r.read("u32")
        if op.tuning is not None:
            op.tuning = r.read("f64")
    rng = (r.read("u128"), r.read("u128"))
    return step, posterior, rng
