"""MCMC over the GPU likelihood: the native step driver, ESS, and the MLP
surrogate of delayed-acceptance MH (include/b3mcmc.h).

``Chain`` mirrors b3's ``MCMC`` loop (crates/b3/src/mcmc.rs:270-393) for a
dated tree under GTR+Gamma with b3's NodeSlide, TreeScale, DeltaExchange,
ScaleReal and NarrowExchange operators.  ``Surrogate`` is the 404 -> 256 ->
256 -> 1 GELU network (BF16 weights, FP32 accumulation) trained on
(state features, exact ln L) pairs from burn-in; ``Chain.run(mode="delayed")``
uses it as the first stage of Christen & Fox's delayed acceptance.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib
from .._lib import check, lib


def ess(x) -> float:
    """Geyer's initial monotone sequence ESS of a scalar trace."""
    x = np.ascontiguousarray(x, np.float64)
    return float(lib().b3m_ess(x, x.size))


class Surrogate:
    """MLP ln L surrogate: inputs -> hidden -> hidden -> 1, GELU."""

    def __init__(self, inputs: int, hidden: int = 256, seed: int = 4, device: int = 0):
        self.inputs, self.hidden = int(inputs), int(hidden)
        h = C.c_void_p()
        check(lib().b3m_mlp_create(int(device), self.inputs, self.hidden, int(seed), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().b3m_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_params(self) -> int:
        return int(lib().b3m_mlp_num_params(self._h))

    def set_scaling(self, x_mean, x_std, y_mean: float, y_std: float):
        check(lib().b3m_mlp_set_scaling(self._h, np.ascontiguousarray(x_mean, np.float32),
                                        np.ascontiguousarray(x_std, np.float32),
                                        float(y_mean), float(y_std)))

    def fit_scaling(self, x, y, min_std: float = 1e-3):
        """Standardise inputs and output with the training set's moments."""
        x = np.asarray(x, np.float64)
        y = np.asarray(y, np.float64)
        self.set_scaling(x.mean(0), np.maximum(x.std(0), min_std), y.mean(), max(y.std(), 1e-9))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, self.inputs)
        out = np.empty(x.shape[0], np.float32)
        check(lib().b3m_mlp_forward(self._h, x, x.shape[0], out))
        return out

    def loss_grad(self, x, y):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, self.inputs)
        y = np.ascontiguousarray(y, np.float32)
        g = np.empty(self.num_params, np.float32)
        loss = C.c_float()
        check(lib().b3m_mlp_loss_grad(self._h, x, y, x.shape[0], C.byref(loss), g))
        return float(loss.value), g

    def params(self):
        p = np.empty(self.num_params, np.float32)
        check(lib().b3m_mlp_get_params(self._h, p))
        return p

    def set_params(self, p):
        check(lib().b3m_mlp_set_params(self._h, np.ascontiguousarray(p, np.float32)))

    def train(self, x, y, epochs=50, batch=256, lr=1e-3) -> float:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, self.inputs)
        y = np.ascontiguousarray(y, np.float32)
        loss = C.c_float()
        check(lib().b3m_mlp_train(self._h, x, y, x.shape[0], int(epochs), int(batch), float(lr),
                                  C.byref(loss)))
        return float(loss.value)

    def unpack(self, p):
        """Split a flat parameter vector into W1, b1, W2, b2, w3, b3."""
        D, H = self.inputs, self.hidden
        o = np.cumsum([0, D * H, H, H * H, H, H, 1])
        return (p[o[0]:o[1]].reshape(D, H), p[o[1]:o[2]], p[o[2]:o[3]].reshape(H, H),
                p[o[3]:o[4]], p[o[4]:o[5]], p[o[5]:o[6]])


class Chain:
    """GTR+Gamma MCMC on a dated tree, likelihood on the GPU."""

    def __init__(self, codes, parents, children, heights, gtr_rates, freqs, alpha,
                 weights=None, **config):
        codes = np.ascontiguousarray(codes, np.uint8)
        L, P = codes.shape
        cfg = _lib.ChainConfig()
        lib().b3m_chain_config_default(C.byref(cfg))
        for k, v in config.items():
            if not hasattr(cfg, k):
                raise TypeError(f"unknown chain setting {k}")
            setattr(cfg, k, v)
        self.config = cfg
        self.L, self.P = L, P
        w = None if weights is None else np.ascontiguousarray(weights, np.uint32)
        self._w = w
        h = C.c_void_p()
        parents = np.asarray(parents).astype(np.int64)
        parents[parents < 0] = 0xFFFFFFFF
        check(lib().b3m_chain_create(
            C.byref(cfg), codes.reshape(-1), L, P,
            None if w is None else w.ctypes.data_as(C.c_void_p),
            np.ascontiguousarray(parents, np.uint32),
            np.ascontiguousarray(children, np.uint32).reshape(-1),
            np.ascontiguousarray(heights, np.float64),
            np.ascontiguousarray(gtr_rates, np.float64),
            np.ascontiguousarray(freqs, np.float64), float(alpha), C.byref(h)))
        self._h = h
        self.surrogate = None

    def close(self):
        if getattr(self, "_h", None):
            lib().b3m_chain_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_features(self) -> int:
        return int(lib().b3m_chain_num_features(self._h))

    def features(self):
        out = np.empty(self.num_features, np.float32)
        check(lib().b3m_chain_features(self._h, out))
        return out

    def state(self):
        ll, lp, post = C.c_double(), C.c_double(), C.c_double()
        check(lib().b3m_chain_state(self._h, C.byref(ll), C.byref(lp), C.byref(post)))
        return ll.value, lp.value, post.value

    def params(self):
        """(parents, children, heights, gtr_rates, alpha) of the current state"""
        N = 2 * self.L - 1
        par = np.empty(N, np.uint32)
        ch = np.empty(2 * (self.L - 1), np.uint32)
        h = np.empty(N, np.float64)
        r = np.empty(6, np.float64)
        a = C.c_double()
        check(lib().b3m_chain_params(self._h, par, ch, h, r, C.byref(a)))
        return par, ch.reshape(-1, 2), h, r, a.value

    def dump_state(self) -> bytes:
        """The chain's state in b3's verbatim state format (mcmc.rs:250-268)."""
        n = C.c_size_t()
        check(lib().b3m_chain_dump_state(self._h, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        check(lib().b3m_chain_dump_state(self._h, buf, n.value, C.byref(n)))
        return buf.raw[:n.value]

    def load_state(self, state: bytes):
        """Mcmc::load_state (mcmc.rs:183-213): parameters, scheduler and
        generator; the likelihood is evaluated from scratch."""
        state = bytes(state)
        check(lib().b3m_chain_load_state(self._h, state, len(state)))

    def run_traced(self, steps, every, writer, tree=None, rates=None, alpha=None,
                   mode="plain"):
        """Run ``steps`` steps and hand the state to a ``TraceWriter`` every
        ``every`` steps (trace_writer.rs:347-369).  ``tree`` (DatedTree),
        ``rates`` (RealVector) and ``alpha`` (Real) are the mirrors the
        writer's items reference; they are set from the chain before each
        row.  Returns the concatenated run statistics of the chunks."""
        from .parameters import Real, RealVector
        from .tree import DatedTree
        done, stats = 0, []
        while done < steps:
            n = min(every, steps - done)
            _, st, _ = self.run(n, mode=mode, record_every=n)
            stats.append(st)
            done += n
            par, ch, h, r, a = self.params()
            if isinstance(tree, DatedTree):
                tree.load_parts(par, h)
                tree.accept()
            if isinstance(rates, RealVector):
                for j in range(6):
                    rates[j] = r[j]
                rates.accept()
            if isinstance(alpha, Real):
                alpha.set(a)
                alpha.accept()
            ll, lp, post = self.state()
            life = int(np.frombuffer(self.dump_state()[8:16], "<u8")[0])
            writer.call(life, post, lp, ll)
        return stats

    def set_surrogate(self, s: Surrogate | None):
        self.surrogate = s
        check(lib().b3m_chain_set_surrogate(self._h, None if s is None else s._h))

    def run(self, steps, mode="plain", record_every=1, collect=0):
        """Run `steps` steps; returns (trace [rows][13], stats dict, (X, y))."""
        steps = int(steps)
        m = _lib.B3M_DELAYED if mode == "delayed" else _lib.B3M_PLAIN
        rows = (steps + record_every - 1) // record_every
        trace = np.zeros((rows, _lib.B3M_TRACE_COLS), np.float64)
        nf = self.num_features
        cx = np.zeros((collect, nf), np.float32) if collect else None
        cy = np.zeros(collect, np.float32) if collect else None
        st = _lib.RunStats()
        check(lib().b3m_chain_run(
            self._h, steps, m, int(record_every), trace.ctypes.data_as(C.c_void_p),
            None if cx is None else cx.ctypes.data_as(C.c_void_p),
            None if cy is None else cy.ctypes.data_as(C.c_void_p), int(collect), C.byref(st)))
        arrays = ("op_proposed", "op_accepted", "tuning")
        stats = {f: getattr(st, f) for f, _ in _lib.RunStats._fields_ if f not in arrays}
        for f in arrays:
            stats[f] = list(getattr(st, f))
        data = None
        if collect:
            n = int(st.collected)
            data = (cx[:n], cy[:n])
        return trace, stats, data
