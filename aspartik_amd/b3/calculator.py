"""Calculator configuration and the HIP calculator handle.

Mirrors b3's ``Calculator`` configuration object (crates/b3/src/calculator/
mod.rs:94-156: ``Calculator.CPU()``, ``.CUDAPartials()``, ``.with_scale()``)
with MI355X backends, and the ``Calculator<f64>`` trait (mod.rs:54-76)
through the C ABI of libb3hip.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib
from .._lib import addr, check, lib


class Calculator:
    """Backend choice + scaling, like b3's ``CalculatorConfig``."""

    def __init__(self, precision: int, device: int = 0, tile_patterns: int = 0,
                 scale_ln: int = 30, states: int = 4, lane_patterns: int = 0,
                 lane_states: int = 0):
        self.precision = int(precision)
        self.device = int(device)
        self.tile_patterns = int(tile_patterns)
        self.lane_patterns = int(lane_patterns)
        self.lane_states = int(lane_states)
        self.scale_ln = int(scale_ln)
        self.states = int(states)

    @classmethod
    def HIP(cls, device: int = 0, tile_patterns: int = 0, lane_patterns: int = 0,
            lane_states: int = 0):
        """FP32 partials with exact power-of-two scaling (the product path).
        tile_patterns / 64: wavefronts per category in a workgroup;
        lane_patterns: patterns per lane (1, 2; 0 = by size); lane_states: 4
        (default) or 1, four lanes per pattern for small pattern counts."""
        return cls(_lib.B3H_F32_SCALED, device, tile_patterns, lane_patterns=lane_patterns,
                   lane_states=lane_states)

    @classmethod
    def HIPPlainF64(cls, device: int = 0):
        """FP64 plain pruning: b3's GPU algorithm at FP64 (the baseline)."""
        return cls(_lib.B3H_F64_PLAIN, device)

    @classmethod
    def HIPCodon(cls, states: int = 61, device: int = 0):
        """FP16 partials on the MFMA matrix cores (32x32x16 f16 tiles, FP32
        accumulation) with per-node power-of-two scaling, for 2..64 states
        (GY94 codon models: 61)."""
        return cls(_lib.B3H_F16_MFMA, device, states=states)

    @classmethod
    def HIPPropagationsF64(cls, device: int = 0):
        """Propagation storage at FP64: b3's default GPU calculator
        (calculator/cuda_propagations, mod.rs:114-128 ``CUDAPropagations``)."""
        return cls(_lib.B3H_F64_PROPAGATION, device)

    @classmethod
    def HIPPropagations(cls, device: int = 0):
        """Propagation storage with FP32 projections and power-of-two
        exponents: the storage alternative to ``HIP()``."""
        return cls(_lib.B3H_F32_PROPAGATION, device)

    @property
    def propagations(self) -> bool:
        return self.precision in (_lib.B3H_F64_PROPAGATION, _lib.B3H_F32_PROPAGATION)

    def with_scale(self, scale_ln: int):
        """calculator/mod.rs:149-155"""
        self.scale_ln = int(scale_ln)
        return self

    def make(self, samples, num_patterns: int, num_categories: int = 1):
        """CalculatorConfig::make4 (calculator/mod.rs:159-200) for K
        categories in one handle."""
        return HipCalculator(self, samples, num_patterns, num_categories)

    def __repr__(self):
        names = {0: "HIP", 1: "HIPPlainF64", 2: "HIPCodon", 3: "HIPPropagationsF64",
                 4: "HIPPropagations"}
        return (f"Calculator.{names.get(self.precision, self.precision)}(device={self.device}, "
                f"scale_ln={self.scale_ln})")


def _sizes(*ok):
    """The arrays of a propose are passed as bare pointers: their sizes are
    checked here."""
    if not all(ok):
        raise ValueError("array sizes do not match the update list")


class HipCalculator:
    """One libb3hip handle: K rate categories of one alignment."""

    def __init__(self, config: Calculator, samples, num_patterns: int, num_categories: int):
        samples = np.ascontiguousarray(samples, np.uint8).reshape(-1)
        self.P = int(num_patterns)
        if samples.size % self.P:
            raise ValueError("samples size is not a multiple of num_patterns")
        self.L = samples.size // self.P
        self.K = int(num_categories)
        self.S = int(config.states)
        cfg = _lib.Config()
        lib().b3h_config_default(C.byref(cfg))
        cfg.device = config.device
        cfg.precision = config.precision
        cfg.states = config.states
        cfg.scale_ln = config.scale_ln
        cfg.tile_patterns = config.tile_patterns
        cfg.lane_patterns = config.lane_patterns
        cfg.lane_states = config.lane_states
        h = C.c_void_p()
        check(lib().b3h_create(C.byref(cfg), samples, self.L, self.P, self.K, C.byref(h)))
        self._h = h
        self.propagations = config.propagations

    def close(self):
        if getattr(self, "_h", None):
            lib().b3h_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- Calculator<f64> -------------------------------------------------
    def propose(self, internals, children, matrices, frequencies):
        nodes = np.ascontiguousarray(internals, np.uint32)
        ch = np.ascontiguousarray(children, np.uint32)
        m = np.ascontiguousarray(matrices, np.float64)
        fr = np.ascontiguousarray(frequencies, np.float64)
        _sizes(ch.size == 2 * nodes.size, m.size == self.K * 2 * nodes.size * self.S * self.S,
               fr.size >= self.S)
        check(lib().b3h_propose(self._h, nodes.size, addr(nodes), addr(ch), addr(m), addr(fr)))

    def propose_propagations(self, nodes, children, num_leaves, matrices, frequencies):
        """CudaPropagationsCalculator::propose: propagation lists and one
        matrix per listed node but the root, [K][len(nodes) - 1][4][4]."""
        nodes = np.ascontiguousarray(nodes, np.uint32)
        ch = np.ascontiguousarray(children, np.uint32)
        m = np.ascontiguousarray(matrices, np.float64)
        fr = np.ascontiguousarray(frequencies, np.float64)
        check(lib().b3h_propose_propagations(self._h, nodes.size, addr(nodes), int(num_leaves),
                                             addr(ch), addr(m), addr(fr)))

    def propose_propagations_times(self, nodes, children, num_leaves, edge_times,
                                   category_rates):
        """edge_times: time length of each listed node's edge but the root's."""
        nodes = np.ascontiguousarray(nodes, np.uint32)
        ch = np.ascontiguousarray(children, np.uint32)
        t = np.ascontiguousarray(edge_times, np.float64)
        r = np.ascontiguousarray(category_rates, np.float64)
        check(lib().b3h_propose_propagations_times(self._h, nodes.size, addr(nodes),
                                                   int(num_leaves), addr(ch), addr(t), addr(r)))

    def propose_times(self, internals, children, edge_times, category_rates):
        nodes = np.ascontiguousarray(internals, np.uint32)
        ch = np.ascontiguousarray(children, np.uint32)
        t = np.ascontiguousarray(edge_times, np.float64)
        r = np.ascontiguousarray(category_rates, np.float64)
        _sizes(ch.size == 2 * nodes.size, t.size == 2 * nodes.size, r.size == self.K)
        check(lib().b3h_propose_times(self._h, nodes.size, addr(nodes), addr(ch), addr(t),
                                      addr(r)))

    def set_model(self, q, frequencies, special_cases: bool = True):
        flags = _lib.B3H_MODEL_SPECIAL_CASES if special_cases else 0
        check(lib().b3h_set_model(self._h, np.ascontiguousarray(q, np.float64).reshape(-1),
                                  np.ascontiguousarray(frequencies, np.float64), flags))

    def set_category_model(self, category, q, frequencies, special_cases: bool = True):
        """Model of one category alone (HeteroLikelihood, hetero.rs:122-135)."""
        flags = _lib.B3H_MODEL_SPECIAL_CASES if special_cases else 0
        check(lib().b3h_set_category_model(
            self._h, int(category), np.ascontiguousarray(q, np.float64).reshape(-1),
            np.ascontiguousarray(frequencies, np.float64), flags))

    def propose_lengths(self, internals, children, lengths):
        """lengths [K][2n]: branch length of each child edge per category."""
        nodes = np.ascontiguousarray(internals, np.uint32)
        lengths = np.ascontiguousarray(lengths, np.float64)
        if lengths.size != self.K * 2 * nodes.size:
            raise ValueError("lengths must hold K x 2n values")
        ch = np.ascontiguousarray(children, np.uint32)
        _sizes(ch.size == 2 * nodes.size)
        check(lib().b3h_propose_lengths(self._h, nodes.size, addr(nodes), addr(ch),
                                        addr(lengths)))

    def set_classes(self, classes):
        """Per-pattern category (hetero.rs:150-160); None: category mean."""
        if classes is None:
            check(lib().b3h_set_classes(self._h, None))
            return
        classes = np.ascontiguousarray(classes, np.uint32)
        if classes.size != self.P:
            raise ValueError("classes must hold one value per pattern")
        check(lib().b3h_set_classes(self._h, classes.ctypes.data_as(C.c_void_p)))

    def accept(self):
        check(lib().b3h_accept(self._h))

    def reject(self):
        check(lib().b3h_reject(self._h))

    def likelihoods(self, category: int = 0):
        out = np.empty(self.P, np.float64)
        check(lib().b3h_likelihoods(self._h, int(category), out))
        return out

    def site_mixed(self):
        out = np.empty(self.P, np.float64)
        check(lib().b3h_site_mixed(self._h, out))
        return out

    def total(self) -> float:
        v = C.c_double()
        check(lib().b3h_total(self._h, C.byref(v)))
        return v.value

    def num_patterns(self) -> int:
        return int(lib().b3h_num_patterns(self._h))

    def set_weights(self, weights):
        check(lib().b3h_set_weights(self._h, np.ascontiguousarray(weights, np.uint32)))

    # ---- measurement -----------------------------------------------------
    def profile_enable(self, on: bool = True):
        check(lib().b3h_profile_enable(self._h, int(bool(on))))

    def profile_read(self):
        ms = C.c_double()
        n = C.c_uint64()
        check(lib().b3h_profile_read(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def profile_launch_times(self):
        """Per-launch kernel times (ms) since profile_enable, in launch order."""
        n = C.c_uint64()
        check(lib().b3h_profile_launch_times(self._h, None, 0, C.byref(n)))
        out = np.zeros(n.value, np.float64)
        if n.value:
            check(lib().b3h_profile_launch_times(self._h, out.ctypes.data, n.value, C.byref(n)))
        return out

    def device_bytes(self) -> int:
        return int(lib().b3h_device_bytes(self._h))


def gamma_rates(alpha: float, num_categories: int):
    """Discrete-gamma category rates (likelihood/gamma.rs:205-225)."""
    out = np.empty(int(num_categories), np.float64)
    check(lib().b3h_gamma_rates(float(alpha), int(num_categories), out))
    return out
