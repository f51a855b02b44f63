"""Batched evaluation of many independent likelihood instances (b3h_batch_*
in include/b3hip.h): one persistent kernel launch evaluates every instance's
update list.  b3 has no batched calculator; each instance follows the same
semantics as one ``HipCalculator`` with K rate categories."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib
from .._lib import check, lib


class Batch:
    def __init__(self, instances, num_categories: int = 4, device: int = 0,
                 single_slot: bool = False, tile_patterns: int = 0, lane_patterns: int = 0,
                 lane_states: int = 0):
        """instances: sequence of (codes [L, P] u8, weights [P] u32 or None).
        Tile shape: tile_patterns / 64 wavefronts per category, lane_patterns
        patterns per lane (1 or 2; 0 = 2), lane_states states per lane (4, or
        1: four lanes per pattern; 0 = 4)."""
        self._keep = []
        arr = (_lib.Instance * len(instances))()
        for i, (codes, weights) in enumerate(instances):
            codes = np.ascontiguousarray(codes, np.uint8)
            L, P = codes.shape
            self._keep.append(codes)
            arr[i].samples = codes.ctypes.data
            arr[i].num_leaves = L
            arr[i].num_patterns = P
            if weights is not None:
                w = np.ascontiguousarray(weights, np.uint32)
                self._keep.append(w)
                arr[i].weights = w.ctypes.data
            else:
                arr[i].weights = None
        self.count = len(instances)
        self.K = int(num_categories)
        self.P = [np.asarray(c).shape[1] for c, _ in instances]
        cfg = _lib.Config()
        lib().b3h_config_default(C.byref(cfg))
        cfg.device = device
        cfg.tile_patterns = tile_patterns
        cfg.lane_patterns = lane_patterns
        cfg.lane_states = lane_states
        h = C.c_void_p()
        flags = _lib.B3H_BATCH_SINGLE_SLOT if single_slot else 0
        check(lib().b3h_batch_create(C.byref(cfg), self.count, arr, self.K, flags, C.byref(h)))
        self._h = h
        self._keep = []  # the library copied codes and weights

    def close(self):
        if getattr(self, "_h", None):
            lib().b3h_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_model(self, i, q, frequencies, special_cases=True):
        flags = _lib.B3H_MODEL_SPECIAL_CASES if special_cases else 0
        check(lib().b3h_batch_set_model(self._h, int(i),
                                        np.ascontiguousarray(q, np.float64).reshape(-1),
                                        np.ascontiguousarray(frequencies, np.float64), flags))

    def propose_times(self, updates, graph: bool = False):
        """updates: per instance None (not updated) or (internals, children,
        edge_times, category_rates)."""
        arr = (_lib.Update * self.count)()
        keep = []
        for i, u in enumerate(updates):
            if u is None:
                arr[i].n = 0
                continue
            internals, children, times, rates = (
                np.ascontiguousarray(u[0], np.uint32), np.ascontiguousarray(u[1], np.uint32),
                np.ascontiguousarray(u[2], np.float64), np.ascontiguousarray(u[3], np.float64))
            keep += [internals, children, times, rates]
            arr[i].n = internals.size
            arr[i].internals = internals.ctypes.data
            arr[i].children = children.ctypes.data
            arr[i].edge_times = times.ctypes.data
            arr[i].category_rates = rates.ctypes.data
        check(lib().b3h_batch_propose_times(self._h, arr, _lib.B3H_BATCH_GRAPH if graph else 0))

    def totals(self):
        out = np.empty(self.count, np.float64)
        check(lib().b3h_batch_totals(self._h, out))
        return out

    def likelihoods(self, i, category=0):
        out = np.empty(self.P[i], np.float64)
        check(lib().b3h_batch_likelihoods(self._h, int(i), int(category), out))
        return out

    def accept(self):
        check(lib().b3h_batch_accept(self._h))

    def reject(self):
        check(lib().b3h_batch_reject(self._h))

    def profile_enable(self, on=True):
        check(lib().b3h_batch_profile_enable(self._h, int(bool(on))))

    def profile_read(self):
        ms = C.c_double()
        n = C.c_uint64()
        check(lib().b3h_batch_profile_read(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def device_bytes(self):
        return int(lib().b3h_batch_device_bytes(self._h))
