"""ctypes wrapper of the FP64 CPU oracle (oracle/b3_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
Each wrapper names the reference function it restates (file:line, relative to
the aspartik repository root).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
szp = np.ctypeslib.ndpointer(np.uintp, flags="C_CONTIGUOUS")
TOMBSTONE = 0xFFFFFFFF


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class _Model(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("n", C.c_int),
        ("kappa", C.c_double),
        ("freqs", C.c_double * 64),
        ("q", C.c_double * 4096),
        ("evec", C.c_double * 4096),
        ("ievc", C.c_double * 4096),
        ("eval", C.c_double * 64),
    ]


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    vp = C.c_void_p
    sig = {
        "or_editbuf_set_edited": (None, [u8p, C.c_size_t]),
        "or_editbuf_accept": (None, [u8p, C.c_size_t]),
        "or_editbuf_reject": (None, [u8p, C.c_size_t]),
        "or_tree_new": (vp, [C.c_uint32, u32p, u32p, f64p]),
        "or_tree_free": (None, [vp]),
        "or_tree_root": (C.c_uint32, [vp]),
        "or_tree_mark_edge_updated": (None, [vp, C.c_uint32]),
        "or_tree_mark_all_edges_updated": (None, [vp]),
        "or_tree_edges_to_update": (C.c_uint32, [vp, u32p]),
        "or_tree_partials_lists": (C.c_uint32, [vp, u32p, u32p]),
        "or_tree_propagation_lists": (C.c_uint32, [vp, u32p, u32p, C.POINTER(C.c_uint32)]),
        "or_tree_set_height": (None, [vp, C.c_uint32, C.c_double]),
        "or_tree_replace_child": (C.c_int, [vp, C.c_uint32, C.c_uint32, C.c_uint32]),
        "or_tree_edge_length": (C.c_double, [vp, C.c_uint32]),
        "or_tree_is_changed": (C.c_int, [vp]),
        "or_tree_accept": (None, [vp]),
        "or_tree_reject": (None, [vp]),
        "or_tree_get": (None, [vp, u32p, u32p, f64p]),
        "or_model_jc": (None, [C.POINTER(_Model)]),
        "or_model_k80": (None, [C.POINTER(_Model), C.c_double]),
        "or_model_hky": (None, [C.POINTER(_Model), C.c_double, f64p]),
        "or_model_gtr": (None, [C.POINTER(_Model), f64p, f64p]),
        "or_model_reversible": (C.c_int, [C.POINTER(_Model), C.c_int, f64p, f64p]),
        "or_model_transition": (None, [C.POINTER(_Model), C.c_double, f64p]),
        "or_model_frequencies": (None, [C.POINTER(_Model), f64p]),
        "or_sym_eigen": (C.c_int, [C.c_int, f64p, f64p, f64p]),
        "or_gamma_p": (C.c_double, [C.c_double, C.c_double]),
        "or_gamma_inv_cdf": (C.c_double, [C.c_double, C.c_double]),
        "or_gamma_categories": (None, [C.c_double, C.c_int, f64p]),
        "or_weighted_sum": (C.c_double, [f64p, u32p, C.c_size_t]),
        "or_log_sum_exp": (C.c_double, [f64p, C.c_size_t]),
        "or_calc4_new": (vp, [C.c_size_t, u8p, C.c_uint32, C.c_uint32, C.c_int]),
        "or_calc4_free": (None, [vp]),
        "or_calc4_propose": (None, [vp, C.c_uint32, u32p, u32p, f64p, f64p]),
        "or_calc4_accept": (None, [vp]),
        "or_calc4_reject": (None, [vp]),
        "or_calc4_likelihoods": (C.POINTER(C.c_double), [vp]),
        "or_calc4_num_threads": (C.c_int, [vp]),
        "or_calcn_new": (vp, [C.c_size_t, u8p, C.c_uint32, C.c_int, C.c_uint32, C.c_int]),
        "or_calcn_free": (None, [vp]),
        "or_calcn_propose": (None, [vp, C.c_uint32, u32p, u32p, f64p, f64p]),
        "or_calcn_accept": (None, [vp]),
        "or_calcn_reject": (None, [vp]),
        "or_calcn_likelihoods": (C.POINTER(C.c_double), [vp]),
        "or_gamma_lik_new": (vp, [vp, C.POINTER(_Model), u8p, C.c_uint32, C.c_size_t, u32p,
                                  C.c_int, C.c_double, C.c_double, C.c_uint32, C.c_int, C.c_int]),
        "or_gamma_lik_free": (None, [vp]),
        "or_gamma_lik_likelihood": (C.c_double, [vp]),
        "or_gamma_lik_accept": (None, [vp]),
        "or_gamma_lik_reject": (None, [vp]),
        "or_gamma_lik_set_model": (None, [vp, C.POINTER(_Model)]),
        "or_gamma_lik_set_alpha": (None, [vp, C.c_double]),
        "or_gamma_lik_site": (C.POINTER(C.c_double), [vp, C.c_int]),
        "or_gamma_lik_rates": (C.POINTER(C.c_double), [vp]),
        "or_deduplicate": (C.c_size_t, [C.c_uint32, C.c_size_t, u8p, u8p, u32p, szp]),
        "or_prop4_new": (vp, [C.c_size_t, u8p, C.c_uint32, C.c_uint32]),
        "or_prop4_free": (None, [vp]),
        "or_prop4_propose": (None, [vp, C.c_uint32, u32p, C.c_uint32, u32p, f64p, f64p]),
        "or_prop4_accept": (None, [vp]),
        "or_prop4_reject": (None, [vp]),
        "or_prop4_likelihoods": (C.POINTER(C.c_double), [vp]),
        "or_components_eval": (C.c_double, [vp, C.c_int, vp, f64p, u8p, C.c_size_t, u32p, vp,
                                            C.c_uint32, f64p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


# --------------------------------------------------------------------------
# EditBuf (crates/sk/src/editbuf.rs)
def editbuf_set_edited(m: np.ndarray, index: int) -> None:
    lib().or_editbuf_set_edited(m, index)


def editbuf_accept(m: np.ndarray) -> None:
    lib().or_editbuf_accept(m, m.size)


def editbuf_reject(m: np.ndarray) -> None:
    lib().or_editbuf_reject(m, m.size)


# --------------------------------------------------------------------------
class Tree:
    """DatedTree update bookkeeping (crates/b3/src/parameters/tree.rs:350-528)."""

    def __init__(self, num_leaves, parents, children, heights):
        self.L = int(num_leaves)
        self.N = 2 * self.L - 1
        self._h = lib().or_tree_new(
            self.L,
            np.ascontiguousarray(parents, np.uint32),
            np.ascontiguousarray(children, np.uint32).reshape(-1),
            np.ascontiguousarray(heights, np.float64),
        )

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_tree_free(self._h)
            self._h = None

    @property
    def root(self):
        return lib().or_tree_root(self._h)

    def mark_edge_updated(self, e):
        lib().or_tree_mark_edge_updated(self._h, int(e))

    def mark_all_edges_updated(self):
        lib().or_tree_mark_all_edges_updated(self._h)

    def edges_to_update(self):
        out = np.zeros(self.N, np.uint32)
        k = lib().or_tree_edges_to_update(self._h, out)
        return out[:k].copy()

    def partials_lists(self):
        internals = np.zeros(self.N, np.uint32)
        children = np.zeros(2 * self.N, np.uint32)
        n = lib().or_tree_partials_lists(self._h, internals, children)
        return internals[:n].copy(), children[: 2 * n].reshape(n, 2).copy()

    def propagation_lists(self):
        nodes = np.zeros(2 * self.N, np.uint32)
        children = np.zeros(2 * self.N, np.uint32)
        nl = C.c_uint32(0)
        n = lib().or_tree_propagation_lists(self._h, nodes, children, C.byref(nl))
        ni = n - nl.value
        return nodes[:n].copy(), children[: 2 * ni].reshape(ni, 2).copy(), nl.value

    def set_height(self, node, h):
        lib().or_tree_set_height(self._h, int(node), float(h))

    def replace_child(self, parent, old, new):
        if lib().or_tree_replace_child(self._h, int(parent), int(old), int(new)) != 0:
            raise ValueError("replace_child: old child is not a child of parent")

    def edge_length(self, e):
        return lib().or_tree_edge_length(self._h, int(e))

    def is_changed(self):
        return bool(lib().or_tree_is_changed(self._h))

    def accept(self):
        lib().or_tree_accept(self._h)

    def reject(self):
        lib().or_tree_reject(self._h)

    def get(self):
        p = np.zeros(self.N, np.uint32)
        c = np.zeros(2 * (self.L - 1), np.uint32)
        h = np.zeros(self.N, np.float64)
        lib().or_tree_get(self._h, p, c, h)
        return p, c.reshape(-1, 2), h


# --------------------------------------------------------------------------
class Model:
    """Substitution models (crates/b3/src/substitution.rs)."""

    def __init__(self):
        self._m = _Model()

    @classmethod
    def jc(cls):
        m = cls()
        lib().or_model_jc(C.byref(m._m))
        return m

    @classmethod
    def k80(cls, kappa):
        m = cls()
        lib().or_model_k80(C.byref(m._m), float(kappa))
        return m

    @classmethod
    def hky(cls, kappa, freqs):
        m = cls()
        lib().or_model_hky(C.byref(m._m), float(kappa), np.ascontiguousarray(freqs, np.float64))
        return m

    @classmethod
    def gtr(cls, rates, freqs):
        m = cls()
        lib().or_model_gtr(C.byref(m._m), np.ascontiguousarray(rates, np.float64),
                           np.ascontiguousarray(freqs, np.float64))
        return m

    @classmethod
    def reversible(cls, R, pi):
        m = cls()
        R = np.ascontiguousarray(R, np.float64)
        rc = lib().or_model_reversible(C.byref(m._m), R.shape[0], R.reshape(-1),
                                       np.ascontiguousarray(pi, np.float64))
        if rc != 0:
            raise RuntimeError("eigendecomposition did not converge")
        return m

    @property
    def n(self):
        return self._m.n

    @property
    def q(self):
        n = self.n
        return np.ctypeslib.as_array(self._m.q)[: n * n].reshape(n, n).copy()

    @property
    def eigen(self):
        n = self.n
        V = np.ctypeslib.as_array(self._m.evec)[: n * n].reshape(n, n).copy()
        Vi = np.ctypeslib.as_array(self._m.ievc)[: n * n].reshape(n, n).copy()
        w = np.ctypeslib.as_array(self._m.eval)[:n].copy()
        return V, w, Vi

    def transition(self, t):
        out = np.zeros(self.n * self.n, np.float64)
        lib().or_model_transition(C.byref(self._m), float(t), out)
        return out.reshape(self.n, self.n)

    def frequencies(self):
        out = np.zeros(self.n, np.float64)
        lib().or_model_frequencies(C.byref(self._m), out)
        return out


def sym_eigen(A):
    A = np.ascontiguousarray(A, np.float64)
    n = A.shape[0]
    w = np.zeros(n)
    U = np.zeros((n, n))
    rc = lib().or_sym_eigen(n, A.reshape(-1), w, U.reshape(-1))
    return rc, w, U


# --------------------------------------------------------------------------
def gamma_p(a, x):
    return lib().or_gamma_p(float(a), float(x))


def gamma_inv_cdf(a, p):
    return lib().or_gamma_inv_cdf(float(a), float(p))


def gamma_categories(alpha, k):
    """crates/b3/src/likelihood/gamma.rs:205-225"""
    out = np.zeros(k)
    lib().or_gamma_categories(float(alpha), int(k), out)
    return out


def weighted_sum(values, weights):
    v = np.ascontiguousarray(values, np.float64)
    return lib().or_weighted_sum(v, np.ascontiguousarray(weights, np.uint32), v.size)


def log_sum_exp(values):
    v = np.ascontiguousarray(values, np.float64)
    return lib().or_log_sum_exp(v, v.size)


# --------------------------------------------------------------------------
class Calc4:
    """Cpu4Calculator (crates/b3/src/calculator/cpu4.rs)."""

    def __init__(self, samples, num_patterns, scale_ln=30, num_threads=1):
        s = np.ascontiguousarray(samples, np.uint8).reshape(-1)
        self.P = int(num_patterns)
        self.L = s.size // self.P
        self._h = lib().or_calc4_new(self.P, s, self.L, int(scale_ln), int(num_threads))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_calc4_free(self._h)
            self._h = None

    @property
    def num_threads(self):
        return lib().or_calc4_num_threads(self._h)

    def propose(self, internals, children, mats, freqs):
        internals = np.ascontiguousarray(internals, np.uint32)
        lib().or_calc4_propose(self._h, internals.size, internals,
                               np.ascontiguousarray(children, np.uint32).reshape(-1),
                               np.ascontiguousarray(mats, np.float64).reshape(-1),
                               np.ascontiguousarray(freqs, np.float64))

    def accept(self):
        lib().or_calc4_accept(self._h)

    def reject(self):
        lib().or_calc4_reject(self._h)

    def likelihoods(self):
        p = lib().or_calc4_likelihoods(self._h)
        return np.ctypeslib.as_array(p, shape=(self.P,)).copy()


class Prop4:
    """CudaPropagationsCalculator's algorithm on the CPU
    (crates/b3/src/calculator/cuda_propagations/kernels.cu:73-187, mod.rs:77-164)."""

    def __init__(self, samples, num_patterns, scale_ln=30):
        s = np.ascontiguousarray(samples, np.uint8).reshape(-1)
        self.P = int(num_patterns)
        self.L = s.size // self.P
        self._h = lib().or_prop4_new(self.P, s, self.L, int(scale_ln))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_prop4_free(self._h)
            self._h = None

    def propose(self, nodes, children, num_leaves, mats, freqs):
        """mats: [len(nodes) - 1][4][4], one per listed node but the root."""
        nodes = np.ascontiguousarray(nodes, np.uint32)
        lib().or_prop4_propose(self._h, nodes.size, nodes, int(num_leaves),
                               np.ascontiguousarray(children, np.uint32).reshape(-1),
                               np.ascontiguousarray(mats, np.float64).reshape(-1),
                               np.ascontiguousarray(freqs, np.float64))

    def accept(self):
        lib().or_prop4_accept(self._h)

    def reject(self):
        lib().or_prop4_reject(self._h)

    def likelihoods(self):
        p = lib().or_prop4_likelihoods(self._h)
        return np.ctypeslib.as_array(p, shape=(self.P,)).copy()


class CalcN:
    """cpu4.rs semantics for S states (no reference n-state calculator)."""

    def __init__(self, samples, num_patterns, states, scale_ln=30, num_threads=1):
        s = np.ascontiguousarray(samples, np.uint8).reshape(-1)
        self.P = int(num_patterns)
        self.L = s.size // self.P
        self.S = int(states)
        self._h = lib().or_calcn_new(self.P, s, self.L, self.S, int(scale_ln), int(num_threads))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_calcn_free(self._h)
            self._h = None

    def propose(self, internals, children, mats, freqs):
        internals = np.ascontiguousarray(internals, np.uint32)
        lib().or_calcn_propose(self._h, internals.size, internals,
                               np.ascontiguousarray(children, np.uint32).reshape(-1),
                               np.ascontiguousarray(mats, np.float64).reshape(-1),
                               np.ascontiguousarray(freqs, np.float64))

    def accept(self):
        lib().or_calcn_accept(self._h)

    def reject(self):
        lib().or_calcn_reject(self._h)

    def likelihoods(self):
        p = lib().or_calcn_likelihoods(self._h)
        return np.ctypeslib.as_array(p, shape=(self.P,)).copy()


class GammaLikelihood:
    """GammaLikelihood / DnaLikelihood (crates/b3/src/likelihood/gamma.rs:97-203,
    dna.rs:68-124) over K Cpu4Calculators.  K == 1 is DnaLikelihood."""

    def __init__(self, tree: Tree, model: Model, codes, num_patterns, weights,
                 num_categories=4, alpha=0.5, clock_rate=1.0, scale_ln=30,
                 threads_per_category=1, parallel_categories=False):
        codes = np.ascontiguousarray(codes, np.uint8).reshape(-1)
        self.P = int(num_patterns)
        self.K = int(num_categories)
        self.tree = tree
        self._model = model
        self._h = lib().or_gamma_lik_new(
            tree._h, C.byref(model._m), codes, tree.L, self.P,
            np.ascontiguousarray(weights, np.uint32), self.K, float(alpha),
            float(clock_rate), int(scale_ln), int(threads_per_category),
            int(bool(parallel_categories)))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_gamma_lik_free(self._h)
            self._h = None

    def likelihood(self):
        return lib().or_gamma_lik_likelihood(self._h)

    def accept(self):
        lib().or_gamma_lik_accept(self._h)

    def reject(self):
        lib().or_gamma_lik_reject(self._h)

    def set_model(self, model: Model):
        self._model = model
        lib().or_gamma_lik_set_model(self._h, C.byref(model._m))

    def set_alpha(self, alpha):
        lib().or_gamma_lik_set_alpha(self._h, float(alpha))

    def site(self, c):
        p = lib().or_gamma_lik_site(self._h, int(c))
        return np.ctypeslib.as_array(p, shape=(self.P,)).copy()

    def rates(self):
        p = lib().or_gamma_lik_rates(self._h)
        return np.ctypeslib.as_array(p, shape=(self.K,)).copy()


def components_eval(tree: Tree, models, edge_rates, codes, weights, classes=None, scale_ln=30):
    """Per-category models and per-edge clock rates from scratch: HeteroLikelihood's
    class-select total (crates/b3/src/likelihood/hetero.rs:122-160) when
    ``classes`` is given, else GammaLikelihood's mix (gamma.rs:126-166).
    edge_rates: [K][2L-1] (clock rate of each edge, times the category rate for
    the gamma mix).  Returns (total, site [K][P])."""
    K = len(models)
    arr = (_Model * K)()
    for c, m in enumerate(models):
        arr[c] = m._m
    codes = np.ascontiguousarray(codes, np.uint8)
    P = codes.shape[1]
    er = np.ascontiguousarray(edge_rates, np.float64).reshape(K, tree.N)
    cls = None if classes is None else np.ascontiguousarray(classes, np.uint32)
    site = np.zeros(K * P, np.float64)
    tot = lib().or_components_eval(tree._h, K, C.cast(arr, C.c_void_p), er.reshape(-1),
                                   codes.reshape(-1), P, np.ascontiguousarray(weights, np.uint32),
                                   None if cls is None else cls.ctypes.data_as(C.c_void_p),
                                   int(scale_ln), site)
    return tot, site.reshape(K, P)


def deduplicate(msa):
    """crates/b3/src/likelihood/mod.rs:74-132.  msa: [num_seqs, num_sites] u8."""
    msa = np.ascontiguousarray(msa, np.uint8)
    ns, nsites = msa.shape
    leaves = np.zeros(ns * nsites, np.uint8)
    weights = np.zeros(nsites, np.uint32)
    s2p = np.zeros(nsites, np.uintp)
    P = lib().or_deduplicate(ns, nsites, msa.reshape(-1), leaves, weights, s2p)
    return leaves[: ns * P].reshape(ns, P).copy(), weights[:P].copy(), s2p.astype(np.int64)
