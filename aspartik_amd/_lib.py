"""Loader of the in-tree HIP library ``libb3hip.so`` (C ABI: include/b3hip.h).

There is no CPU fallback: if the library is missing or cannot be loaded,
:func:`lib` raises.  Build it with ``__graft_entry__.build()`` or
``make -C aspartik_amd/csrc``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libb3hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "b3hip.h")
MCMC_HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "b3mcmc.h")
DATA_HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "b3data.h")

B3H_OK = 0
STATUS_NAMES = {
    0: "B3H_OK",
    1: "B3H_ERR_ARG",
    2: "B3H_ERR_NO_DEVICE",
    3: "B3H_ERR_OOM",
    4: "B3H_ERR_HIP",
    5: "B3H_ERR_STATE",
    6: "B3H_ERR_UNSUPPORTED",
}

B3H_F32_SCALED = 0
B3H_F64_PLAIN = 1
B3H_F16_MFMA = 2
B3H_F64_PROPAGATION = 3
B3H_F32_PROPAGATION = 4
B3H_MODEL_SPECIAL_CASES = 1


class B3hError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class Config(C.Structure):
    _fields_ = [
        ("device", C.c_int32),
        ("precision", C.c_uint32),
        ("states", C.c_uint32),
        ("scale_ln", C.c_uint32),
        ("tile_patterns", C.c_uint32),
        ("lane_patterns", C.c_uint32),
        ("lane_states", C.c_uint32),
        ("reserved", C.c_uint32 * 1),
    ]


class Instance(C.Structure):
    _fields_ = [
        ("samples", C.c_void_p),
        ("num_leaves", C.c_uint32),
        ("num_patterns", C.c_uint32),
        ("weights", C.c_void_p),
    ]


class Update(C.Structure):
    _fields_ = [
        ("n", C.c_uint32),
        ("internals", C.c_void_p),
        ("children", C.c_void_p),
        ("edge_times", C.c_void_p),
        ("category_rates", C.c_void_p),
    ]


class ChainConfig(C.Structure):
    """b3m_chain_config (include/b3mcmc.h)"""
    _fields_ = [
        ("device", C.c_int32),
        ("categories", C.c_uint32),
        ("seed", C.c_uint64),
        ("birth_rate", C.c_double),
        ("alpha_prior_rate", C.c_double),
        ("w_node_slide", C.c_double),
        ("w_tree_scale", C.c_double),
        ("w_rates", C.c_double),
        ("w_alpha", C.c_double),
        ("w_narrow", C.c_double),
        ("scale_factor", C.c_double),
        ("alpha_factor", C.c_double),
        ("delta_tuning", C.c_double),
        ("tune_every", C.c_uint32),
        ("tune_cutoff", C.c_uint64),
        ("tune_magnitude", C.c_double),
        ("full_evaluations", C.c_uint32), ("screen_topology", C.c_uint32),
        ("surrogate_beta", C.c_double),
        ("narrow_reversible", C.c_uint32),
    ]


class RunStats(C.Structure):
    """b3m_run_stats (include/b3mcmc.h)"""
    _fields_ = [
        ("steps", C.c_uint64), ("accepted", C.c_uint64), ("exact_evals", C.c_uint64),
        ("surrogate_evals", C.c_uint64), ("stage1_rejects", C.c_uint64),
        ("stage2_rejects", C.c_uint64), ("aborts", C.c_uint64), ("prior_rejects", C.c_uint64),
        ("op_proposed", C.c_uint64 * 5), ("op_accepted", C.c_uint64 * 5),
        ("collected", C.c_uint64), ("seconds", C.c_double),
        ("likelihood_seconds", C.c_double), ("surrogate_seconds", C.c_double),
        ("tuning", C.c_double * 5),
    ]


B3M_PLAIN = 0
B3M_DELAYED = 1
B3M_TRACE_COLS = 13
TRACE_COLUMNS = ["step", "posterior", "likelihood", "prior", "root_height", "tree_length",
                 "alpha", "rate_ac", "rate_ag", "rate_at", "rate_cg", "rate_ct", "rate_gt"]

B3H_BATCH_SINGLE_SLOT = 1
B3H_BATCH_GRAPH = 1

u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")

_lib = None

def addr(a) -> int:
    """Address of the first element of a C-contiguous array, for the
    pointer arguments of the per-evaluation calls (declared c_void_p: numpy's
    ndpointer check costs microseconds per argument).  The caller converts
    dtype and layout and keeps the array alive across the call."""
    try:
        return C.addressof(C.c_char.from_buffer(a))
    except (TypeError, ValueError):  # read-only or empty
        return a.ctypes.data


# name -> (restype, argtypes); every symbol declared in include/b3hip.h
SIGNATURES = {
    "b3h_config_default": (None, [C.POINTER(Config)]),
    "b3h_create": (C.c_int, [C.POINTER(Config), u8p, C.c_uint32, C.c_uint32, C.c_uint32,
                             C.POINTER(C.c_void_p)]),
    "b3h_destroy": (C.c_int, [C.c_void_p]),
    "b3h_set_weights": (C.c_int, [C.c_void_p, u32p]),
    "b3h_propose": (C.c_int, [C.c_void_p, C.c_uint32] + [C.c_void_p] * 4),
    "b3h_set_model": (C.c_int, [C.c_void_p, f64p, f64p, C.c_uint32]),
    "b3h_propose_times": (C.c_int, [C.c_void_p, C.c_uint32] + [C.c_void_p] * 4),
    "b3h_propose_propagations": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32]
                                 + [C.c_void_p] * 3),
    "b3h_propose_propagations_times": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p,
                                                 C.c_uint32] + [C.c_void_p] * 3),
    "b3h_set_category_model": (C.c_int, [C.c_void_p, C.c_uint32, f64p, f64p, C.c_uint32]),
    "b3h_propose_lengths": (C.c_int, [C.c_void_p, C.c_uint32] + [C.c_void_p] * 3),
    "b3h_set_classes": (C.c_int, [C.c_void_p, C.c_void_p]),
    "b3h_accept": (C.c_int, [C.c_void_p]),
    "b3h_reject": (C.c_int, [C.c_void_p]),
    "b3h_likelihoods": (C.c_int, [C.c_void_p, C.c_uint32, f64p]),
    "b3h_site_mixed": (C.c_int, [C.c_void_p, f64p]),
    "b3h_total": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "b3h_num_patterns": (C.c_uint32, [C.c_void_p]),
    "b3h_last_error": (C.c_char_p, []),
    "b3h_device_count": (C.c_int, []),
    "b3h_profile_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "b3h_profile_read": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "b3h_profile_launch_times": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64,
                                           C.POINTER(C.c_uint64)]),
    "b3h_device_bytes": (C.c_uint64, [C.c_void_p]),
    "b3h_gamma_rates": (C.c_int, [C.c_double, C.c_uint32, f64p]),
    "b3h_batch_create": (C.c_int, [C.POINTER(Config), C.c_uint32, C.POINTER(Instance), C.c_uint32,
                                   C.c_uint32, C.POINTER(C.c_void_p)]),
    "b3h_batch_destroy": (C.c_int, [C.c_void_p]),
    "b3h_batch_set_model": (C.c_int, [C.c_void_p, C.c_uint32, f64p, f64p, C.c_uint32]),
    "b3h_batch_propose_times": (C.c_int, [C.c_void_p, C.POINTER(Update), C.c_uint32]),
    "b3h_batch_totals": (C.c_int, [C.c_void_p, f64p]),
    "b3h_batch_likelihoods": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, f64p]),
    "b3h_batch_accept": (C.c_int, [C.c_void_p]),
    "b3h_batch_reject": (C.c_int, [C.c_void_p]),
    "b3h_batch_profile_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "b3h_batch_profile_read": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "b3h_batch_device_bytes": (C.c_uint64, [C.c_void_p]),
}

# every symbol declared in include/b3mcmc.h
MCMC_SIGNATURES = {
    "b3m_ess": (C.c_double, [f64p, C.c_size_t]),
    "b3m_da_stage1": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                C.c_double]),
    "b3m_da_stage2": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double]),
    "b3m_mlp_create": (C.c_int, [C.c_int, C.c_uint32, C.c_uint32, C.c_uint64,
                                 C.POINTER(C.c_void_p)]),
    "b3m_mlp_destroy": (C.c_int, [C.c_void_p]),
    "b3m_mlp_set_scaling": (C.c_int, [C.c_void_p, f32p, f32p, C.c_float, C.c_float]),
    "b3m_mlp_forward": (C.c_int, [C.c_void_p, f32p, C.c_uint32, f32p]),
    "b3m_mlp_num_params": (C.c_size_t, [C.c_void_p]),
    "b3m_mlp_loss_grad": (C.c_int, [C.c_void_p, f32p, f32p, C.c_uint32, C.POINTER(C.c_float),
                                    f32p]),
    "b3m_mlp_get_params": (C.c_int, [C.c_void_p, f32p]),
    "b3m_mlp_set_params": (C.c_int, [C.c_void_p, f32p]),
    "b3m_mlp_train": (C.c_int, [C.c_void_p, f32p, f32p, C.c_uint32, C.c_uint32, C.c_uint32,
                                C.c_float, C.POINTER(C.c_float)]),
    "b3m_chain_config_default": (None, [C.POINTER(ChainConfig)]),
    "b3m_chain_create": (C.c_int, [C.POINTER(ChainConfig), u8p, C.c_uint32, C.c_uint32, C.c_void_p,
                                   u32p, u32p, f64p, f64p, f64p, C.c_double,
                                   C.POINTER(C.c_void_p)]),
    "b3m_chain_destroy": (C.c_int, [C.c_void_p]),
    "b3m_chain_num_features": (C.c_uint32, [C.c_void_p]),
    "b3m_chain_features": (C.c_int, [C.c_void_p, f32p]),
    "b3m_chain_run": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_void_p,
                                C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(RunStats)]),
    "b3m_chain_set_surrogate": (C.c_int, [C.c_void_p, C.c_void_p]),
    "b3m_chain_params": (C.c_int, [C.c_void_p, u32p, u32p, f64p, f64p, C.POINTER(C.c_double)]),
    "b3m_chain_dump_state": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t,
                                       C.POINTER(C.c_size_t)]),
    "b3m_chain_load_state": (C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t]),
    "b3m_chain_state": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                  C.POINTER(C.c_double)]),
}


# every symbol declared in include/b3data.h
DATA_SIGNATURES = {
    "b3d_msa_from_fasta": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    "b3d_msa_from_fasta_file": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "b3d_msa_destroy": (C.c_int, [C.c_void_p]),
    "b3d_msa_num_sequences": (C.c_uint32, [C.c_void_p]),
    "b3d_msa_num_sites": (C.c_uint64, [C.c_void_p]),
    "b3d_msa_name": (C.c_char_p, [C.c_void_p, C.c_uint32]),
    "b3d_msa_data": (C.POINTER(C.c_uint8), [C.c_void_p]),
    "b3d_msa_base_frequencies": (C.c_int, [C.c_void_p, f64p]),
    "b3d_deduplicate": (C.c_int, [u8p, C.c_uint32, C.c_uint64, C.c_int, u8p, C.c_uint64, u32p,
                                  np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS"),
                                  C.POINTER(C.c_uint64)]),
    "b3d_deduplicate_noalloc": (C.c_int, [u8p, C.c_uint32, C.c_uint64, C.c_int, u8p, C.c_uint64,
                                          u32p,
                                          np.ctypeslib.ndpointer(np.uint64,
                                                                 flags="C_CONTIGUOUS"),
                                          C.POINTER(C.c_uint64)]),
}


def lib():
    """Load libb3hip.so; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is not built; run __graft_entry__.build() "
            "(the HIP library is required: there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in {**SIGNATURES, **MCMC_SIGNATURES, **DATA_SIGNATURES}.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(status: int) -> None:
    if status != B3H_OK:
        msg = lib().b3h_last_error()
        raise B3hError(status, msg.decode() if msg else "")


def device_count() -> int:
    return int(lib().b3h_device_count())
