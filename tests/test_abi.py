"""The C-ABI library loads and exports every symbol include/b3hip.h declares;
argument errors come back as statuses (no GPU compute needed)."""
import ctypes as C
import re
import subprocess

import numpy as np
import pytest

from aspartik_amd import _lib


def _declared(path=None, prefix="b3h"):
    text = open(path or _lib.HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    names = _declared()
    for must in ["b3h_create", "b3h_destroy", "b3h_propose", "b3h_propose_times", "b3h_accept",
                 "b3h_reject", "b3h_likelihoods", "b3h_total", "b3h_num_patterns",
                 "b3h_last_error", "b3h_set_model"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (b3h_\w+)", out))
    for name in _declared():
        assert name in exported, name
        getattr(lib, name)
    assert set(_lib.SIGNATURES) == set(_declared())
    mcmc = _declared(_lib.MCMC_HEADER_PATH, "b3m")
    exported_m = set(re.findall(r"\bT (b3m_\w+)", out))
    for name in mcmc:
        assert name in exported_m, name
        getattr(lib, name)
    assert set(_lib.MCMC_SIGNATURES) == set(mcmc)
    data = _declared(_lib.DATA_HEADER_PATH, "b3d")
    exported_d = set(re.findall(r"\bT (b3d_\w+)", out))
    for name in data:
        assert name in exported_d, name
        getattr(lib, name)
    assert set(_lib.DATA_SIGNATURES) == set(data)


def test_mcmc_struct_layouts_match_the_header():
    """ctypes mirrors of b3m_chain_config / b3m_run_stats have the C sizes"""
    src = _lib.MCMC_HEADER_PATH
    prog = ("#include <stdio.h>\n#include \"" + src + "\"\n"
            "int main(void){printf(\"%zu %zu\\n\", sizeof(b3m_chain_config),"
            " sizeof(b3m_run_stats));return 0;}\n")
    import os
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(prog)
        exe = os.path.join(d, "t")
        subprocess.run(["gcc", "-I", os.path.dirname(src), "-o", exe, c], check=True)
        a, b = map(int, subprocess.run([exe], capture_output=True, text=True,
                                       check=True).stdout.split())
    assert a == C.sizeof(_lib.ChainConfig)
    assert b == C.sizeof(_lib.RunStats)


def test_create_reports_missing_device_as_status():
    lib = _lib.lib()
    cfg = _lib.Config()
    lib.b3h_config_default(C.byref(cfg))
    cfg.device = _lib.device_count() + 7
    h = C.c_void_p()
    st = lib.b3h_create(C.byref(cfg), np.ones(8, np.uint8), 2, 4, 1, C.byref(h))
    assert st == 2  # B3H_ERR_NO_DEVICE
    assert h.value is None
    assert lib.b3h_last_error()


@pytest.mark.parametrize("kwargs,status", [
    (dict(L=1), 1), (dict(P=0), 1), (dict(K=0), 1), (dict(K=17), 1),
    (dict(scale_ln=0), 1), (dict(states=7), 6), (dict(precision=9), 6),
])
def test_create_argument_errors(kwargs, status):
    lib = _lib.lib()
    cfg = _lib.Config()
    lib.b3h_config_default(C.byref(cfg))
    cfg.scale_ln = kwargs.get("scale_ln", 30)
    cfg.states = kwargs.get("states", 4)
    cfg.precision = kwargs.get("precision", 0)
    L, P, K = kwargs.get("L", 3), kwargs.get("P", 5), kwargs.get("K", 1)
    h = C.c_void_p()
    st = lib.b3h_create(C.byref(cfg), np.ones(max(1, L * P), np.uint8), L, P, K, C.byref(h))
    assert st == status
    assert h.value is None


def test_null_handle_is_an_error_not_a_crash():
    lib = _lib.lib()
    assert lib.b3h_accept(None) == 1
    assert lib.b3h_reject(None) == 1
    assert lib.b3h_num_patterns(None) == 0
    v = C.c_double()
    assert lib.b3h_total(None, C.byref(v)) == 1
    assert lib.b3h_destroy(None) == 0
