"""aspartik_amd: MI355X-native Felsenstein-pruning likelihood for b3.

The compute path is the HIP library ``libb3hip.so`` (C ABI in
include/b3hip.h).  ``aspartik_amd.b3`` mirrors the b3 Python surface that
drives it (calculator configuration, dated tree update lists, substitution
models, DNA and gamma likelihoods).
"""
from ._lib import B3hError, device_count, lib  # noqa: F401

__version__ = "0.1.0"

__all__ = ["B3hError", "device_count", "lib", "__version__"]
