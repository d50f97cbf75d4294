"""MI355X-native residual-and-update path of OpenSTC (HIP kernels behind a C-ABI).

The compute lives in ``openstc_amd/libostc.so`` (built from
``openstc_amd/csrc``); this package holds the ctypes binding, the Python
time-loop handle and the synthetic workload generators.
"""
from . import _lib  # noqa: F401

__all__ = ["_lib"]
