"""gowalker_amd: MI355X-native second-order random-walk engine (Node2vec, SOPR, SOSR).

The compute lives in libgw.so (HIP kernels for gfx950 behind the C ABI of
include/gw.h); `gowalker_amd.engine` is a thin numpy/ctypes view of that ABI.
"""
from ._abi import LIB_PATH, GwError, lib  # noqa: F401
