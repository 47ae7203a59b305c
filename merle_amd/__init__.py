"""merle_amd -- MI355X-native fused bitmap-index query path.

The compute path is libmbx.so (hand-written HIP for gfx950 behind the C ABI in
include/mbx.h).  This package only loads it (ctypes) and offers thin host
helpers used by the tests, the benchmark and the examples.  There is no
Python or CPU fallback: if the library is missing, loading fails loudly.
"""
import ctypes as _C
import os as _os

from . import abi

__all__ = ["lib", "load", "abi", "LibraryMissing"]

_lib = None


class LibraryMissing(RuntimeError):
    pass


def load():
    """Load libmbx.so (built in-tree by merle_amd/csrc/Makefile)."""
    global _lib
    if _lib is not None:
        return _lib
    if not _os.path.exists(abi.LIB_PATH):
        raise LibraryMissing(
            f"{abi.LIB_PATH} not found: build it with "
            "`make -C merle_amd/csrc` (or __graft_entry__.build())")
    lib = _C.CDLL(abi.LIB_PATH)
    missing = abi.bind(lib)
    if missing:
        raise LibraryMissing(f"libmbx.so lacks exports: {missing}")
    abi.check_layout(lib)
    _lib = lib
    return lib


def lib():
    return load()
