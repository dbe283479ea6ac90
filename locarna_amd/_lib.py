"""Loader for the native library (ctypes).  There is no fallback: if the
library has not been built, loading fails loudly."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# LOCARNA_AMD_LIB selects another build of the same library (the sanitizer
# build of locarna_amd/build.py)
LIB_PATH = os.environ.get("LOCARNA_AMD_LIB") or os.path.join(_HERE, "_build", "liblocarna_amd.so")
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("liblocarna_amd.so not built; run __graft_entry__.build()")
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib
