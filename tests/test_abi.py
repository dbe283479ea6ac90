"""The C-ABI library loads and exports every entry point that
include/locarna_amd.h declares (no GPU needed)."""
import ctypes
import os
import re

from locarna_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "locarna_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lcg_[a-z0-9_]+)\s*\(", text)))


def test_all_declared_symbols_exported():
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_prepare_reports_malformed_input():
    from locarna_amd import api
    import pytest
    with pytest.raises(api.LcgError, match="-2"):
        api.Pair("not a pp file", "#PP 2.0\n", api.locarna_params())
    bad = "#PP 2.0\n\nx ACGUACGU\n\n#END\n\n#SECTION BASEPAIRS\n\n1 20 0.5\n\n#END\n"
    with pytest.raises(api.LcgError, match="Invalid indices"):
        api.Pair(bad, bad, api.locarna_params())
