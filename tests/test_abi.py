"""The C-ABI library loads and exports every symbol its headers declare
(no GPU needed: nothing is called)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(bh_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_declared_symbols():
    import bem3_emf_b2_amd as bem
    if not os.path.exists(bem.LIB_PATH):
        pytest.skip("library not built")
    L = ctypes.CDLL(bem.LIB_PATH)
    names = _declared("bem3_hip.h")
    assert len(names) > 5
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
