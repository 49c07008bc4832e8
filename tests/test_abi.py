"""The C-ABI library: builds, loads and exports every symbol include/gral_pf.h declares."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "gral_pf.h")).read()
    return sorted(set(re.findall(r"\b(gral_pf_[a-z_]+)\s*\(", src)))


def test_header_declares_expected_entry_points():
    syms = declared_symbols()
    for s in ("gral_pf_create", "gral_pf_solve", "gral_pf_iterate", "gral_pf_download_fields", "gral_pf_destroy"):
        assert s in syms


def test_library_exports_all_declared_symbols():
    from gral_amd import pf
    if not os.path.exists(pf.LIB_PATH):
        pytest.skip("libgral_pf.so not built")
    import ctypes
    lib = ctypes.CDLL(pf.LIB_PATH)
    for s in declared_symbols():
        assert hasattr(lib, s), s
    assert sorted(pf.EXPORTS) == declared_symbols()
