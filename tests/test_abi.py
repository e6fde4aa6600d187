"""The C-ABI library loads and exports every entry point include/ela.h declares (no GPU calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "ela.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|size_t|const char\*)\s+(ela_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "ela_fapply" in names and "ela_solve" in names and len(names) >= 25


def test_library_exports_every_declared_symbol():
    import anisotropicfracture_amd as ela
    if not os.path.exists(ela.LIB_PATH):
        pytest.fail("libela.so not built: run `python -m anisotropicfracture_amd.build`")
    L = ctypes.CDLL(ela.LIB_PATH)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing
    # every declared symbol has a ctypes signature in the Python binding
    assert not [n for n in declared() if n not in ela.SIGNATURES]
