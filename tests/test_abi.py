"""The C ABI: libumi_hip.so loads and exports every function include/*.h declares.

No compute call is made here (no GPU needed).
"""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"\b(umi_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "umi_me_full_search_u8" in names and "umi_malloc" in names


def test_library_exports_every_declared_symbol():
    from umi_amd._lib import LIB_PATH, SIGNATURES

    assert LIB_PATH.exists(), "libumi_hip.so not built"
    lib = ctypes.CDLL(str(LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"
    unbound = [n for n in declared_functions() if n not in SIGNATURES and n != "umi_last_error"]
    assert not unbound, f"declared but not bound in umi_amd._lib: {unbound}"


def test_errors_are_status_codes_not_aborts():
    # argument validation happens before any device call
    from umi_amd._lib import lib

    L = lib()
    assert L.umi_me_full_search_u8(None, None, 8, 8, 8, 8, 4, None, None, None) == 1
    assert b"null" in L.umi_last_error()
    assert L.umi_me_full_search_u8(1, 1, 64, 64, 64, 12, 4, 1, None, None) == 3
    assert L.umi_cost_volume_u8(1, 1, 8, 8, 8, 4, 0, 2, 1, 1, 1, None) == 1
