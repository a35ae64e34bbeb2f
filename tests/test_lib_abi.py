"""The C-ABI library loads and exports every symbol include/matcha_amd.h declares (CPU only)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def header_symbols():
    text = (ROOT / "include" / "matcha_amd.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(mch_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    syms = header_symbols()
    assert "mch_search_orientations" in syms and len(syms) >= 8


def test_library_exports_all_header_symbols():
    lib_path = ROOT / "matcha_amd" / "libmatcha_amd.so"
    if not lib_path.exists():
        pytest.fail("libmatcha_amd.so not built (run python -m matcha_amd.build)")
    lib = ctypes.CDLL(str(lib_path))
    missing = [s for s in header_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_header():
    from matcha_amd import _lib
    assert set(header_symbols()) <= set(_lib.exported_symbols())
