"""The C-ABI library loads and exports every entry point include/*.h declares (CPU only)."""
import ctypes
import re
from pathlib import Path

from ahead_amd import capi

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(ahead_hip_\w+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "ahead_hip_select_an" in names
    assert len(names) >= 10


def test_library_loads_and_exports_every_declared_symbol():
    lib = capi.load()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_signatures_cover_the_header():
    # every declared entry point has an explicit ctypes signature in the binding
    assert set(declared_symbols()) <= set(capi._SIGS)


def test_status_and_version_without_gpu():
    lib = capi.load()
    assert b"gfx950" in lib.ahead_hip_version()
    # a NULL column is rejected on the host before any device work
    st = lib.ahead_hip_select_an(None, 0, 0, 0, 0, 0, 1, 1, None, 0, None, 0, None, None, None, None)
    assert st == 1
    assert b"NULL" in lib.ahead_hip_last_error()


def test_column_struct_layout_matches_header():
    assert ctypes.sizeof(capi.AheadCol) == 64
