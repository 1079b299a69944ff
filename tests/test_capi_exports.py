"""The C-ABI library loads and exports every entry point include/*.h declares (CPU only)."""
import ctypes
import re
from pathlib import Path

from ahead_amd import an, capi

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(ahead_(?:hip|serve)_\w+)\s*\(", text))
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


def test_noalloc_select_rejects_columns_beyond_its_scratch():
    # one 2048-row tile per 16 KiB sub-tile (tiling 1 x 4096): 400M 64-bit codes need ~195k tiles,
    # more than the 4 MiB static scratch holds; the call is refused on the host, before any device
    # work, so the column pointer is never dereferenced
    lib = capi.load()
    old = capi.get_tiling()
    capi.set_tiling(1, 4096)
    try:
        meta = an.an_meta(an.RESINT, 65117)
        d = capi.AheadCol(0x1000, 400_000_000, 8, 1, 32, 0, meta.a, meta.a_inv, meta.type.max_u, 0)
        st = lib.ahead_hip_select_an_noalloc(ctypes.byref(d), capi.GE, 0, capi.LE, 1, capi.COMB_AND, 1, 1, 0x1000,
                                             0, None, 0, 0x1000, None, None)
        assert st == 5  # AHEAD_ERR_TOO_LARGE
        assert b"static scratch" in lib.ahead_hip_last_error()
    finally:
        capi.set_tiling(*old)


def test_status_constants_match_the_header():
    import re
    from pathlib import Path

    from ahead_amd import capi

    h = (Path(__file__).resolve().parent.parent / "include" / "ahead_hip.h").read_text()
    for name in ("AHEAD_ERR_INVALID", "AHEAD_ERR_SIZE_MISMATCH", "AHEAD_ERR_HIP", "AHEAD_ERR_NOMEM",
                 "AHEAD_ERR_TOO_LARGE", "AHEAD_ERR_DEVICE"):
        m = re.search(rf"{name}\s*=\s*(\d+)", h)
        assert m and int(m.group(1)) == getattr(capi, name), name
