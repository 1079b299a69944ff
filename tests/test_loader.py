"""AHEAD's table format (header CSV + raw .ahead column files, lib/column_storage/Transaction.cpp):
the reference's own SSB headers (tests/golden/headers, from the reference's headers.tgz) parse to
the expected column types, and tables written in the format read back unchanged."""
from pathlib import Path

import numpy as np
import pytest

from ahead_amd import an, datagen, loader

HEADERS = Path(__file__).resolve().parent / "golden" / "headers"


def test_reference_headers_parse():
    tables = sorted(p.name[: -len("_header.csv")] for p in HEADERS.glob("*_header.csv"))
    assert len(tables) == 10
    for tbl in tables:
        cols = loader.read_header(HEADERS / tbl)
        assert len(cols) >= 7
        for name, typ, t in cols:
            if typ.startswith("RES"):
                assert t is not None and t.encoded
            elif typ.startswith(("STRING", "CHAR", "FIXED")):
                assert t is None
            else:
                assert t is not None and not t.encoded
    lo = {n: (typ, t) for n, typ, t in loader.read_header(HEADERS / "lineorderAN")}
    assert lo["quantity"] == ("RESTINY", an.RESTINY)
    assert lo["extendedprice"] == ("RESINT", an.RESINT)
    assert lo["orderpriority"][1] is None
    d = {n: t for n, _, t in loader.read_header(HEADERS / "dateAN")}
    assert d["year"] is an.RESSHORT and d["datekey"] is an.RESINT


def test_unknown_type_raises(tmp_path):
    (tmp_path / "t_header.csv").write_text("a|b\nRESINT|WIDGET\n")
    with pytest.raises(ValueError):
        loader.read_header(tmp_path / "t")


def test_write_read_roundtrip(tmp_path):
    n = 10_000
    q = datagen.encode(datagen.uniform(n, 1, 1, 1, 50), 233, 2)
    p = datagen.encode(datagen.uniform(n, 1, 2, 0, 10**7), 65117, 8)
    k = datagen.uniform(n, 1, 3, 0, 2**31).astype(np.uint32)
    loader.write_table(tmp_path / "lo", {"quantity": ("RESTINY", q), "price": ("RESINT", p), "key": ("INTEGER", k)})
    cols = loader.read_header(tmp_path / "lo")
    assert [c[1] for c in cols] == ["RESTINY", "RESINT", "INTEGER"]
    assert (tmp_path / "lo_price.ahead").stat().st_size == 8 * n
    assert np.array_equal(loader.read_column(tmp_path / "lo", "quantity", an.RESTINY), q)
    assert np.array_equal(loader.read_column(tmp_path / "lo", "price", an.RESINT), p)
    assert np.array_equal(loader.read_column(tmp_path / "lo", "key", an.INT), k)
    m = loader.column_meta(an.RESSHORT)
    assert m.a == 63877 and (m.a * m.a_inv) % (1 << 32) == 1


def test_string_columns_are_written_but_not_loaded(tmp_path):
    n = 100
    s = np.frombuffer(b"".join(b"abcde\0" for _ in range(n)), dtype=np.uint8)
    loader.write_table(tmp_path / "t", {"name": ("STRING:5", s), "k": ("RESINT", np.arange(n, dtype=np.uint64) * 65117)})
    assert (tmp_path / "t_name.ahead").stat().st_size == 6 * n
    cols = {c[0]: c[2] for c in loader.read_header(tmp_path / "t")}
    assert cols["name"] is None and cols["k"] is an.RESINT


def test_bare_string_defaults_to_maxlen_64(tmp_path):
    # Transaction.cpp:249-255: STRING without ":len" takes MAXLEN_STRING = 64 (+1 terminator) bytes per row
    n = 7
    s = np.zeros(65 * n, dtype=np.uint8)
    loader.write_table(tmp_path / "t", {"name": ("STRING", s), "k": ("RESINT", np.arange(n, dtype=np.uint64) * 65117)})
    assert (tmp_path / "t_name.ahead").stat().st_size == 65 * n
    assert loader._type_of("STRING") == (None, 65)
    assert loader._type_of("STRING:25") == (None, 26)


def test_resbigint_is_refused_at_load(tmp_path):
    # Transaction.cpp:378-381: RESBIGINT columns are listed by the header but not loaded
    n = 4
    loader.write_table(tmp_path / "t", {"v": ("RESBIGINT", np.arange(n, dtype=np.uint64) * 65117)})
    cols = loader.read_header(tmp_path / "t")
    assert cols[0][2] is an.RESBIGINT
    with pytest.raises(ValueError, match="unsupported for load"):
        loader.load_table(tmp_path / "t")
