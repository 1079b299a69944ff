"""AHEAD's on-disk table format, loaded straight into device-resident BATs.

A table `<prefix>` is a header file `<prefix>_header.csv` and one raw file per column,
`<prefix>_<column>.ahead` (lib/column_storage/Transaction.cpp:108-112, 291-300, 407-409):
  line 1 of the header: column names, '|'-separated;
  line 2: column types -- INTEGER/INT, TINYINT, SHORTINT, LARGEINT, STRING[:len], FIXED, CHAR,
          RESTINY, RESSHORT, RESINT, RESBIGINT (Transaction.cpp:233-280);
  a column file: the column's values as little-endian fixed-width integers, no header
          (ColumnManager.cpp:283-320); AN-encoded types hold their codes.
AN-encoded columns are created with the largest A of their type's table (Transaction.cpp:356-376:
restiny 233, resshort 63877, resint 65117).  With a minimum bit-flip weight the codes are
re-encoded while loading to the type's AsBFW[min_bfw] (Transaction.cpp:421-450,
ColumnManager.cpp:320-375): code * (A_old^-1 * A_new) in the code width, done on the device.
String, fixed-point and char columns are not part of the hardened integer path and are not loaded.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from . import an

MAXLEN_STRING = 64

# header type name -> (column type, bytes per value); None marks types this loader does not load
_TYPES = {
    "INTEGER": (an.INT, 4), "INT": (an.INT, 4), "TINYINT": (an.TINYINT, 1), "SHORTINT": (an.SHORTINT, 2),
    "LARGEINT": (an.BIGINT, 8), "RESTINY": (an.RESTINY, 2), "RESSHORT": (an.RESSHORT, 4),
    "RESINT": (an.RESINT, 8), "RESBIGINT": (an.RESBIGINT, 8), "FIXED": (None, 8), "CHAR": (None, 1),
}
# the A a freshly loaded AN-encoded column carries (largest of its table)
_LOAD_A = {"restiny": an.RESTINY_AS[-1], "resshort": an.RESSHORT_AS[-1], "resint": an.RESINT_AS[-1],
           "resbigint": an.RESINT_AS[-1]}
_BFW = {"restiny": an.RESTINY_AS_BFW, "resshort": an.RESSHORT_AS_BFW, "resint": an.RESINT_AS_BFW}


def _type_of(name: str):
    # the reference matches prefixes in this order (Transaction.cpp:236-276)
    for key in ("INTEGER", "TINYINT", "SHORTINT", "LARGEINT", "RESTINY", "RESSHORT", "RESINT", "RESBIGINT",
                "FIXED", "CHAR"):
        if name.startswith(key):
            return _TYPES[key]
    if name.startswith("INT"):
        return _TYPES["INT"]
    if name.startswith("STRING"):
        # STRING[:len]: len + 1 bytes per row, len defaults to MAXLEN_STRING = 64
        # (Transaction.cpp:249-255, MetaRepositoryManager.cpp:41)
        n = int(name[7:]) if len(name) > 6 and name[6] == ":" else MAXLEN_STRING
        return (None, n + 1)
    raise ValueError(f"data type {name} in header unknown")


def read_header(prefix: str | Path, delim: str = "|") -> list[tuple[str, str, an.ColType | None]]:
    """[(column name, header type, column type or None if not loadable)] of a table."""
    lines = Path(f"{prefix}_header.csv").read_text().splitlines()
    names = [x for x in lines[0].strip().split(delim) if x]
    types = [x for x in lines[1].strip().split(delim) if x]
    if len(names) != len(types):
        raise ValueError(f"{prefix}_header.csv: {len(names)} names but {len(types)} types")
    return [(n, t, _type_of(t)[0]) for n, t in zip(names, types)]


def column_meta(t: an.ColType, a: int | None = None) -> an.ColMeta:
    if not t.encoded:
        return an.plain_meta(t)
    a = _LOAD_A[t.name] if a is None else a
    return an.an_meta(t, a)


def read_column(prefix: str | Path, name: str, t: an.ColType) -> np.ndarray:
    raw = np.fromfile(f"{prefix}_{name}.ahead", dtype=np.dtype(f"<u{t.width}"))
    return raw.astype(np.dtype(f"u{t.width}"), copy=False)


def write_table(prefix: str | Path, columns: dict[str, tuple[str, np.ndarray]], delim: str = "|") -> None:
    """Writes a table in the same format: columns = {name: (header type, values or codes)}."""
    names = list(columns)
    Path(f"{prefix}_header.csv").write_text(delim.join(names) + "\n" + delim.join(columns[n][0] for n in names) + "\n")
    for n in names:
        typ, arr = columns[n]
        t, w = _type_of(typ)
        if w not in (1, 2, 4, 8):  # fixed-width strings: rows of w bytes
            np.ascontiguousarray(arr, dtype=np.uint8).reshape(-1, w).tofile(f"{prefix}_{n}.ahead")
            continue
        np.ascontiguousarray(arr).astype(np.dtype(f"<u{w}")).tofile(f"{prefix}_{n}.ahead")


def load_table(prefix: str | Path, columns: list[str] | None = None, min_bfw: dict[str, int] | None = None) -> dict:
    """Loads the integer columns of a table into device BATs (<void, T>).

    min_bfw: per AN type name ("restiny", "resshort", "resint") the index into the type's
    minimum-bit-flip-weight A list to re-encode to while loading (the reference's --AN-minbfw)."""
    from . import bat as B
    from . import capi

    out = {}
    for name, typ, t in read_header(prefix):
        if columns is not None and name not in columns:
            continue
        if t is not None and t.name == "resbigint":
            # Transaction.cpp:378-381: RESBIGINT columns are "currently unsupported for load"
            raise ValueError(f"column {name}: data type {typ} in header currently unsupported for load")
        if t is None:
            if columns is not None:
                raise ValueError(f"column {name} has type {typ}, which the hardened integer path does not load")
            continue
        codes = read_column(prefix, name, t)
        meta = column_meta(t)
        col = capi.Column.from_numpy(codes, meta)
        k = (min_bfw or {}).get(t.name)
        if k:
            a_new = _BFW[t.name][k]
            factor = (an.inverse(meta.a, 8 * t.width) * a_new) & ((1 << (8 * t.width)) - 1)
            plain = {2: an.SHORTINT, 4: an.INT, 8: an.BIGINT}[t.width]
            src = capi.Column(col.buf, col.n, an.plain_meta(plain))
            dst = capi.DeviceBuffer(t.width * col.n)
            capi.encode_an_into(src, factor, t.width, dst)
            col = capi.Column(dst, col.n, column_meta(t, a_new))
        out[name] = B.BAT(capi.Column.void(col.n), col)
    return out
