"""Star-schema queries on the hardened operators: the thirteen SSB queries of the reference's
continuous drivers (tools/ssbm/ssbm-q11_continuous.cpp .. ssbm-q43_continuous.cpp) with
integer-coded dimension attributes, in two forms over the same device-resident BATs
(ahead_amd.bat):

  chain   the reference's operator chain, operator by operator (the drop-in path);
  fused   the dimension selections as in the chain, then one fact scan
          (ahead_hip_star_query) in place of the chain's fact-side operators, with the chain's
          row semantics; when a dimension makes the fused form inapplicable (a corrupt or repeated
          key, too many groups ...), the chain runs instead.

Schema (seeded; the fact columns are generated on the device; `host_tables` builds the same tables
with numpy):
  date      n_date rows, one per day from 1992-01-01: datekey = row id (resint, A=63877),
            d_year (resshort, A=63877), d_yearmonthnum = year * 100 + month (resint, A=65117),
            d_weeknuminyear = 1 + day of year / 7 (restiny, A=233)
  customer  n_cust rows: custkey = row id, region in [0,5), nation = 5*region + [0,5),
            city = 10*nation + [0,10)  (restiny, A=233)
  supplier  n_supp rows: the same attributes as customer
  part      n_part rows: partkey = row id, mfgr in [0,5), category = 5*mfgr + [0,5) (restiny,
            A=233), brand = 40*category + [0,40) (resshort, A=63877)
  lineorder n_fact rows: orderdate, custkey, suppkey, partkey uniform over the dimension keys
            (resint, A=63877); quantity in [1,50], discount in [0,10] (restiny, A=233);
            extendedprice, revenue in [90000, 10500000], supplycost in [1000, 80000] (resint,
            A=65117)
The reference's string constants map to these codes: regions AFRICA 0, AMERICA 1, ASIA 2,
EUROPE 3, MIDDLE EAST 4; the k-th nation of a region (k = 0..4 in dbgen's order) is 5*region + k,
so UNITED STATES = 9 and UNITED KINGDOM = 19; city "UNITED KIk" = 10*19 + k; MFGR#m = m - 1,
MFGR#mc = 5*(m-1) + c-1, MFGR#mcbb = 40*category + bb-1; d_yearmonth "Dec1997" is
d_yearmonthnum 199712.

Seeded single-bit flips at `rate` go into the columns the queries only select on: lo_quantity and
the region, mfgr and category attributes.  selectAN checks every row it scans and reports it, and a
corrupt code takes part in the predicate on its raw value (selectAN_scalar.tcc:84-90).  Flips
anywhere else make the reference's chains throw: a fetch drops a row whose fetched code is
corrupt, and the later operators then see inputs of different sizes (groupbyAN.tcc:105-107).
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace

import numpy as np

from . import an, datagen
from . import bat as B

A_KEY, A_TINY, A_SHORT, A_MEAS = 63877, 233, 63877, 65117
A_OIDX = an.A_OID
A_OIDX_INV = an.A_OID_INV

REGION = {"AFRICA": 0, "AMERICA": 1, "ASIA": 2, "EUROPE": 3, "MIDDLE EAST": 4}
US, UK = 5 * REGION["AMERICA"] + 4, 5 * REGION["EUROPE"] + 4


def uk_city(k: int) -> int:
    return 10 * UK + k


def category(m: int, c: int) -> int:
    return 5 * (m - 1) + (c - 1)


def brand(m: int, c: int, b: int) -> int:
    return 40 * category(m, c) + (b - 1)


@dataclass
class SsbSpec:
    n_fact: int = 600_000_000
    n_date: int = 2556
    n_cust: int = 3_000_000
    n_supp: int = 200_000
    n_part: int = 1_400_000
    rate: float = 1e-7
    seed: int = 777
    flip_cols: tuple | None = None  # columns that receive flips (None: the COLUMNS defaults)

    def flipped(self, name: str) -> bool:
        return COLUMNS[name][4] if self.flip_cols is None else name in self.flip_cols


META = {
    "key": an.an_meta(an.RESINT, A_KEY),
    "tiny": an.an_meta(an.RESTINY, A_TINY),
    "short": an.an_meta(an.RESSHORT, A_SHORT),
    "meas": an.an_meta(an.RESINT, A_MEAS),
}

# column name -> (table, kind, stream, value spec, flipped)
#   value spec: ("rowid",) | ("uniform", lo, hi) | ("cal", field) | ("sub", parent column, mult, span)
COLUMNS = {
    "d_datekey": ("date", "key", 1, ("rowid",), False),
    "d_year": ("date", "short", 2, ("cal", "year"), False),
    "d_yearmonthnum": ("date", "meas", 24, ("cal", "yearmonthnum"), False),
    "d_weeknuminyear": ("date", "tiny", 25, ("cal", "weeknum"), False),
    "c_custkey": ("cust", "key", 3, ("rowid",), False),
    "c_region": ("cust", "tiny", 4, ("uniform", 0, 4), True),
    "c_nation": ("cust", "tiny", 5, ("sub", "c_region", 5, 5), False),
    "c_city": ("cust", "tiny", 6, ("sub", "c_nation", 10, 10), False),
    "s_suppkey": ("supp", "key", 7, ("rowid",), False),
    "s_region": ("supp", "tiny", 8, ("uniform", 0, 4), True),
    "s_nation": ("supp", "tiny", 9, ("sub", "s_region", 5, 5), False),
    "s_city": ("supp", "tiny", 10, ("sub", "s_nation", 10, 10), False),
    "p_partkey": ("part", "key", 11, ("rowid",), False),
    "p_mfgr": ("part", "tiny", 12, ("uniform", 0, 4), True),
    "p_category": ("part", "tiny", 13, ("sub", "p_mfgr", 5, 5), True),
    "p_brand": ("part", "short", 14, ("sub", "p_category", 40, 40), False),
    "lo_orderdate": ("fact", "key", 15, ("fk", "date"), False),
    "lo_custkey": ("fact", "key", 16, ("fk", "cust"), False),
    "lo_suppkey": ("fact", "key", 17, ("fk", "supp"), False),
    "lo_partkey": ("fact", "key", 18, ("fk", "part"), False),
    "lo_quantity": ("fact", "tiny", 19, ("uniform", 1, 50), True),
    "lo_discount": ("fact", "tiny", 20, ("uniform", 0, 10), False),
    "lo_extendedprice": ("fact", "meas", 21, ("uniform", 90_000, 10_500_000), False),
    "lo_revenue": ("fact", "meas", 22, ("uniform", 90_000, 10_500_000), False),
    "lo_supplycost": ("fact", "meas", 23, ("uniform", 1000, 80_000), False),
}

# dimension -> (fact FK column, dimension key column)
DIM_KEYS = {"date": ("lo_orderdate", "d_datekey"), "cust": ("lo_custkey", "c_custkey"),
            "supp": ("lo_suppkey", "s_suppkey"), "part": ("lo_partkey", "p_partkey")}
DIM_KEYS_SET = {c for pair in DIM_KEYS.values() for c in pair}


def _rows(spec: SsbSpec, table: str) -> int:
    return {"date": spec.n_date, "cust": spec.n_cust, "supp": spec.n_supp, "part": spec.n_part,
            "fact": spec.n_fact}[table]


def _calendar(n: int, what: str) -> np.ndarray:
    days = np.datetime64("1992-01-01") + np.arange(n)
    year = days.astype("datetime64[Y]").astype(np.int64) + 1970
    if what == "year":
        return year.astype(np.uint64)
    if what == "yearmonthnum":
        month = days.astype("datetime64[M]").astype(np.int64) % 12 + 1
        return (year * 100 + month).astype(np.uint64)
    doy = (days - days.astype("datetime64[Y]")).astype(np.int64)
    return (doy // 7 + 1).astype(np.uint64)


def host_values(spec: SsbSpec, name: str, cache: dict, start: int = 0, count: int | None = None) -> np.ndarray:
    """Unencoded values of one column (numpy twin of the device generator).  Fact columns can be
    taken in row ranges [start, start + count)."""
    table, kind, stream, vs, _ = COLUMNS[name]
    n = _rows(spec, table)
    if table == "fact" and (start != 0 or count is not None):
        count = n - start if count is None else count
        if vs[0] == "fk":
            lo, hi = 0, _rows(spec, vs[1]) - 1
        else:
            lo, hi = vs[1], vs[2]
        r = datagen.rand_u64(spec.seed, stream, count, start)
        with np.errstate(over="ignore"):
            return np.uint64(lo) + (((r >> np.uint64(32)) * np.uint64(hi - lo + 1)) >> np.uint64(32))
    if name in cache:
        return cache[name]
    if vs[0] == "rowid":
        v = np.arange(n, dtype=np.uint64)
    elif vs[0] == "cal":
        v = _calendar(n, vs[1])
    elif vs[0] == "uniform":
        v = datagen.uniform(n, spec.seed, stream, vs[1], vs[2])
    elif vs[0] == "fk":
        v = datagen.uniform(n, spec.seed, stream, 0, _rows(spec, vs[1]) - 1)
    else:  # ("sub", parent, mult, span)
        v = host_values(spec, vs[1], cache) * np.uint64(vs[2]) + datagen.uniform(n, spec.seed, stream, 0, vs[3] - 1)
    cache[name] = v
    return v


def host_tables(spec: SsbSpec) -> dict:
    """name -> (codes numpy array, meta, flipped positions) with the device generator's streams."""
    cache, out = {}, {}
    th = datagen.flip_threshold(spec.rate)
    for name, (table, kind, stream, vs, _) in COLUMNS.items():
        meta = META[kind]
        codes = datagen.encode(host_values(spec, name, cache), meta.a, meta.width)
        pos = datagen.inject_flips(codes, spec.seed, 100 + stream, th) if spec.flipped(name) else np.zeros(0, np.uint64)
        out[name] = (codes, meta, pos)
    return out


def device_tables(spec: SsbSpec) -> dict:
    """name -> <void, T> BAT on the device.  Values come from host_values (the derived attributes
    need their parents), except the fact columns, which are generated on the device."""
    from . import capi

    cache, out = {}, {}
    th = datagen.flip_threshold(spec.rate)
    for name, (table, kind, stream, vs, _) in COLUMNS.items():
        meta = META[kind]
        n = _rows(spec, table)
        flipped = spec.flipped(name)
        if table == "fact" and vs[0] in ("fk", "uniform"):
            hi = _rows(spec, vs[1]) - 1 if vs[0] == "fk" else vs[2]
            lo = 0 if vs[0] == "fk" else vs[1]
            buf = capi.DeviceBuffer(meta.width * n)
            capi.gen_uniform(buf, meta.width, n, spec.seed, stream, lo, hi, meta.a)
            if flipped:
                capi.inject_flips(buf, meta.width, n, spec.seed, 100 + stream, th)
            col = capi.Column(buf, n, meta)
        else:
            codes = datagen.encode(host_values(spec, name, cache), meta.a, meta.width)
            if flipped:
                datagen.inject_flips(codes, spec.seed, 100 + stream, th)
            col = capi.Column.from_numpy(codes, meta)
        out[name] = B.BAT(capi.Column.void(n), col)
    return out


# ------------------------------------------------------------------------------------------------
# query specifications
# ------------------------------------------------------------------------------------------------

PROBE, REFETCH = 1, 2  # ahead_hip_star_query dimension flags
SUM, SUBTRACT, MULSUM = 0, 1, 2  # measure kinds


@dataclass
class Dim:
    """One dimension of a query, in the chain's join order.  sels: (attribute column, predicate)
    pairs whose selections the chain intersects; a predicate is (op, v) or (op1, v1, op2, v2, comb)
    on unencoded values.  attr: the group attribute.  flags: PROBE (the chain probes the fact FK
    column directly), REFETCH (the chain fetches the FK again for the attribute join)."""
    name: str
    sels: list = field(default_factory=list)
    attr: str | None = None
    flags: int = 0


@dataclass
class Query:
    """A query shape: family picks the chain (q1: ssbm-q1x, q2: ssbm-q2x, q3: ssbm-q3x, q41, q42,
    q43); preds are fact-column predicates (Q1.x); dims in chain order; meas and kind the measure
    (SUM of meas[0], SUBTRACT meas[0] - meas[1], MULSUM meas[0] * meas[1]); pre_mask flags the
    measures fetched right after the fact predicates; groups lists the dimensions the chain groups
    by, in its groupbyAN order."""
    family: str
    dims: list
    meas: list
    kind: int
    preds: list = field(default_factory=list)
    pre_mask: int = 0
    groups: list = field(default_factory=list)


def q1(qty: tuple, disc: tuple, date_sels: list) -> Query:
    return Query("q1", [Dim("date", date_sels)], ["lo_extendedprice", "lo_discount"], MULSUM,
                 preds=[("lo_quantity", qty), ("lo_discount", disc)], pre_mask=0b10)


def q2(s_sel: tuple, p_sel: tuple) -> Query:
    return Query("q2", [Dim("supp", [s_sel], None, PROBE), Dim("part", [p_sel], "p_brand", REFETCH),
                        Dim("date", [], "d_year")], ["lo_revenue"], SUM, groups=[2, 1])


def q3(s_sel: tuple, c_sel: tuple, d_sel: tuple, s_attr: str, c_attr: str) -> Query:
    return Query("q3", [Dim("supp", [s_sel], s_attr, PROBE | REFETCH), Dim("cust", [c_sel], c_attr, REFETCH),
                        Dim("date", [d_sel], "d_year", REFETCH)], ["lo_revenue"], SUM, groups=[2, 0, 1])


def q_p1(year: int = 1993, disc_lo: int = 1, disc_hi: int = 3, qty_lt: int = 25) -> Query:
    """Q1.1 (ssbm-q11_continuous.cpp:69-121) with its constants as parameters."""
    return q1((B.LT, qty_lt), (B.GE, disc_lo, B.LE, disc_hi, B.AND), [("d_year", (B.EQ, year))])


def q_p2(category: int = category(1, 2), s_region: int = REGION["AMERICA"]) -> Query:
    """Q2.1 (ssbm-q21_continuous.cpp:84-181)."""
    return q2(("s_region", (B.EQ, s_region)), ("p_category", (B.EQ, category)))


def q_p3(c_region: int = REGION["ASIA"], s_region: int = REGION["ASIA"], year_lo: int = 1992,
         year_hi: int = 1997) -> Query:
    """Q3.1 (ssbm-q31_continuous.cpp:89-212)."""
    return q3(("s_region", (B.EQ, s_region)), ("c_region", (B.EQ, c_region)),
              ("d_year", (B.GE, year_lo, B.LE, year_hi, B.AND)), "s_nation", "c_nation")


def q_p4(c_region: int = REGION["AMERICA"], s_region: int = REGION["AMERICA"], mfgr: tuple = (0, 1)) -> Query:
    """Q4.1 (ssbm-q41_continuous.cpp:98-212)."""
    return Query("q41", [Dim("part", [("p_mfgr", (B.EQ, mfgr[0], B.EQ, mfgr[1], B.OR))], None, PROBE),
                         Dim("supp", [("s_region", (B.EQ, s_region))]),
                         Dim("cust", [("c_region", (B.EQ, c_region))], "c_nation", REFETCH),
                         Dim("date", [], "d_year", REFETCH)],
                 ["lo_revenue", "lo_supplycost"], SUBTRACT, groups=[3, 2])


_YEARS_97_98 = ("d_year", (B.EQ, 1997, B.EQ, 1998, B.OR))
_UK_CITIES = (B.EQ, uk_city(1), B.EQ, uk_city(5), B.OR)

QUERIES = {
    "Q1.1": q_p1(),
    "Q1.2": q1((B.GE, 26, B.LE, 35, B.AND), (B.GE, 4, B.LE, 6, B.AND), [("d_yearmonthnum", (B.EQ, 199401))]),
    "Q1.3": q1((B.GE, 26, B.LE, 35, B.AND), (B.GE, 5, B.LE, 7, B.AND),
               [("d_year", (B.EQ, 1994)), ("d_weeknuminyear", (B.EQ, 6))]),
    "Q2.1": q_p2(),
    "Q2.2": q2(("s_region", (B.EQ, REGION["ASIA"])), ("p_brand", (B.GE, brand(2, 2, 21), B.LE, brand(2, 2, 28), B.AND))),
    "Q2.3": q2(("s_region", (B.EQ, REGION["EUROPE"])), ("p_brand", (B.EQ, brand(2, 2, 39)))),
    "Q3.1": q_p3(),
    "Q3.2": q3(("s_nation", (B.EQ, US)), ("c_nation", (B.EQ, US)), ("d_year", (B.GE, 1992, B.LE, 1997, B.AND)),
               "s_city", "c_city"),
    "Q3.3": q3(("s_city", _UK_CITIES), ("c_city", _UK_CITIES), ("d_year", (B.GE, 1992, B.LE, 1997, B.AND)),
               "s_city", "c_city"),
    "Q3.4": q3(("s_city", _UK_CITIES), ("c_city", _UK_CITIES), ("d_yearmonthnum", (B.EQ, 199712)),
               "s_city", "c_city"),
    "Q4.1": q_p4(),
    # ssbm-q42_continuous.cpp:103-264
    "Q4.2": Query("q42", [Dim("part", [("p_mfgr", (B.EQ, 0, B.EQ, 1, B.OR))], "p_category", PROBE | REFETCH),
                          Dim("supp", [("s_region", (B.EQ, REGION["AMERICA"]))], "s_nation", REFETCH),
                          Dim("cust", [("c_region", (B.EQ, REGION["AMERICA"]))]),
                          Dim("date", [_YEARS_97_98], "d_year", REFETCH)],
                  ["lo_revenue", "lo_supplycost"], SUBTRACT, groups=[3, 1, 0]),
    # ssbm-q43_continuous.cpp:99-235
    "Q4.3": Query("q43", [Dim("part", [("p_category", (B.EQ, category(1, 4)))], "p_brand", PROBE | REFETCH),
                          Dim("supp", [("s_nation", (B.EQ, US))], "s_city", REFETCH),
                          Dim("date", [_YEARS_97_98], "d_year", REFETCH)],
                  ["lo_revenue", "lo_supplycost"], SUBTRACT, groups=[2, 1, 0]),
}

SHAPES = {"P1": q_p1, "P2": q_p2, "P3": q_p3, "P4": q_p4}


# ------------------------------------------------------------------------------------------------
# the operator chains
# ------------------------------------------------------------------------------------------------

def _decode_sum(col_bat) -> int:
    v = int(col_bat.tail.to_numpy()[0])
    return (v * col_bat.tail.meta.a_inv) & an.MASK64


class Errors:
    """Collects every error vector a query's operators return, in call order."""

    def __init__(self):
        self.items: list[tuple[str, np.ndarray]] = []

    def add(self, name: str, *vecs):
        for k, v in enumerate(vecs):
            if v is not None and v.size:
                self.items.append((f"{name}.e{k + 1}", v))

    def total(self) -> int:
        return sum(v.size for _, v in self.items)


def _enc(col, v: int) -> int:
    return (v * col.tail.meta.a) & an.MASK64


def _measure(ops, E: Errors, bat, name: str):
    """A fetched attribute or measure as grouping, arithmetic and aggregation take it: unchanged,
    or checked and decoded here when the strategy decodes late (ops.late_decode;
    ssbm-q11_late.cpp:110-120, ssbm-q21_late.cpp:146-153)."""
    late = getattr(ops, "late_decode", None)
    return bat if late is None else late(bat, E, name)


def _select(ops, t: dict, E: Errors, attr: str, pred: tuple):
    col = t[attr]
    if len(pred) == 2:
        sel, e = ops.selectAN(col, pred[0], _enc(col, pred[1]))
    else:
        sel, e = ops.selectAN(col, pred[0], _enc(col, pred[1]), pred[2], _enc(col, pred[3]), pred[4])
    E.add(f"select_{attr}", e)
    return sel


def _select_dim(ops, t: dict, E: Errors, key: str, attr: str, pred: tuple):
    """selectAN on a dimension attribute, then matchjoinAN(reverse(key), mirror_head(sel)):
    key | OID dimension (ssbm-q21_continuous.cpp:87-93)."""
    sel = _select(ops, t, E, attr, pred)
    r = ops.matchjoinAN(t[key].reverse(), sel.mirror_head())
    E.add(f"match_{attr}", *r[1:])
    return r[0]


def chain_q1(t: dict, q: Query, ops=B) -> dict:
    """Q1.1 / Q1.2 / Q1.3 (ssbm-q11_continuous.cpp:69-121, -q12:69-121, -q13:72-140)."""
    E = Errors()
    (qc, qp), (dc, dp) = q.preds
    LQ, LD, LO, LE = t[qc], t[dc], t["lo_orderdate"], t["lo_extendedprice"]
    s1 = _select(ops, t, E, qc, qp)
    s2 = _select(ops, t, E, dc, dp)
    r4 = ops.matchjoinAN(s1.mirror_head(), s2.mirror_head())
    E.add("match_selections", *r4[1:])
    bat5 = r4[0].mirror_head()
    r6 = ops.matchjoinAN(bat5, LO)
    E.add("fetch_orderdate", *r6[1:])
    sels = q.dims[0].sels
    s7 = _select(ops, t, E, *sels[0])
    if len(sels) > 1:  # d_year and d_weeknuminyear (ssbm-q13_continuous.cpp:100-110)
        s8 = _select(ops, t, E, *sels[1])
        rA = ops.matchjoinAN(s7.mirror_head(), s8.mirror_head())
        E.add("match_date_selections", *rA[1:])
        s7 = rA[0]
    r9 = ops.matchjoinAN(s7.mirror_head(), t["d_datekey"])
    E.add("fetch_datekey", *r9[1:])
    rB = ops.hashjoinAN(r6[0], r9[0].reverse())
    E.add("join_date", *rB[1:])
    batC = rB[0].mirror_head()
    rD = ops.matchjoinAN(batC, LE)
    E.add("fetch_price", *rD[1:])
    rE = ops.matchjoinAN(bat5, LD)
    E.add("fetch_discount", *rE[1:])
    rF = ops.matchjoinAN(batC, rE[0])
    E.add("match_discount", *rF[1:])
    rG = ops.aggregate_mul_sumAN(_measure(ops, E, rD[0], "price"), _measure(ops, E, rF[0], "discount"))
    E.add("mul_sum", *rG[1:])
    return {"result": _decode_sum(rG[0]), "rows": rD[0].size(), "errors": E}


def _groups(ops, E: Errors, attrs: list, measure, measure_ready: bool = False) -> dict:
    """groupbyAN over the attributes (unary, then binary), grouped sum, and the group keys."""
    attrs = [_measure(ops, E, a, f"attr_{k}") for k, a in enumerate(attrs)]
    if not measure_ready:
        measure = _measure(ops, E, measure, "measure")
    g = ops.groupbyAN(attrs[0])
    E.add("group_0", *g[2:])
    for k, a in enumerate(attrs[1:]):
        g = ops.groupbyAN(a, g[0], g[1].size())
        E.add(f"group_{k + 1}", *g[2:])
    s = ops.aggregate_sum_groupedAN(measure, g[0], g[1].size())
    E.add("grouped_sum", *s[1:])
    keys = []
    for k, a in enumerate(attrs):
        r = ops.fetchjoinAN(g[1], a)
        E.add(f"group_key_{k}", *r[1:])
        m = r[0].tail.meta
        keys.append((r[0].tail.to_numpy().astype(np.uint64) * np.uint64(m.a_inv)) & np.uint64((1 << (8 * m.width)) - 1))
    sums = (s[0].tail.to_numpy() * np.uint64(s[0].tail.meta.a_inv))
    return {"sums": sums, "keys": keys, "groups": int(g[1].size())}


def chain_q2(t: dict, q: Query, ops=B) -> dict:
    """Q2.1 / Q2.2 / Q2.3 (ssbm-q21_continuous.cpp:84-181, -q22 and -q23 alike)."""
    E = Errors()
    supp = _select_dim(ops, t, E, "s_suppkey", *q.dims[0].sels[0])
    r5 = ops.hashjoinAN(t["lo_suppkey"], supp, (A_OIDX, A_OIDX_INV), (supp.tail.meta.a, supp.tail.meta.a_inv))
    E.add("join_supplier", *r5[1:])
    r7 = ops.matchjoinAN(r5[0].mirror_head(), t["lo_partkey"])
    E.add("fetch_partkey", *r7[1:])
    part = _select_dim(ops, t, E, "p_partkey", *q.dims[1].sels[0])
    rC = ops.hashjoinAN(r7[0], part)
    E.add("join_part", *rC[1:])
    rF = ops.matchjoinAN(rC[0].mirror_head(), t["lo_orderdate"])
    E.add("fetch_orderdate", *rF[1:])
    rI = ops.hashjoinAN(rF[0], t["d_datekey"].reverse(), (rF[0].head.meta.a, rF[0].head.meta.a_inv),
                        (A_OIDX, A_OIDX_INV))
    E.add("join_date", *rI[1:])
    batW = rI[0].mirror_head()
    rX = ops.matchjoinAN(batW, t["lo_partkey"])
    E.add("fetch_partkey_2", *rX[1:])
    rZ = ops.hashjoinAN(rX[0], t["p_partkey"].reverse(), (rX[0].head.meta.a, rX[0].head.meta.a_inv),
                        (A_OIDX, A_OIDX_INV))
    E.add("join_part_2", *rZ[1:])
    ay = ops.fetchjoinAN(rI[0].clear_head(), t["d_year"])
    E.add("fetch_year", *ay[1:])
    ab = ops.fetchjoinAN(rZ[0].clear_head(), t["p_brand"])
    E.add("fetch_brand", *ab[1:])
    ar = ops.fetchjoinAN(batW.clear_head(), t["lo_revenue"])
    E.add("fetch_revenue", *ar[1:])
    g = _groups(ops, E, [ay[0], ab[0]], ar[0])
    g.update({"rows": ar[0].size(), "errors": E})
    return g


def _join_dim_oids(ops, t: dict, E: Errors, rows, fact_fk: str, build, tag: str):
    """fetchjoinAN(rows, fact FK) then hashjoinAN with the dimension's (key | OID) BAT, re-encoding
    both sides: VOID | OID dimension for every row of `rows` (ssbm-q31_continuous.cpp:153-160)."""
    fk = ops.fetchjoinAN(rows, t[fact_fk])
    E.add(f"fetch_{tag}", *fk[1:])
    j = ops.hashjoinAN(fk[0], build, (A_OIDX, A_OIDX_INV), (A_OIDX, A_OIDX_INV))
    E.add(f"join_{tag}", *j[1:])
    return j[0].clear_head()


def chain_q3(t: dict, q: Query, ops=B) -> dict:
    """Q3.1 .. Q3.4 (ssbm-q31_continuous.cpp:89-212, -q32 .. -q34 alike)."""
    E = Errors()
    ds, dc, dd = q.dims
    supp = _select_dim(ops, t, E, "s_suppkey", *ds.sels[0])
    r5 = ops.hashjoinAN(t["lo_suppkey"], supp, (A_OIDX, A_OIDX_INV), (supp.tail.meta.a, supp.tail.meta.a_inv))
    E.add("join_supplier", *r5[1:])
    cust = _select_dim(ops, t, E, "c_custkey", *dc.sels[0])
    r11 = ops.matchjoinAN(r5[0].mirror_head(), t["lo_custkey"])
    E.add("fetch_custkey", *r11[1:])
    r12 = ops.hashjoinAN(r11[0], cust)
    E.add("join_customer", *r12[1:])
    date = _select_dim(ops, t, E, "d_datekey", *dd.sels[0])
    r18 = ops.matchjoinAN(r12[0].mirror_head(), t["lo_orderdate"])
    E.add("fetch_orderdate", *r18[1:])
    r19 = ops.hashjoinAN(r18[0], date)
    E.add("join_date", *r19[1:])
    rows = r19[0].mirror_head().clear_head()
    ar = ops.fetchjoinAN(rows, t["lo_revenue"])
    E.add("fetch_revenue", *ar[1:])
    s_oid = _join_dim_oids(ops, t, E, rows, "lo_suppkey", supp, "suppkey")
    as_ = ops.fetchjoinAN(s_oid, t[ds.attr])
    E.add(f"fetch_{ds.attr}", *as_[1:])
    c_oid = _join_dim_oids(ops, t, E, rows, "lo_custkey", cust, "custkey")
    ac = ops.fetchjoinAN(c_oid, t[dc.attr])
    E.add(f"fetch_{dc.attr}", *ac[1:])
    d_oid = _join_dim_oids(ops, t, E, rows, "lo_orderdate", date, "orderdate")
    ad = ops.fetchjoinAN(d_oid, t["d_year"])
    E.add("fetch_year", *ad[1:])
    g = _groups(ops, E, [ad[0], as_[0], ac[0]], ar[0])
    g.update({"rows": ar[0].size(), "errors": E})
    return g


def _profit(ops, t: dict, E: Errors, rows):
    ar = ops.fetchjoinAN(rows, t["lo_revenue"])
    E.add("fetch_revenue", *ar[1:])
    asc = ops.fetchjoinAN(rows, t["lo_supplycost"])
    E.add("fetch_supplycost", *asc[1:])
    ap = ops.arithmeticAN(ops.SUB, _measure(ops, E, ar[0], "revenue"), _measure(ops, E, asc[0], "supplycost"),
                          a_result=A_MEAS, result_type=an.RESINT)
    E.add("profit", *ap[1:])
    return ap[0]


def _probe_dim(ops, t: dict, E: Errors, fact_fk: str, dim):
    """hashjoinAN of the whole fact FK column with the dimension's (key | OID) BAT."""
    r = ops.hashjoinAN(t[fact_fk], dim, (A_OIDX, A_OIDX_INV), (dim.tail.meta.a, dim.tail.meta.a_inv))
    E.add(f"join_{fact_fk}", *r[1:])
    return r[0].mirror_head()


def _fetch_join(ops, t: dict, E: Errors, rows, fact_fk: str, dim, tag: str):
    """matchjoinAN(rows, fact FK) then hashjoinAN with the dimension: the rows that join."""
    f = ops.matchjoinAN(rows, t[fact_fk])
    E.add(f"fetch_{tag}", *f[1:])
    j = ops.hashjoinAN(f[0], dim)
    E.add(f"join_{tag}", *j[1:])
    return j[0].mirror_head()


def chain_q41(t: dict, q: Query, ops=B) -> dict:
    """Q4.1 (ssbm-q41_continuous.cpp:98-212)."""
    E = Errors()
    dp, dsu, dcu, _ = q.dims
    part = _select_dim(ops, t, E, "p_partkey", *dp.sels[0])
    bat6 = _probe_dim(ops, t, E, "lo_partkey", part)
    supp = _select_dim(ops, t, E, "s_suppkey", *dsu.sels[0])
    r12 = _fetch_join(ops, t, E, bat6, "lo_suppkey", supp, "suppkey")
    cust = _select_dim(ops, t, E, "c_custkey", *dcu.sels[0])
    r19 = _fetch_join(ops, t, E, r12, "lo_custkey", cust, "custkey")
    rows = r19.clear_head()
    profit = _profit(ops, t, E, rows)
    c_oid = _join_dim_oids(ops, t, E, rows, "lo_custkey", cust, "custkey_2")
    an_ = ops.fetchjoinAN(c_oid, t[dcu.attr])
    E.add(f"fetch_{dcu.attr}", *an_[1:])
    d_oid = _join_dim_oids(ops, t, E, rows, "lo_orderdate", t["d_datekey"].reverse(), "orderdate")
    ay = ops.fetchjoinAN(d_oid, t["d_year"])
    E.add("fetch_year", *ay[1:])
    g = _groups(ops, E, [ay[0], an_[0]], profit, measure_ready=True)
    g.update({"rows": profit.size(), "errors": E})
    return g


def chain_q42(t: dict, q: Query, ops=B) -> dict:
    """Q4.2 (ssbm-q42_continuous.cpp:103-264)."""
    E = Errors()
    dp, dsu, dcu, dd = q.dims
    part = _select_dim(ops, t, E, "p_partkey", *dp.sels[0])
    bat6 = _probe_dim(ops, t, E, "lo_partkey", part)
    supp = _select_dim(ops, t, E, "s_suppkey", *dsu.sels[0])
    bat13 = _fetch_join(ops, t, E, bat6, "lo_suppkey", supp, "suppkey")
    cust = _select_dim(ops, t, E, "c_custkey", *dcu.sels[0])
    bat20 = _fetch_join(ops, t, E, bat13, "lo_custkey", cust, "custkey")
    date = _select_dim(ops, t, E, "d_datekey", *dd.sels[0])
    bat27 = _fetch_join(ops, t, E, bat20, "lo_orderdate", date, "orderdate")
    rows = bat27.clear_head()
    profit = _profit(ops, t, E, rows)
    s_oid = _join_dim_oids(ops, t, E, rows, "lo_suppkey", supp, "suppkey_2")
    an_ = ops.fetchjoinAN(s_oid, t[dsu.attr])
    E.add(f"fetch_{dsu.attr}", *an_[1:])
    d_oid = _join_dim_oids(ops, t, E, rows, "lo_orderdate", date, "orderdate_2")
    ay = ops.fetchjoinAN(d_oid, t["d_year"])
    E.add("fetch_year", *ay[1:])
    p_oid = _join_dim_oids(ops, t, E, rows, "lo_partkey", t["p_partkey"].reverse(), "partkey_2")
    ac = ops.fetchjoinAN(p_oid, t[dp.attr])
    E.add(f"fetch_{dp.attr}", *ac[1:])
    g = _groups(ops, E, [ay[0], an_[0], ac[0]], profit, measure_ready=True)
    g.update({"rows": profit.size(), "errors": E})
    return g


def chain_q43(t: dict, q: Query, ops=B) -> dict:
    """Q4.3 (ssbm-q43_continuous.cpp:99-235)."""
    E = Errors()
    dp, dsu, dd = q.dims
    part = _select_dim(ops, t, E, "p_partkey", *dp.sels[0])
    bat6 = _probe_dim(ops, t, E, "lo_partkey", part)
    supp = _select_dim(ops, t, E, "s_suppkey", *dsu.sels[0])
    bat13 = _fetch_join(ops, t, E, bat6, "lo_suppkey", supp, "suppkey")
    date = _select_dim(ops, t, E, "d_datekey", *dd.sels[0])
    bat27 = _fetch_join(ops, t, E, bat13, "lo_orderdate", date, "orderdate")
    rows = bat27.clear_head()
    profit = _profit(ops, t, E, rows)
    s_oid = _join_dim_oids(ops, t, E, rows, "lo_suppkey", supp, "suppkey_2")
    ac = ops.fetchjoinAN(s_oid, t[dsu.attr])
    E.add(f"fetch_{dsu.attr}", *ac[1:])
    d_oid = _join_dim_oids(ops, t, E, rows, "lo_orderdate", date, "orderdate_2")
    ay = ops.fetchjoinAN(d_oid, t["d_year"])
    E.add("fetch_year", *ay[1:])
    p_oid = _join_dim_oids(ops, t, E, rows, "lo_partkey", t["p_partkey"].reverse(), "partkey_2")
    ab = ops.fetchjoinAN(p_oid, t[dp.attr])
    E.add(f"fetch_{dp.attr}", *ab[1:])
    g = _groups(ops, E, [ay[0], ac[0], ab[0]], profit, measure_ready=True)
    g.update({"rows": profit.size(), "errors": E})
    return g


CHAINS = {"q1": chain_q1, "q2": chain_q2, "q3": chain_q3, "q41": chain_q41, "q42": chain_q42, "q43": chain_q43}


def chain(t: dict, q: Query, ops=B) -> dict:
    return CHAINS[q.family](t, q, ops)


# ------------------------------------------------------------------------------------------------
# the fused form
# ------------------------------------------------------------------------------------------------

FUSED_KINDS = ["pred0", "pred1", "meas0", "meas1", "fk0", "fk1", "fk2", "fk3", "refetch0", "refetch1", "refetch2",
               "refetch3", "attr0", "attr1", "attr2", "attr3"]


def _fused_args(t: dict, q: Query, ops, E: "Errors") -> dict:
    """The chain's dimension selections (their error vectors into E) and the fused entry point's
    arguments."""
    dims = []
    for dq in q.dims:
        fk, key = DIM_KEYS[dq.name]
        sel = None
        if dq.sels:
            sel = _select(ops, t, E, *dq.sels[0])
            if len(dq.sels) > 1:
                s2 = _select(ops, t, E, *dq.sels[1])
                r = ops.matchjoinAN(sel.mirror_head(), s2.mirror_head())
                E.add(f"match_{dq.name}_selections", *r[1:])
                sel = r[0]
        dims.append({"fk": t[fk].tail, "key": t[key].tail, "attr": t[dq.attr].tail if dq.attr else None,
                     "sel": sel.head if sel is not None else None, "flags": dq.flags})
    preds = []
    for col, p in q.preds:
        c = t[col]
        if len(p) == 2:
            preds.append((c.tail, p[0], _enc(c, p[1]), B.LT, 0, B.NONE_))
        else:
            preds.append((c.tail, p[0], _enc(c, p[1]), p[2], _enc(c, p[3]), p[4]))
    meas = [t[m].tail for m in q.meas]
    return {"n_fact": t[q.meas[0]].tail.n, "dims": dims, "preds": preds, "meas": meas, "kind": q.kind,
            "pre_mask": q.pre_mask}


def _fused_result(t: dict, q: Query, r: dict, E: "Errors", ops, fallback: bool) -> dict:
    if r["status"] != 0:
        if not fallback:
            return {"status": r["status"], "errors": E}
        out = chain(t, q, ops)
        out["status"] = r["status"]
        return out
    for k, v in enumerate(r["err"]):
        # the fact predicates' vectors are exactly their selectAN's (every row checked)
        E.add(f"select_{q.preds[k][0]}" if k < len(q.preds) else f"fused_{FUSED_KINDS[k]}", v)
    inv = an.inverse(an.A_RESULT, 64)
    sums = (r["sums"] * np.uint64(inv))
    out = {"status": 0, "rows": r["kept"], "errors": E, "first_row": r["first_row"]}
    if q.family == "q1":
        out["result"] = int(sums[0]) if sums.size else 0
        return out
    keys = []
    for d in q.groups:
        m = t[q.dims[d].attr].tail.meta
        keys.append((r["keys"][d] * np.uint64(m.a_inv)) & np.uint64((1 << (8 * m.width)) - 1))
    out.update({"sums": sums, "keys": keys, "groups": r["groups"]})
    return out


def fused(t: dict, q: Query, ops=B, fallback: bool = True, ev_scan: tuple | None = None) -> dict:
    """The query with one fact scan (ops.star_query) after the chain's dimension selections.  The
    result has the chain's form; `status` is the fused entry point's (0: it applied).  With a
    status other than 0 and `fallback`, the chain's result is returned (with that status).  On the
    device operators with one selection per dimension at most, the host waits once (the
    selections' counts are read on the device; every output comes back in one copy)."""
    if ops is B and all(len(d.sels) <= 1 for d in q.dims):
        return _fused_batch_device(t, [q], fallback, ev_scan, single=True)[0]
    return _fused_ops(t, q, ops, fallback, ev_scan)


def _fused_ops(t: dict, q: Query, ops=B, fallback: bool = True, ev_scan: tuple | None = None) -> dict:
    """fused() through the operator layer: each dimension selection returns its OID list (and
    count) to the host before the star query runs."""
    E = Errors()
    r = ops.star_query(**_fused_args(t, q, ops, E), ev_scan=ev_scan)
    return _fused_result(t, q, r, E, ops, fallback)


def fused_batch(t: dict, qs: list, ops=B, fallback: bool = True, ev_scan: tuple | None = None) -> list:
    """Several queries over the same tables with shared fact scans (ops.star_query_batch): each
    result is the one fused() returns for that query alone."""
    if ops is B and qs and all(len(d.sels) <= 1 for q in qs for d in q.dims):
        return _fused_batch_device(t, qs, fallback, ev_scan)
    Es = [Errors() for _ in qs]
    calls = [_fused_args(t, q, ops, E) for q, E in zip(qs, Es)]
    rs = ops.star_query_batch(calls, ev_scan=ev_scan)
    return [_fused_result(t, q, r, E, ops, fallback) for q, r, E in zip(qs, rs, Es)]


# a device batch's read-back arena: the selections' error vectors and every query's outputs
_SEL_ERR_CAP = 1024
_BATCH_GROUPS_CAP = 2048   # the fused form's group limit (more groups: its status, then the chain)
_BATCH_ERR_CAP = 256


def _fused_batch_device(t: dict, qs: list, fallback: bool, ev_scan: tuple | None, single: bool = False) -> list:
    """fused_batch on the device operators with one wait: every dimension selection (one per
    dimension at most) is launched without reading its counts back (the star queries read the OID
    count on the device), then the batch (single: the one query through the single-query entry
    point); the selections' error vectors and all the queries' outputs come back in one copy.  A
    query whose selection errors or outputs overflow the arena's capacities runs again alone
    (through the operator layer / StarQuery.results())."""
    import ctypes

    import torch

    from . import capi

    lib, ws, stream = capi.load(), capi.workspace(), capi.current_stream()
    sels, q_off, off = [], [], 0
    for qi, q in enumerate(qs):
        for dq in q.dims:
            if dq.sels:
                sels.append((qi, dq.sels[0], off))
                off += (16 + 8 * _SEL_ERR_CAP + 255) // 256 * 256
    for q in qs:
        q_off.append(off)
        off += capi.StarQuery.out_bytes(sum(1 for d in q.dims if d.attr), _BATCH_GROUPS_CAP, _BATCH_ERR_CAP)
    arena = capi.DeviceBuffer(off)
    oids = {}
    for qi, (attr, pred), so in sels:
        col = t[attr]
        n = col.tail.n
        buf = capi.DeviceBuffer(8 * max(n, 1))
        op2, th2, comb = (pred[2], _enc(col, pred[3]), pred[4]) if len(pred) > 2 else (B.LT, 0, B.NONE_)
        d = col.tail.desc()
        capi.check(lib.ahead_hip_select_an(ctypes.byref(d), pred[0], _enc(col, pred[1]), op2, th2 & an.MASK64, comb,
                                           B.A_OID, B.A_OID_ERR, buf.ptr, n, arena.ptr + so + 16, _SEL_ERR_CAP,
                                           arena.ptr + so, None, ws.h, stream))
        oids[(qi, attr)] = (buf, so)
    sqs = []
    for qi, q in enumerate(qs):
        dims = []
        for dq in q.dims:
            fk, key = DIM_KEYS[dq.name]
            x = {"fk": t[fk].tail, "key": t[key].tail, "attr": t[dq.attr].tail if dq.attr else None,
                 "flags": dq.flags, "key_index": B.key_index(t[key].tail)}
            if dq.sels:
                buf, so = oids[(qi, dq.sels[0][0])]
                x.update(sel=buf, sel_count=arena.view(so, 16), sel_a_inv=B.RESOID_META.a_inv)
            dims.append(x)
        preds = []
        for col, p in q.preds:
            c = t[col]
            if len(p) == 2:
                preds.append((c.tail, p[0], _enc(c, p[1]), B.LT, 0, B.NONE_))
            else:
                preds.append((c.tail, p[0], _enc(c, p[1]), p[2], _enc(c, p[3]), p[4]))
        sqs.append(capi.StarQuery(t[q.meas[0]].tail.n, dims, preds, [t[m].tail for m in q.meas], q.kind, q.pre_mask,
                                  groups_cap=_BATCH_GROUPS_CAP, err_cap=_BATCH_ERR_CAP, out_arena=(arena, q_off[qi])))
    if ev_scan is not None:
        sqs[0].args.ev_scan[0], sqs[0].args.ev_scan[1] = ev_scan
    if single and len(sqs) == 1:
        sqs[0].launch()
    else:
        capi.star_query_batch(sqs)
    host = torch.empty(off, dtype=torch.uint8, pin_memory=True)
    host.copy_(arena.t[:off], non_blocking=True)
    torch.cuda.current_stream().synchronize()
    h = host.numpy()
    out = []
    for qi, q in enumerate(qs):
        E = Errors()
        rerun = False
        for qj, (attr, _), so in sels:
            if qj != qi:
                continue
            ne = int(h[so + 8: so + 16].view(np.uint64)[0])
            if ne > _SEL_ERR_CAP:
                rerun = True
            E.add(f"select_{attr}", h[so + 16: so + 16 + 8 * min(ne, _SEL_ERR_CAP)].view(np.uint64).copy())
        if rerun:
            out.append(_fused_ops(t, q, fallback=fallback))
            continue
        r = sqs[qi].results_from(h, q_off[qi])
        if r is None:
            r = sqs[qi].results()
        out.append(_fused_result(t, q, r, E, B, fallback))
    return out


def serve_query(t: dict, q: Query):
    """The query as the C++ request loop takes it (ahead_serve_query): its star query arguments
    and, per dimension, the selection the server runs (one per dimension at most)."""
    from . import capi

    dims = []
    for dq in q.dims:
        if len(dq.sels) > 1:
            raise ValueError(f"{q.family}: the request loop runs one selection per dimension")
        fk, key = DIM_KEYS[dq.name]
        dims.append({"fk": t[fk].tail, "key": t[key].tail, "attr": t[dq.attr].tail if dq.attr else None,
                     "flags": dq.flags, "key_index": B.key_index(t[key].tail)})
    preds = []
    for col, p in q.preds:
        c = t[col]
        if len(p) == 2:
            preds.append((c.tail, p[0], _enc(c, p[1]), B.LT, 0, B.NONE_))
        else:
            preds.append((c.tail, p[0], _enc(c, p[1]), p[2], _enc(c, p[3]), p[4]))
    sq = capi.ServeQuery()
    sq.q = capi.star_args(t[q.meas[0]].tail.n, dims, preds, [t[m].tail for m in q.meas], q.kind, q.pre_mask)
    for d, dq in enumerate(q.dims):
        if not dq.sels:
            continue
        attr, pred = dq.sels[0]
        col = t[attr]
        op2, th2, comb = (pred[2], _enc(col, pred[3]), pred[4]) if len(pred) > 2 else (B.LT, 0, B.NONE_)
        sq.has_sel[d] = 1
        sq.sel[d] = capi.ServeSel(col.tail.desc(), pred[0], op2, comb, 0, _enc(col, pred[1]), th2 & an.MASK64,
                                  B.A_OID, B.RESOID_META.a_inv)
    return sq


def serve_result(t: dict, q: Query, r: dict, fallback: bool = True) -> dict:
    """A request-loop result in fused()'s form: the selections' error vectors first, in the
    dimensions' order, then the star query's (the chain's result with fallback when its status is
    not 0)."""
    E = Errors()
    for d, dq in enumerate(q.dims):
        if dq.sels:
            E.add(f"select_{dq.sels[0][0]}", r["sel_err"][d])
    return _fused_result(t, q, r, E, B, fallback)


def fused_ws_bytes(t: dict, q: Query, ops=B) -> int:
    """Device workspace bytes of the query's fused form (its dimension maps and tables)."""
    dims = [{"fk": t[DIM_KEYS[d.name][0]].tail, "key": t[DIM_KEYS[d.name][1]].tail,
             "attr": t[d.attr].tail if d.attr else None, "flags": d.flags} for d in q.dims]
    return ops.star_query_ws_bytes(t[q.meas[0]].tail.n, dims, [t[m].tail for m in q.meas], q.kind)


def _p(shape: str):
    def run(t: dict, ops=B, **kw) -> dict:
        return chain(t, SHAPES[shape](**kw), ops)

    run.__doc__ = SHAPES[shape].__doc__
    return run


p1, p2, p3, p4 = _p("P1"), _p("P2"), _p("P3"), _p("P4")
PIPELINES = {"P1": p1, "P2": p2, "P3": p3, "P4": p4}


# ------------------------------------------------------------------------------------------------
# textbook evaluation on decoded host values (fault-free data), for tests
# ------------------------------------------------------------------------------------------------

_OPS = {0: np.less, 1: np.less_equal, 2: np.greater, 3: np.greater_equal, 4: np.equal, 5: np.not_equal}


def _pred_mask(v: np.ndarray, p: tuple) -> np.ndarray:
    m = _OPS[p[0]](v, np.uint64(p[1]))
    if len(p) == 2:
        return m
    m2 = _OPS[p[2]](v, np.uint64(p[3]))
    return (m & m2) if p[4] == B.AND else (m | m2)


def codes_chunk(spec: SsbSpec, name: str, s0: int, cnt: int) -> np.ndarray:
    """Stored codes of fact rows [s0, s0 + cnt) of one column: encoded values with the generator's
    flips applied (numpy twin of the device columns)."""
    _, kind, stream, _, _ = COLUMNS[name]
    meta = META[kind]
    codes = datagen.encode(host_values(spec, name, {}, s0, cnt), meta.a, meta.width)
    th = datagen.flip_threshold(spec.rate)
    if spec.flipped(name) and th:
        local = np.nonzero(datagen.rand_u64(spec.seed, 100 + stream, cnt, s0) < np.uint64(th))[0]
        if local.size:
            r = datagen.mix(local.astype(np.uint64) + np.uint64(s0) +
                            np.uint64(datagen.stream_key(spec.seed, 100 + stream + 1)))
            bits = ((r >> np.uint64(32)) * np.uint64(8 * meta.width)) >> np.uint64(32)
            codes[local] ^= np.ones(local.size, dtype=codes.dtype) << bits.astype(codes.dtype)
    return codes


def _pred_codes(codes: np.ndarray, meta, p: tuple) -> np.ndarray:
    """A selectAN predicate on stored codes against encoded thresholds (selectAN_scalar.tcc:89)."""
    m = (1 << (8 * meta.width)) - 1
    enc = [((v * meta.a) & m) for v in (p[1],) + ((p[3],) if len(p) > 2 else ())]
    q = (p[0], enc[0]) + ((p[2], enc[1], p[4]) if len(p) > 2 else ())
    return _pred_mask(codes.astype(np.uint64), q)


def _textbook_chunk(spec: SsbSpec, q: Query, dsel: list, dattr: list, s0: int, cnt: int):
    """One fact row range of the textbook evaluation: (rows, total) for Q1.x, else (rows, {key tuple:
    [sum, first row]})."""
    fv = lambda c: host_values(spec, c, {}, s0, cnt)  # noqa: E731
    m = np.ones(cnt, dtype=bool)
    for col, p in q.preds:
        m &= _pred_codes(codes_chunk(spec, col, s0, cnt), META[COLUMNS[col][1]], p)
    fks = []
    for i, dq in enumerate(q.dims):
        fk = fv(DIM_KEYS[dq.name][0]).astype(np.int64)
        m &= dsel[i][fk]
        fks.append(fk)
    idx = np.nonzero(m)[0]
    m0 = fv(q.meas[0])[idx]
    with np.errstate(over="ignore"):
        if q.kind == SUM:
            val = m0
        elif q.kind == SUBTRACT:
            val = m0 - fv(q.meas[1])[idx]
        else:
            val = m0 * fv(q.meas[1])[idx]
        if q.family == "q1":
            return idx.size, int(val.sum(dtype=np.uint64))
    if idx.size == 0:
        return 0, {}
    keys = np.stack([dattr[d][fks[d][idx]] for d in q.groups], axis=1)
    uk, firsts, inv = np.unique(keys, axis=0, return_index=True, return_inverse=True)
    acc = np.zeros(uk.shape[0], dtype=np.uint64)
    np.add.at(acc, inv.reshape(-1), val)
    return idx.size, {tuple(int(x) for x in uk[j]): [int(acc[j]), s0 + int(idx[firsts[j]])] for j in range(uk.shape[0])}


def textbook(spec: SsbSpec, q: Query | str, chunk: int = 50_000_000, workers: int = 1, **kw) -> dict:
    """GROUP BY / SUM of the query with numpy, the fact table taken in chunks of `chunk` rows
    (evaluated by `workers` threads): predicates on the stored codes (flips included, as selectAN
    evaluates them), group attributes and measures on decoded values (columns the chains fetch
    must be fault-free).  q: a Query, a QUERIES name, or a SHAPES name with keyword arguments.
    Groups come in first-occurrence order over the qualifying fact rows."""
    if isinstance(q, str):
        q = QUERIES[q] if q in QUERIES else SHAPES[q](**kw)
    cache = {}
    v = lambda c: host_values(spec, c, cache)  # noqa: E731
    th = datagen.flip_threshold(spec.rate)
    # per dimension: selected rows and the group attribute per key (keys are row ids)
    dsel, dattr = [], []
    for dq in q.dims:
        table = COLUMNS[DIM_KEYS[dq.name][1]][0]
        m = np.ones(_rows(spec, table), dtype=bool)
        for col, p in dq.sels:
            meta = META[COLUMNS[col][1]]
            codes = datagen.encode(v(col), meta.a, meta.width)
            if spec.flipped(col):
                datagen.inject_flips(codes, spec.seed, 100 + COLUMNS[col][2], th)
            m &= _pred_codes(codes, meta, p)
        dsel.append(m)
        dattr.append(v(dq.attr) if dq.attr else None)
    starts = list(range(0, spec.n_fact, chunk))
    job = lambda s0: _textbook_chunk(spec, q, dsel, dattr, s0, min(chunk, spec.n_fact - s0))  # noqa: E731
    if workers > 1 and len(starts) > 1:
        from concurrent.futures import ThreadPoolExecutor

        with ThreadPoolExecutor(workers) as ex:
            parts = list(ex.map(job, starts))
    else:
        parts = [job(s0) for s0 in starts]
    rows = sum(p[0] for p in parts)
    if q.family == "q1":
        return {"result": sum(p[1] for p in parts) & an.MASK64, "rows": rows}
    sums: dict = {}
    for _, d in parts:  # chunk order: the first chunk that holds a key holds its first row
        for k, (sm, f) in d.items():
            if k not in sums:
                sums[k] = [0, f]
            sums[k][0] = (sums[k][0] + sm) & an.MASK64
    order = sorted(sums, key=lambda k: sums[k][1])
    return {"sums": np.array([sums[k][0] for k in order], dtype=np.uint64),
            "keys": [np.array([k[j] for k in order], dtype=np.uint64) for j in range(len(q.groups))],
            "first_row": np.array([sums[k][1] for k in order], dtype=np.uint64),
            "groups": len(order), "rows": rows}


def flip_positions(spec: SsbSpec, name: str, chunk: int = 50_000_000, workers: int = 1) -> np.ndarray:
    """The rows of column `name` that the generator flips (sorted), computed in chunks."""
    table, _, stream, _, _ = COLUMNS[name]
    n = _rows(spec, table)
    th = np.uint64(datagen.flip_threshold(spec.rate))
    if not spec.flipped(name) or th == 0:
        return np.zeros(0, dtype=np.uint64)

    def job(s0):
        r = datagen.rand_u64(spec.seed, 100 + stream, min(chunk, n - s0), s0)
        return np.nonzero(r < th)[0].astype(np.uint64) + np.uint64(s0)

    starts = list(range(0, n, chunk))
    if workers > 1 and len(starts) > 1:
        from concurrent.futures import ThreadPoolExecutor

        with ThreadPoolExecutor(workers) as ex:
            parts = list(ex.map(job, starts))
    else:
        parts = [job(s0) for s0 in starts]
    return np.concatenate(parts) if parts else np.zeros(0, dtype=np.uint64)


def with_spec(spec: SsbSpec, **kw) -> SsbSpec:
    return replace(spec, **kw)
