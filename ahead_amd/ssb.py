"""Star-schema pipelines on the hardened operators: the SSB query chains of the reference's
continuous drivers (tools/ssbm/ssbm-q11_continuous.cpp, -q21, -q31, -q41) with integer-coded
dimension attributes, run operator by operator on device-resident BATs (ahead_amd.bat).

Schema (seeded, generated on the device; `host_tables` builds the same tables with numpy):
  date      n_date rows: datekey = row id (resint, A=63877), year = 1992 + 7*row/n_date (resshort,
            A=63877)
  customer  n_cust rows: custkey = row id, region in [0,5), nation = 5*region + [0,5),
            city = 10*nation + [0,10)  (restiny, A=233)
  supplier  n_supp rows: the same attributes as customer
  part      n_part rows: partkey = row id, mfgr in [0,5), category = 5*mfgr + [0,5) (restiny,
            A=233), brand = 40*category + [0,40) (resshort, A=63877)
  lineorder n_fact rows: orderdate, custkey, suppkey, partkey uniform over the dimension keys
            (resint, A=63877); quantity in [1,50], discount in [0,10] (restiny, A=233);
            extendedprice, revenue in [90000, 10500000], supplycost in [1000, 80000] (resint,
            A=65117)
Seeded single-bit flips at `rate` go into the columns the pipelines only select on: lo_quantity
and the region, mfgr and category attributes.  selectAN checks every row it scans and
reports it, and a corrupt code takes part in the predicate on its raw value
(selectAN_scalar.tcc:84-90).  Flips anywhere else make the reference's chains throw.  A fetch drops
a row whose fetched code is corrupt, and the later operators then see inputs of different sizes
(groupbyAN.tcc:105-107).  Join keys are not safe either: a flip in the upper 32 bits of a 64-bit
key code leaves the truncated decoded key unchanged, so the hash join still matches the row
(hashjoinAN.tcc:199), and a later fetch of the same key drops it.

Pipelines (predicates on attribute codes, as the reference passes encoded constants):
  P1 (Q1.1)  sum(extendedprice * discount) where d_year = 1993, 1 <= discount <= 3, quantity < 25
  P2 (Q2.1)  sum(revenue) group by d_year, p_brand where p_category = 12, s_region = 1
  P3 (Q3.1)  sum(revenue) group by c_nation, s_nation, d_year where c_region = 2, s_region = 2,
             1992 <= d_year <= 1997
  P4 (Q4.1)  sum(revenue - supplycost) group by d_year, c_nation where c_region = 1,
             s_region = 1, p_mfgr in {0, 1}
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import an, datagen
from . import bat as B

A_KEY, A_TINY, A_SHORT, A_MEAS = 63877, 233, 63877, 65117
A_OIDX = an.A_OID
A_OIDX_INV = an.A_OID_INV


@dataclass
class SsbSpec:
    n_fact: int = 600_000_000
    n_date: int = 2556
    n_cust: int = 3_000_000
    n_supp: int = 200_000
    n_part: int = 1_400_000
    rate: float = 1e-7
    seed: int = 777


META = {
    "key": an.an_meta(an.RESINT, A_KEY),
    "tiny": an.an_meta(an.RESTINY, A_TINY),
    "short": an.an_meta(an.RESSHORT, A_SHORT),
    "meas": an.an_meta(an.RESINT, A_MEAS),
}

# column name -> (table, kind, stream, value spec, flipped)
#   value spec: ("rowid",) | ("uniform", lo, hi) | ("year",) | ("sub", parent column, mult, span)
COLUMNS = {
    "d_datekey": ("date", "key", 1, ("rowid",), False),
    "d_year": ("date", "short", 2, ("year",), False),
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


def _rows(spec: SsbSpec, table: str) -> int:
    return {"date": spec.n_date, "cust": spec.n_cust, "supp": spec.n_supp, "part": spec.n_part,
            "fact": spec.n_fact}[table]


def host_values(spec: SsbSpec, name: str, cache: dict) -> np.ndarray:
    """Unencoded values of one column (numpy twin of the device generator)."""
    if name in cache:
        return cache[name]
    table, kind, stream, vs, _ = COLUMNS[name]
    n = _rows(spec, table)
    if vs[0] == "rowid":
        v = np.arange(n, dtype=np.uint64)
    elif vs[0] == "year":
        v = 1992 + (np.arange(n, dtype=np.uint64) * np.uint64(7)) // np.uint64(max(n, 1))
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
    for name, (table, kind, stream, vs, flipped) in COLUMNS.items():
        meta = META[kind]
        codes = datagen.encode(host_values(spec, name, cache), meta.a, meta.width)
        pos = datagen.inject_flips(codes, spec.seed, 100 + stream, th) if flipped else np.zeros(0, np.uint64)
        out[name] = (codes, meta, pos)
    return out


def device_tables(spec: SsbSpec) -> dict:
    """name -> <void, T> BAT on the device.  Values come from host_values (the derived attributes
    need their parents), except the fact columns, which are generated on the device."""
    from . import capi

    cache, out = {}, {}
    th = datagen.flip_threshold(spec.rate)
    for name, (table, kind, stream, vs, flipped) in COLUMNS.items():
        meta = META[kind]
        n = _rows(spec, table)
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


def _decode_sum(col_bat) -> int:
    v = int(col_bat.tail.to_numpy()[0])
    return (v * col_bat.tail.meta.a_inv) & an.MASK64


class Errors:
    """Collects every error vector a pipeline's operators return, in call order."""

    def __init__(self):
        self.items: list[tuple[str, np.ndarray]] = []

    def add(self, name: str, *vecs):
        for k, v in enumerate(vecs):
            if v is not None and v.size:
                self.items.append((f"{name}.e{k + 1}", v))

    def total(self) -> int:
        return sum(v.size for _, v in self.items)


def p1(t: dict, ops=B, year: int = 1993, disc_lo: int = 1, disc_hi: int = 3, qty_lt: int = 25) -> dict:
    """Q1.1 (ssbm-q11_continuous.cpp:69-121)."""
    E = Errors()
    LQ, LD, LO, LE = t["lo_quantity"], t["lo_discount"], t["lo_orderdate"], t["lo_extendedprice"]
    DY, DD = t["d_year"], t["d_datekey"]
    s1, e = ops.selectAN(LQ, ops.LT, qty_lt * LQ.tail.meta.a)
    E.add("select_quantity", e)
    s2, e = ops.selectAN(LD, ops.GE, disc_lo * LD.tail.meta.a, ops.LE, disc_hi * LD.tail.meta.a, ops.AND)
    E.add("select_discount", e)
    r4 = ops.matchjoinAN(s1.mirror_head(), s2.mirror_head())
    E.add("match_selections", *r4[1:])
    bat5 = r4[0].mirror_head()
    r6 = ops.matchjoinAN(bat5, LO)
    E.add("fetch_orderdate", *r6[1:])
    s7, e = ops.selectAN(DY, ops.EQ, year * DY.tail.meta.a)
    E.add("select_year", e)
    r9 = ops.matchjoinAN(s7.mirror_head(), DD)
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
    rG = ops.aggregate_mul_sumAN(rD[0], rF[0])
    E.add("mul_sum", *rG[1:])
    return {"result": _decode_sum(rG[0]), "rows": rD[0].size(), "errors": E}


def _select_dim(ops, t: dict, E: Errors, key: str, attr: str, pred: tuple, tag: str):
    """selectAN on a dimension attribute, then matchjoinAN(reverse(key), mirror_head(sel)):
    key | OID dimension (ssbm-q21_continuous.cpp:87-93)."""
    col = t[attr]
    a = col.tail.meta.a
    if len(pred) == 2:
        sel, e = ops.selectAN(col, pred[0], pred[1] * a)
    else:
        sel, e = ops.selectAN(col, pred[0], pred[1] * a, pred[2], pred[3] * a, pred[4])
    E.add(f"select_{tag}", e)
    r = ops.matchjoinAN(t[key].reverse(), sel.mirror_head())
    E.add(f"match_{tag}", *r[1:])
    return r[0]


def _groups(ops, E: Errors, attrs: list, measure) -> dict:
    """groupbyAN over the attributes (unary, then binary), grouped sum, and the group keys."""
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


def p2(t: dict, ops=B, category: int = 12, s_region: int = 1) -> dict:
    """Q2.1 (ssbm-q21_continuous.cpp:84-181)."""
    E = Errors()
    supp = _select_dim(ops, t, E, "s_suppkey", "s_region", (ops.EQ, s_region), "s_region")
    r5 = ops.hashjoinAN(t["lo_suppkey"], supp, (A_OIDX, A_OIDX_INV), (supp.tail.meta.a, supp.tail.meta.a_inv))
    E.add("join_supplier", *r5[1:])
    r7 = ops.matchjoinAN(r5[0].mirror_head(), t["lo_partkey"])
    E.add("fetch_partkey", *r7[1:])
    part = _select_dim(ops, t, E, "p_partkey", "p_category", (ops.EQ, category), "p_category")
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


def _join_dim_oids(ops, t: dict, E: Errors, rows, fact_fk: str, dim_key: str, tag: str):
    """fetchjoinAN(rows, fact FK) then hashjoinAN with the dimension key (re-encoding both sides):
    VOID | OID dimension for every row of `rows` (ssbm-q31_continuous.cpp:153-160)."""
    fk = ops.fetchjoinAN(rows, t[fact_fk])
    E.add(f"fetch_{tag}", *fk[1:])
    j = ops.hashjoinAN(fk[0], t[dim_key].reverse(), (A_OIDX, A_OIDX_INV), (A_OIDX, A_OIDX_INV))
    E.add(f"join_{tag}", *j[1:])
    return j[0].clear_head()


def p3(t: dict, ops=B, c_region: int = 2, s_region: int = 2, year_lo: int = 1992, year_hi: int = 1997) -> dict:
    """Q3.1 (ssbm-q31_continuous.cpp:89-212)."""
    E = Errors()
    supp = _select_dim(ops, t, E, "s_suppkey", "s_region", (ops.EQ, s_region), "s_region")
    r5 = ops.hashjoinAN(t["lo_suppkey"], supp, (A_OIDX, A_OIDX_INV), (supp.tail.meta.a, supp.tail.meta.a_inv))
    E.add("join_supplier", *r5[1:])
    cust = _select_dim(ops, t, E, "c_custkey", "c_region", (ops.EQ, c_region), "c_region")
    r11 = ops.matchjoinAN(r5[0].mirror_head(), t["lo_custkey"])
    E.add("fetch_custkey", *r11[1:])
    r12 = ops.hashjoinAN(r11[0], cust)
    E.add("join_customer", *r12[1:])
    date = _select_dim(ops, t, E, "d_datekey", "d_year", (ops.GE, year_lo, ops.LE, year_hi, ops.AND), "d_year")
    r18 = ops.matchjoinAN(r12[0].mirror_head(), t["lo_orderdate"])
    E.add("fetch_orderdate", *r18[1:])
    r19 = ops.hashjoinAN(r18[0], date)
    E.add("join_date", *r19[1:])
    rows = r19[0].mirror_head().clear_head()
    ar = ops.fetchjoinAN(rows, t["lo_revenue"])
    E.add("fetch_revenue", *ar[1:])
    s_oid = _join_dim_oids(ops, t, E, rows, "lo_suppkey", "s_suppkey", "suppkey")
    as_ = ops.fetchjoinAN(s_oid, t["s_nation"])
    E.add("fetch_s_nation", *as_[1:])
    c_oid = _join_dim_oids(ops, t, E, rows, "lo_custkey", "c_custkey", "custkey")
    ac = ops.fetchjoinAN(c_oid, t["c_nation"])
    E.add("fetch_c_nation", *ac[1:])
    d_oid = _join_dim_oids(ops, t, E, rows, "lo_orderdate", "d_datekey", "orderdate")
    ad = ops.fetchjoinAN(d_oid, t["d_year"])
    E.add("fetch_year", *ad[1:])
    g = _groups(ops, E, [ad[0], as_[0], ac[0]], ar[0])
    g.update({"rows": ar[0].size(), "errors": E})
    return g


def p4(t: dict, ops=B, c_region: int = 1, s_region: int = 1, mfgr: tuple = (0, 1)) -> dict:
    """Q4.1 (ssbm-q41_continuous.cpp:98-212)."""
    E = Errors()
    part = _select_dim(ops, t, E, "p_partkey", "p_mfgr", (ops.EQ, mfgr[0], ops.EQ, mfgr[1], ops.OR), "p_mfgr")
    r5 = ops.hashjoinAN(t["lo_partkey"], part, (A_OIDX, A_OIDX_INV), (part.tail.meta.a, part.tail.meta.a_inv))
    E.add("join_part", *r5[1:])
    bat6 = r5[0].mirror_head()
    supp = _select_dim(ops, t, E, "s_suppkey", "s_region", (ops.EQ, s_region), "s_region")
    r11 = ops.matchjoinAN(bat6, t["lo_suppkey"])
    E.add("fetch_suppkey", *r11[1:])
    r12 = ops.hashjoinAN(r11[0], supp)
    E.add("join_supplier", *r12[1:])
    cust = _select_dim(ops, t, E, "c_custkey", "c_region", (ops.EQ, c_region), "c_region")
    r18 = ops.matchjoinAN(r12[0].mirror_head(), t["lo_custkey"])
    E.add("fetch_custkey", *r18[1:])
    r19 = ops.hashjoinAN(r18[0], cust)
    E.add("join_customer", *r19[1:])
    rows = r19[0].mirror_head().clear_head()
    ar = ops.fetchjoinAN(rows, t["lo_revenue"])
    E.add("fetch_revenue", *ar[1:])
    asc = ops.fetchjoinAN(rows, t["lo_supplycost"])
    E.add("fetch_supplycost", *asc[1:])
    ap = ops.arithmeticAN(ops.SUB, ar[0], asc[0], a_result=A_MEAS, result_type=an.RESINT)
    E.add("profit", *ap[1:])
    c_oid = _join_dim_oids(ops, t, E, rows, "lo_custkey", "c_custkey", "custkey")
    an_ = ops.fetchjoinAN(c_oid, t["c_nation"])
    E.add("fetch_c_nation", *an_[1:])
    d_oid = _join_dim_oids(ops, t, E, rows, "lo_orderdate", "d_datekey", "orderdate")
    ay = ops.fetchjoinAN(d_oid, t["d_year"])
    E.add("fetch_year", *ay[1:])
    g = _groups(ops, E, [ay[0], an_[0]], ap[0])
    g.update({"rows": ap[0].size(), "errors": E})
    return g


PIPELINES = {"P1": p1, "P2": p2, "P3": p3, "P4": p4}


# ------------------------------------------------------------------------------------------------
# textbook evaluation of the same queries on decoded host values (fault-free data), for tests
# ------------------------------------------------------------------------------------------------

def textbook(spec: SsbSpec, name: str, **q) -> dict:
    cache = {}
    v = lambda c: host_values(spec, c, cache)  # noqa: E731
    if name == "P1":
        year = v("d_year")[v("lo_orderdate").astype(np.int64)]
        qty, d = v("lo_quantity"), v("lo_discount")
        m = (year == q.get("year", 1993)) & (d >= q.get("disc_lo", 1)) & (d <= q.get("disc_hi", 3)) & \
            (qty < q.get("qty_lt", 25))
        return {"result": int((v("lo_extendedprice")[m] * d[m]).sum()) & an.MASK64, "rows": int(m.sum())}
    fk = lambda c: v(c).astype(np.int64)  # noqa: E731
    if name == "P2":
        m = (v("s_region")[fk("lo_suppkey")] == q.get("s_region", 1)) & \
            (v("p_category")[fk("lo_partkey")] == q.get("category", 12))
        keys = [v("d_year")[fk("lo_orderdate")][m], v("p_brand")[fk("lo_partkey")][m]]
        meas = v("lo_revenue")[m]
    elif name == "P3":
        yr = v("d_year")[fk("lo_orderdate")]
        m = (v("c_region")[fk("lo_custkey")] == q.get("c_region", 2)) & \
            (v("s_region")[fk("lo_suppkey")] == q.get("s_region", 2)) & (yr >= q.get("year_lo", 1992)) & \
            (yr <= q.get("year_hi", 1997))
        keys = [yr[m], v("s_nation")[fk("lo_suppkey")][m], v("c_nation")[fk("lo_custkey")][m]]
        meas = v("lo_revenue")[m]
    else:
        mf = v("p_mfgr")[fk("lo_partkey")]
        mp = q.get("mfgr", (0, 1))
        m = (v("c_region")[fk("lo_custkey")] == q.get("c_region", 1)) & \
            (v("s_region")[fk("lo_suppkey")] == q.get("s_region", 1)) & ((mf == mp[0]) | (mf == mp[1]))
        keys = [v("d_year")[fk("lo_orderdate")][m], v("c_nation")[fk("lo_custkey")][m]]
        meas = (v("lo_revenue")[m].astype(np.int64) - v("lo_supplycost")[m].astype(np.int64)).astype(np.uint64)
    # groups in first-occurrence order of the (key tuple) over the qualifying rows in fact order
    order = {}
    sums = []
    for i in range(meas.size):
        k = tuple(int(x[i]) for x in keys)
        g = order.setdefault(k, len(order))
        if g == len(sums):
            sums.append(0)
        sums[g] = (sums[g] + int(meas[i])) & an.MASK64
    ks = list(order.keys())
    return {"sums": np.array(sums, dtype=np.uint64), "keys": [np.array([k[j] for k in ks], dtype=np.uint64)
                                                              for j in range(len(keys))],
            "groups": len(ks), "rows": int(meas.size)}
