"""The fused star query on the GPU (ahead_hip_star_query) against the oracle's (or_star_query): all
thirteen SSB queries, bit for bit in results, group order, first rows and every error vector, with
flips in every fact column and dimension attribute; against the GPU operator chains on data the
chains process (selection flips); and the plan-time statuses."""
import numpy as np
import pytest

import oracle_bat as OB
from ahead_amd import ssb

pytestmark = pytest.mark.gpu

SMALL = dict(n_fact=300_000, n_date=2556, n_cust=30_000, n_supp=2_000, n_part=14_000)
KEY_COLS = {"d_datekey", "c_custkey", "s_suppkey", "p_partkey"}
ALL_FLIPS = tuple(c for c in ssb.COLUMNS if c not in KEY_COLS)


def same(got, ref):
    assert got["rows"] == ref["rows"]
    if "result" in ref:
        assert got["result"] == ref["result"]
    else:
        assert got["groups"] == ref["groups"]
        assert np.array_equal(got["sums"], ref["sums"])
        for a, b in zip(got["keys"], ref["keys"]):
            assert np.array_equal(a, b)
    ge, re = dict(got["errors"].items), dict(ref["errors"].items)
    assert sorted(ge) == sorted(re)
    for k in re:
        assert np.array_equal(ge[k], re[k]), k


@pytest.mark.parametrize("name", list(ssb.QUERIES))
def test_star_query_matches_oracle_with_flips_everywhere(name):
    spec = ssb.SsbSpec(rate=1e-3, seed=31, flip_cols=ALL_FLIPS, **SMALL)
    q = ssb.QUERIES[name]
    print(f"star_query {name} vs oracle, {spec.n_fact} rows", flush=True)
    got = ssb.fused(ssb.device_tables(spec), q, fallback=False)
    ref = ssb.fused(OB.tables(spec), q, OB, fallback=False)
    assert got["status"] == 0 and ref["status"] == 0
    same(got, ref)
    assert np.array_equal(got["first_row"], ref["first_row"])
    assert got["errors"].total() > 0


@pytest.mark.parametrize("name", list(ssb.QUERIES))
def test_star_query_matches_gpu_chain(name):
    q = ssb.QUERIES[name]
    attrs = {d.attr for d in q.dims}
    flips = tuple(c for c, v in ssb.COLUMNS.items() if v[4] and c not in attrs)
    spec = ssb.SsbSpec(rate=1e-3, seed=32, flip_cols=flips, **SMALL)
    print(f"star_query {name} vs chain, {spec.n_fact} rows", flush=True)
    t = ssb.device_tables(spec)
    got = ssb.fused(t, q, fallback=False)
    ref = ssb.chain(t, q)
    assert got["status"] == 0
    same(got, ref)


def test_star_query_falls_back_to_the_chain_on_corrupt_keys():
    spec = ssb.SsbSpec(rate=5e-3, seed=33, flip_cols=("s_suppkey",), **SMALL)
    q = ssb.QUERIES["Q2.1"]
    print("star_query corrupt supplier keys", flush=True)
    t = ssb.device_tables(spec)
    assert ssb.fused(t, q, fallback=False)["status"] == 2
    got = ssb.fused(t, q)
    assert got["status"] == 2
    same(got, ssb.chain(t, q))


def test_star_query_empty_and_tiny_fact_tables():
    for n in (0, 1, 511, 513):
        spec = ssb.SsbSpec(rate=0.0, seed=34, n_fact=n, n_date=2556, n_cust=300, n_supp=50, n_part=140)
        print(f"star_query tiny fact {n}", flush=True)
        for name in ("Q1.1", "Q3.1", "Q4.1"):
            q = ssb.QUERIES[name]
            got = ssb.fused(ssb.device_tables(spec), q, fallback=False)
            ref = ssb.fused(OB.tables(spec), q, OB, fallback=False)
            same(got, ref)


def test_star_query_where_every_row_joins():
    """No dimension selection: every fact row joins, so each wave step stages more rows than one
    batch of 64 and finishes the earlier batches at once; flips in the measures, a key and a group
    attribute send rows of every batch to the exact path."""
    q = ssb.Query("q41", [ssb.Dim("supp", [], "s_nation"), ssb.Dim("date", [], "d_year", ssb.REFETCH)],
                  ["lo_revenue", "lo_supplycost"], ssb.SUBTRACT, groups=[1, 0])
    spec = ssb.SsbSpec(rate=4e-3, seed=35, flip_cols=("lo_revenue", "lo_supplycost", "lo_orderdate", "s_nation"),
                       **SMALL)
    print("star_query every row joins", flush=True)
    got = ssb.fused(ssb.device_tables(spec), q, fallback=False)
    ref = ssb.fused(OB.tables(spec), q, OB, fallback=False)
    assert got["status"] == 0 and ref["status"] == 0
    same(got, ref)
    assert np.array_equal(got["first_row"], ref["first_row"])
    e = dict(ref["errors"].items)
    assert e["fused_meas0.e1"].size > 100 and e["fused_attr0.e1"].size > 100


def test_star_query_key_index_statuses():
    """A key index built over a dimension with a repeated or a corrupt key reports it, and the
    queries over that dimension fall back to the chain with that status."""
    from ahead_amd import an, bat as B, capi

    spec = ssb.SsbSpec(rate=0.0, seed=36, **SMALL)
    t = ssb.device_tables(spec)
    q = ssb.QUERIES["Q2.1"]
    keys = t["s_suppkey"].tail.to_numpy().copy()
    meta = t["s_suppkey"].tail.meta
    print("star_query key index statuses", flush=True)
    assert B.key_index(t["s_suppkey"].tail).status == 0
    dup = keys.copy()
    dup[11] = dup[5]
    t2 = dict(t)
    t2["s_suppkey"] = B.BAT(capi.Column.void(dup.size), capi.Column.from_numpy(dup, meta))
    assert B.key_index(t2["s_suppkey"].tail).status == 3
    got = ssb.fused(t2, q)
    assert got["status"] == 3
    same(got, ssb.chain(t2, q))
    bad = keys.copy()
    bad[7] ^= np.uint64(1 << 45)
    t3 = dict(t)
    t3["s_suppkey"] = B.BAT(capi.Column.void(bad.size), capi.Column.from_numpy(bad, meta))
    assert B.key_index(t3["s_suppkey"].tail).status == 2
    assert ssb.fused(t3, q, fallback=False)["status"] == 2
    assert an.A_OID_ERR == 16089


# ---- shared scans (ahead_hip_star_query_batch) ------------------------------------------------

def test_batch_of_all_queries_matches_oracle_with_flips_everywhere():
    """All thirteen queries in one call: packed into shared scans of up to eight queries (four FK
    and two predicate columns each); every query's result, group order, first rows and error
    vectors equal the oracle's for that query alone."""
    spec = ssb.SsbSpec(rate=1e-3, seed=41, flip_cols=ALL_FLIPS, **SMALL)
    qs = list(ssb.QUERIES.values())
    print(f"star_query_batch of {len(qs)} queries vs oracle, {spec.n_fact} rows", flush=True)
    got = ssb.fused_batch(ssb.device_tables(spec), qs, fallback=False)
    host = OB.tables(spec)
    for name, q, g in zip(ssb.QUERIES, qs, got):
        ref = ssb.fused(host, q, OB, fallback=False)
        assert g["status"] == 0 and ref["status"] == 0, name
        same(g, ref)
        assert np.array_equal(g["first_row"], ref["first_row"]), name


@pytest.mark.parametrize("names", [("Q1.1", "Q2.1", "Q3.1", "Q4.1"), ("Q4.1", "Q4.1"), ("Q2.3", "Q1.3", "Q4.3"),
                                   ("Q3.4",)])
def test_batch_equals_single_queries(names):
    """Small batches (the serve shapes, a query twice, mixed measure kinds, one query) against the
    same queries run alone on the GPU, with flips in every column the chains survive."""
    spec = ssb.SsbSpec(rate=2e-3, seed=42, flip_cols=ALL_FLIPS, **SMALL)
    t = ssb.device_tables(spec)
    qs = [ssb.QUERIES[n] for n in names]
    print(f"star_query_batch {names}", flush=True)
    got = ssb.fused_batch(t, qs, fallback=False)
    for n, q, g in zip(names, qs, got):
        ref = ssb.fused(t, q, fallback=False)
        assert g["status"] == 0 and ref["status"] == 0, n
        same(g, ref)
        assert np.array_equal(g["first_row"], ref["first_row"]), n


def test_batch_where_every_row_joins_and_tiny_tables():
    every = ssb.Query("q41", [ssb.Dim("supp", [], "s_nation"), ssb.Dim("date", [], "d_year", ssb.REFETCH)],
                      ["lo_revenue", "lo_supplycost"], ssb.SUBTRACT, groups=[1, 0])
    spec = ssb.SsbSpec(rate=4e-3, seed=43, flip_cols=("lo_revenue", "lo_supplycost", "lo_orderdate", "s_nation"),
                       **SMALL)
    print("star_query_batch every row joins", flush=True)
    qs = [every, ssb.QUERIES["Q1.2"], every]
    got = ssb.fused_batch(ssb.device_tables(spec), qs, fallback=False)
    host = OB.tables(spec)
    for q, g in zip(qs, got):
        same(g, ssb.fused(host, q, OB, fallback=False))
    for n in (0, 1, 511, 513):
        tiny = ssb.SsbSpec(rate=0.0, seed=44, n_fact=n, n_date=2556, n_cust=300, n_supp=50, n_part=140)
        print(f"star_query_batch tiny fact {n}", flush=True)
        qs = [ssb.QUERIES[k] for k in ("Q1.1", "Q3.1", "Q4.1", "Q2.2")]
        got = ssb.fused_batch(ssb.device_tables(tiny), qs, fallback=False)
        host = OB.tables(tiny)
        for q, g in zip(qs, got):
            same(g, ssb.fused(host, q, OB, fallback=False))


def test_batch_reports_each_querys_status():
    """A query over a dimension whose key index reports a corrupt key gets that status (and the
    chain's result with fallback); the others in the batch are unaffected."""
    from ahead_amd import bat as B, capi

    spec = ssb.SsbSpec(rate=0.0, seed=45, **SMALL)
    t = ssb.device_tables(spec)
    bad = t["s_suppkey"].tail.to_numpy().copy()
    bad[7] ^= np.uint64(1 << 45)
    t3 = dict(t)
    t3["s_suppkey"] = B.BAT(capi.Column.void(bad.size), capi.Column.from_numpy(bad, t["s_suppkey"].tail.meta))
    qs = [ssb.QUERIES[k] for k in ("Q1.1", "Q2.1", "Q4.1", "Q1.3")]  # Q2.x-Q4.x join the supplier
    print("star_query_batch statuses", flush=True)
    got = ssb.fused_batch(t3, qs)
    assert [g["status"] for g in got] == [0, 2, 2, 0]
    for q, g in zip(qs, got):
        same(g, ssb.chain(t3, q))


# ---- same-shape shared scans ------------------------------------------------------------------

def _instances(shape: str, n: int, seed: int) -> list:
    """n queries of one serve shape with seeded random constants."""
    from ahead_amd import serve

    qs = [q for q in serve.random_queries(12 * n, 0.0, seed) if q.shape == shape][:n]
    assert len(qs) == n
    return [q.spec() for q in qs]


@pytest.mark.parametrize("shape", ["P1", "P2", "P3", "P4"])
def test_same_shape_batch_matches_oracle_with_flips_everywhere(shape):
    """Eight queries of one shape with different constants share one scan (combined dimension maps,
    predicate pass-bit tables); each query's result, group order, first rows and error vectors
    equal the oracle's for that query alone, with flips in every fact column and attribute."""
    spec = ssb.SsbSpec(rate=1e-3, seed=51, flip_cols=ALL_FLIPS, **SMALL)
    qs = _instances(shape, 8, 52)
    print(f"star_query_batch 8 x {shape} vs oracle, {spec.n_fact} rows", flush=True)
    got = ssb.fused_batch(ssb.device_tables(spec), qs, fallback=False)
    host = OB.tables(spec)
    for i, (q, g) in enumerate(zip(qs, got)):
        ref = ssb.fused(host, q, OB, fallback=False)
        assert g["status"] == 0 and ref["status"] == 0, i
        same(g, ref)
        assert np.array_equal(g["first_row"], ref["first_row"]), i


def test_same_shape_passes_mixed_with_other_queries():
    """Eleven P2 queries (passes of eight and three), five P3 (two of them the same query), one P4
    (a union pass with the other thirteen queries' leftovers) and Q1.2: each equals the same query
    run alone on the GPU, with flips in every fact column and attribute."""
    spec = ssb.SsbSpec(rate=2e-3, seed=53, flip_cols=ALL_FLIPS, **SMALL)
    t = ssb.device_tables(spec)
    p3 = _instances("P3", 3, 55)
    qs = _instances("P2", 11, 54) + p3 + [p3[1], p3[1]] + _instances("P4", 1, 56) + [ssb.QUERIES["Q1.2"]]
    print(f"star_query_batch of {len(qs)} queries (same-shape passes) vs single queries", flush=True)
    got = ssb.fused_batch(t, qs, fallback=False)
    for i, (q, g) in enumerate(zip(qs, got)):
        ref = ssb.fused(t, q, fallback=False)
        assert g["status"] == 0 and ref["status"] == 0, i
        same(g, ref)
        assert np.array_equal(g["first_row"], ref["first_row"]), i


def test_same_shape_batch_tiny_fact_tables_and_repeated_keys():
    """Same-shape passes over fact tables of 0, 1, 511 and 513 rows against the oracle; over a
    supplier dimension with a repeated key every query of the pass reports it (and gets the
    chain's result with fallback)."""
    from ahead_amd import bat as B, capi

    for n in (0, 1, 511, 513):
        tiny = ssb.SsbSpec(rate=0.0, seed=57, n_fact=n, n_date=2556, n_cust=300, n_supp=50, n_part=140)
        print(f"star_query_batch same shape, tiny fact {n}", flush=True)
        qs = _instances("P3", 3, 58) + _instances("P1", 2, 59)
        got = ssb.fused_batch(ssb.device_tables(tiny), qs, fallback=False)
        host = OB.tables(tiny)
        for q, g in zip(qs, got):
            same(g, ssb.fused(host, q, OB, fallback=False))
    spec = ssb.SsbSpec(rate=0.0, seed=60, **SMALL)
    t = ssb.device_tables(spec)
    dup = t["s_suppkey"].tail.to_numpy().copy()
    dup[11] = dup[5]
    t2 = dict(t)
    t2["s_suppkey"] = B.BAT(capi.Column.void(dup.size), capi.Column.from_numpy(dup, t["s_suppkey"].tail.meta))
    qs = _instances("P3", 3, 61)
    print("star_query_batch same shape, repeated supplier key", flush=True)
    got = ssb.fused_batch(t2, qs)
    assert [g["status"] for g in got] == [3, 3, 3]
    for q, g in zip(qs, got):
        same(g, ssb.chain(t2, q))


def test_same_shape_full_passes_equal_single_queries():
    """Full same-shape passes (AHEAD_SQ_MAX_SHAPE_PASS queries of P3 and of P4, and one more of
    each, which runs outside the pass) against the same queries run alone, with flips in every
    fact column and attribute."""
    from ahead_amd import capi

    spec = ssb.SsbSpec(rate=1e-3, seed=62, flip_cols=ALL_FLIPS, **SMALL)
    t = ssb.device_tables(spec)
    n = capi.SQ_MAX_SHAPE_PASS + 1
    qs = _instances("P3", n, 63) + _instances("P4", n, 64)
    print(f"star_query_batch {len(qs)} queries: full P3 and P4 passes and one more of each", flush=True)
    got = ssb.fused_batch(t, qs, fallback=False)
    for i, (q, g) in enumerate(zip(qs, got)):
        ref = ssb.fused(t, q, fallback=False)
        assert g["status"] == 0 and ref["status"] == 0, i
        same(g, ref)
        assert np.array_equal(g["first_row"], ref["first_row"]), i
