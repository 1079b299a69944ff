"""The oracle's fused star query (or_star_query) pinned on the CPU.

* Fault-free data: for all thirteen SSB queries it equals the textbook GROUP BY / SUM over decoded
  values (numpy) and the query's operator chain run on the oracle's operators.
* Flips in the columns the queries select on: every result and every selectAN error vector equals
  the chain's.
* Flips in foreign keys the chains survive (a probed key keeps probing with its truncated value,
  hashjoinAN.tcc:199; a fetched key drops the row, matchjoinAN.tcc:287-295): results and the
  fused FK vectors equal the chain's join / fetch vectors.
* Flips in measures and group attributes (the chains throw there): the dropped rows and their
  positions follow from the fault-free textbook masks and the generator's flip positions.
* The cases where the fused form does not apply (repeated key, corrupt key, OID out of range)."""
import dataclasses

import numpy as np
import pytest

import oracle_bat as OB
import oracle_lib as O
from ahead_amd import an, datagen, ssb

SMALL = dict(n_fact=60_000, n_date=2556, n_cust=1_500, n_supp=100, n_part=700)
AOE = np.uint64(an.A_OID_ERR)


def same_result(got, ref):
    assert got["rows"] == ref["rows"]
    if "result" in ref:
        assert got["result"] == ref["result"]
        return
    assert got["groups"] == ref["groups"]
    assert np.array_equal(got["sums"], ref["sums"])
    for a, b in zip(got["keys"], ref["keys"]):
        assert np.array_equal(a, b)


def errs(res) -> dict:
    return {k: v for k, v in res["errors"].items}


@pytest.mark.parametrize("name", list(ssb.QUERIES))
def test_fused_equals_textbook_and_chain(name):
    spec = ssb.SsbSpec(rate=0.0, seed=3, **SMALL)
    t = OB.tables(spec)
    q = ssb.QUERIES[name]
    f = ssb.fused(t, q, OB, fallback=False)
    assert f["status"] == 0
    same_result(f, ssb.textbook(spec, q))
    same_result(f, ssb.chain(t, q, OB))
    assert f["errors"].total() == 0


def selection_flips(q) -> tuple:
    """The columns the schema flips, less the query's group attributes (a corrupt group attribute
    makes the chain throw: its fetch drops the row and groupbyAN sees a size mismatch)."""
    attrs = {d.attr for d in q.dims}
    return tuple(c for c, v in ssb.COLUMNS.items() if v[4] and c not in attrs)


@pytest.mark.parametrize("name", list(ssb.QUERIES))
def test_fused_equals_chain_with_selection_flips(name):
    q = ssb.QUERIES[name]
    spec = ssb.SsbSpec(rate=2e-2, seed=4, flip_cols=selection_flips(q), **SMALL)
    t = OB.tables(spec)
    f = ssb.fused(t, q, OB, fallback=False)
    c = ssb.chain(t, q, OB)
    same_result(f, c)
    same_result(f, ssb.textbook(spec, q))  # predicates on the stored (flipped) codes
    fe, ce = errs(f), errs(c)
    assert {k for k in ce if not k.startswith("select_")} == set()
    assert set(fe) == set(ce)
    for k in ce:
        assert np.array_equal(fe[k], ce[k]), k
    if q.preds:
        assert fe["select_lo_quantity.e1"].size > 0


# (query, flip rate, flipped FK columns, chain vector -> fused vector)
SURVIVING = [
    ("Q1.1", 3e-3, ("lo_orderdate",), {"fetch_orderdate.e4": "fused_fk0.e1"}),
    ("Q2.1", 3e-3, ("lo_suppkey", "lo_partkey", "lo_orderdate"),
     {"join_supplier.e2": "fused_fk0.e1", "fetch_partkey.e4": "fused_fk1.e1", "fetch_orderdate.e4": "fused_fk2.e1"}),
    ("Q3.1", 3e-3, ("lo_custkey", "lo_orderdate"),
     {"fetch_custkey.e4": "fused_fk1.e1", "fetch_orderdate.e4": "fused_fk2.e1"}),
    ("Q4.1", 3e-3, ("lo_partkey", "lo_suppkey", "lo_custkey"),
     {"join_lo_partkey.e2": "fused_fk0.e1", "fetch_suppkey.e4": "fused_fk1.e1", "fetch_custkey.e4": "fused_fk2.e1"}),
    ("Q4.3", 5e-2, ("lo_suppkey", "lo_orderdate"),
     {"fetch_suppkey.e4": "fused_fk1.e1", "fetch_orderdate.e4": "fused_fk2.e1"}),
]


@pytest.mark.parametrize("name,rate,cols,mapping", SURVIVING)
def test_fused_equals_chain_with_surviving_key_flips(name, rate, cols, mapping):
    spec = ssb.SsbSpec(rate=rate, seed=5, flip_cols=cols, **SMALL)
    t = OB.tables(spec)
    q = ssb.QUERIES[name]
    f = ssb.fused(t, q, OB, fallback=False)
    c = ssb.chain(t, q, OB)
    same_result(f, c)
    fe, ce = errs(f), errs(c)
    assert set(ce) == set(mapping), sorted(ce)
    assert set(fe) == set(mapping.values()), sorted(fe)
    for k, v in mapping.items():
        assert np.array_equal(fe[v], ce[k]), k


def _final_rows(spec, q):
    """Fact rows that pass every predicate and join every dimension, on the fault-free values."""
    clean = dataclasses.replace(spec, rate=0.0)
    cache = {}
    v = lambda c: ssb.host_values(clean, c, cache)  # noqa: E731
    m = np.ones(spec.n_fact, dtype=bool)
    for col, p in q.preds:
        m &= ssb._pred_mask(v(col), p)
    for dq in q.dims:
        fk, key = ssb.DIM_KEYS[dq.name]
        sel = np.ones(v(key).size, dtype=bool)
        for col, p in dq.sels:
            sel &= ssb._pred_mask(v(col), p)
        m &= sel[v(fk).astype(np.int64)]
    return m


@pytest.mark.parametrize("name", ["Q2.1", "Q3.1", "Q4.2"])
def test_fused_drops_rows_with_corrupt_measures(name):
    q = ssb.QUERIES[name]
    spec = ssb.SsbSpec(rate=5e-3, seed=6, flip_cols=tuple(q.meas), **SMALL)
    host = ssb.host_tables(spec)
    f = ssb.fused(OB.tables(spec), q, OB, fallback=False)
    final = _final_rows(spec, q)
    for m, col in enumerate(q.meas):
        flipped = host[col][2].astype(np.int64)
        expect = np.sort(flipped[final[flipped]]).astype(np.uint64) * AOE
        got = errs(f).get(f"fused_meas{m}.e1", np.zeros(0, np.uint64))
        assert np.array_equal(got, expect), col
        assert expect.size > 0
    bad = np.zeros(spec.n_fact, dtype=bool)
    for col in q.meas:
        bad[host[col][2].astype(np.int64)] = True
    # the textbook evaluation of the kept rows: fault-free values with the dropped rows removed
    keep_spec = dataclasses.replace(spec, rate=0.0)
    ref = ssb.textbook(keep_spec, q)
    assert f["rows"] == ref["rows"] - int((final & bad).sum())
    assert f["rows"] == int((final & ~bad).sum())


def test_fused_reports_corrupt_group_attributes_at_dimension_rows():
    q = ssb.QUERIES["Q3.1"]
    spec = ssb.SsbSpec(rate=5e-2, seed=7, flip_cols=("s_nation",), **SMALL)
    host = ssb.host_tables(spec)
    f = ssb.fused(OB.tables(spec), q, OB, fallback=False)
    cache = {}
    clean = dataclasses.replace(spec, rate=0.0)
    final = _final_rows(spec, q)
    supp = ssb.host_values(clean, "lo_suppkey", cache).astype(np.int64)
    flipped = np.zeros(spec.n_supp, dtype=bool)
    flipped[host["s_nation"][2].astype(np.int64)] = True
    hit = final & flipped[supp]
    expect = supp[hit].astype(np.uint64) * AOE  # one entry per fact row, in fact order
    assert expect.size > 0
    assert np.array_equal(errs(f)["fused_attr0.e1"], expect)
    assert f["rows"] == int((final & ~flipped[supp]).sum())


def _dim_args(spec, q, host):
    """Oracle star-query arguments built straight from host columns (no operators)."""
    cols = {k: O.Col(v[0], v[1]) for k, v in host.items()}
    dims = []
    for dq in q.dims:
        fk, key = ssb.DIM_KEYS[dq.name]
        dims.append({"fk": cols[fk], "key": cols[key], "attr": cols[dq.attr] if dq.attr else None, "sel": None,
                     "flags": dq.flags})
    return cols, dims


def test_star_query_status_when_fused_form_does_not_apply():
    spec = ssb.SsbSpec(rate=0.0, seed=8, **SMALL)
    q = ssb.QUERIES["Q4.1"]
    host = ssb.host_tables(spec)
    cols, dims = _dim_args(spec, q, host)
    meas = [cols[m] for m in q.meas]
    assert O.star_query(spec.n_fact, dims, [], meas, q.kind)["status"] == 0
    # a repeated key
    k2 = host["s_suppkey"][0].copy()
    k2[7] = k2[3]
    dims[1]["key"] = O.Col(k2, host["s_suppkey"][1])
    assert O.star_query(spec.n_fact, dims, [], meas, q.kind)["status"] == 3
    # a corrupt key
    k3 = host["s_suppkey"][0].copy()
    k3[5] ^= np.uint64(1 << 40)
    dims[1]["key"] = O.Col(k3, host["s_suppkey"][1])
    assert O.star_query(spec.n_fact, dims, [], meas, q.kind)["status"] == 2
    # a selected OID beyond the dimension
    dims[1]["key"] = cols["s_suppkey"]
    dims[1]["sel"] = np.array([3, spec.n_supp], dtype=np.uint64) * np.uint64(an.A_OID)
    assert O.star_query(spec.n_fact, dims, [], meas, q.kind)["status"] == 1


def test_flip_positions_are_detected_in_every_fk():
    """Closed form I1 at the fused level: with every FK flipped and no selections, each dimension
    probed straight on the fact column reports exactly the flipped rows."""
    spec = ssb.SsbSpec(rate=1e-3, seed=9, flip_cols=("lo_suppkey",), **SMALL)
    q = ssb.QUERIES["Q2.1"]
    host = ssb.host_tables(spec)
    f = ssb.fused(OB.tables(spec), q, OB, fallback=False)
    assert np.array_equal(errs(f)["fused_fk0.e1"], host["lo_suppkey"][2].astype(np.uint64) * AOE)
    assert datagen.flip_threshold(1e-3) > 0
