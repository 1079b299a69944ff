"""The reference's other strategies (ahead_amd.strategies) on the oracle's operators.

* Fault-free data: every strategy's result equals the hardened chain's and the textbook GROUP BY
  for all thirteen queries (closed form I2: on valid codes every operator is the textbook one).
* The vote follows vote_majority_value (ModularRedundant.hpp:111-180): DMR returns equal results
  and raises on a difference; TMR returns the first result another equals, else the second if the
  third equals it, else raises.
* A flip in one replica's values changes that replica's result (DMR raises) and TMR out-votes it.
* Early: the decode error vectors are the flipped rows of each column the query reads (closed
  form I1: every single-bit flip is detected), x A_OID_ERR.
* Late: flips in the selected columns leave the hardened chain's result and no decode error; a
  flipped measure is dropped by the late decode and the aggregation's size check raises."""
import dataclasses

import numpy as np
import pytest

import oracle_bat as OB
from ahead_amd import an, ssb
from ahead_amd import strategies as S

SMALL = dict(n_fact=40_000, n_date=2556, n_cust=1_200, n_supp=80, n_part=600)
AOE = np.uint64(an.A_OID_ERR)


@pytest.fixture(scope="module")
def clean():
    spec = ssb.SsbSpec(rate=0.0, seed=51, **SMALL)
    t = OB.tables(spec)
    return spec, t, S.plain_tables(t, OB)


@pytest.mark.parametrize("name", list(ssb.QUERIES))
def test_every_strategy_equals_the_hardened_chain(clean, name):
    spec, t, plain = clean
    q = ssb.QUERIES[name]
    ref = ssb.chain(t, q, OB)
    assert S.same(ref, ssb.textbook(spec, q))
    assert S.same(S.normal(plain, q, OB), ref)
    e = S.early(t, q, OB)
    assert S.same(e, ref) and e["errors"].total() == 0
    la = S.late(t, q, OB)
    assert S.same(la, ref) and la["errors"].total() == 0
    assert S.same(S.modular(S.replicas(plain, 2, OB), q, OB), ref)
    assert S.same(S.modular(S.replicas(plain, 3, OB), q, OB), ref)


def _res(v):
    return {"rows": 1, "result": v}


def test_vote_follows_vote_majority_value():
    assert S.vote([_res(5), _res(5)])["result"] == 5
    with pytest.raises(S.ModularityError):
        S.vote([_res(5), _res(6)])
    assert S.vote([_res(5), _res(6), _res(5)])["result"] == 5
    assert S.vote([_res(6), _res(5), _res(5)])["result"] == 5
    assert S.vote([_res(5), _res(5), _res(6)])["result"] == 5
    with pytest.raises(S.ModularityError):
        S.vote([_res(4), _res(5), _res(6)])
    g1 = {"rows": 3, "groups": 2, "sums": np.array([1, 2], np.uint64), "keys": [np.array([7, 8], np.uint64)]}
    g2 = dict(g1, keys=[np.array([7, 9], np.uint64)])
    assert not S.same(g1, g2) and S.same(g1, dict(g1))


def _joined_row(spec, q):
    """A fact row that passes every predicate and joins every dimension (fault-free values)."""
    cache = {}
    v = lambda c: ssb.host_values(spec, c, cache)  # noqa: E731
    m = np.ones(spec.n_fact, dtype=bool)
    for col, p in q.preds:
        m &= ssb._pred_mask(v(col), p)
    for dq in q.dims:
        fk, key = ssb.DIM_KEYS[dq.name]
        sel = np.ones(v(key).size, dtype=bool)
        for col, p in dq.sels:
            sel &= ssb._pred_mask(v(col), p)
        m &= sel[v(fk).astype(np.int64)]
    return int(np.flatnonzero(m)[0])


@pytest.mark.parametrize("name,col", [("Q1.1", "lo_extendedprice"), ("Q2.1", "lo_revenue"), ("Q4.1", "lo_supplycost")])
def test_a_flip_in_one_replica(clean, name, col):
    spec, t, plain = clean
    q = ssb.QUERIES[name]
    row = _joined_row(spec, q)
    ref = ssb.chain(t, q, OB)
    for k, bad in ((2, 1), (3, 0), (3, 2)):
        reps = S.replicas(S.replicas(plain, 2, OB)[1], k, OB)  # the shared tables stay untouched
        reps[bad][col].tail.arr[row] ^= np.uint64(1 << 3)
        assert not S.same(ssb.chain(reps[bad], q, OB), ref)
        if k == 2:
            with pytest.raises(S.ModularityError):
                S.modular(reps, q, OB)
        else:
            assert S.same(S.modular(reps, q, OB), ref)


@pytest.mark.parametrize("name", ["Q1.1", "Q2.1", "Q3.1", "Q4.1"])
def test_early_decode_errors_are_the_flipped_rows(name):
    q = ssb.QUERIES[name]
    spec = ssb.SsbSpec(rate=2e-3, seed=52, flip_cols=tuple(c for c in ssb.COLUMNS if c not in ssb.DIM_KEYS_SET),
                       **SMALL)
    host = ssb.host_tables(spec)
    _, E = S.early_decode(OB.tables(spec), q, OB)
    e = dict(E.items)
    for col in S.columns_read(q):
        expect = np.sort(host[col][2]).astype(np.uint64) * AOE
        assert np.array_equal(e.get(f"decode_{col}.e1", np.zeros(0, np.uint64)), expect), col
    assert all(f"decode_{c}.e1" in e for c in q.meas)


@pytest.mark.parametrize("name", ["Q1.2", "Q2.2", "Q3.1", "Q4.1"])
def test_late_with_flips_in_the_selected_columns(name):
    """Late runs the plain operators on the codes: a flipped code takes part in a predicate on its
    raw value, as in selectAN, and nothing it decodes is flipped: the hardened chain's result, no
    decode errors."""
    q = ssb.QUERIES[name]
    attrs = {d.attr for d in q.dims}
    flips = tuple(c for c, v in ssb.COLUMNS.items() if v[4] and c not in attrs)
    spec = ssb.SsbSpec(rate=2e-2, seed=53, flip_cols=flips, **SMALL)
    t = OB.tables(spec)
    ref = ssb.chain(t, q, OB)
    assert ref["errors"].total() > 0
    r = S.late(t, q, OB)
    assert S.same(r, ref) and r["errors"].total() == 0


def test_late_with_a_flipped_measure_fails_its_size_check():
    """A corrupt fetched measure is dropped by the late decode, so the measure and the grouping
    differ in length and the aggregation's size check raises (aggregateAN.tcc:54-56)."""
    q = ssb.QUERIES["Q2.1"]
    spec = ssb.SsbSpec(rate=5e-2, seed=54, flip_cols=("lo_revenue",), **SMALL)
    with pytest.raises(RuntimeError):
        S.late(OB.tables(spec), q, OB)


@pytest.mark.parametrize("name", list(ssb.QUERIES))
def test_reencoding_chain_equals_the_continuous_chain(name):
    """Re-encoding changes the codes of every intermediate, not the decoded results or the error
    positions: the continuous chain's result and every error vector, with selection flips."""
    q = ssb.QUERIES[name]
    attrs = {d.attr for d in q.dims}
    flips = tuple(c for c, v in ssb.COLUMNS.items() if v[4] and c not in attrs)
    spec = ssb.SsbSpec(rate=1e-2, seed=55, flip_cols=flips, **SMALL)
    t = OB.tables(spec)
    ref = ssb.chain(t, q, OB)
    r = S.continuous_reenc(t, q, OB)
    assert S.same(r, ref)
    er, eref = dict(r["errors"].items), dict(ref["errors"].items)
    assert sorted(er) == sorted(eref)
    for k in eref:
        assert np.array_equal(er[k], eref[k]), k


def test_reencoding_uses_a_different_a_per_operator():
    ops = S._ReencOps(OB)
    side8 = OB.void(3)
    side2 = OB.HCol(np.zeros(3, np.uint16), 3, an.an_meta(an.RESTINY, 233))
    a8 = [ops._a(side8)[0] for _ in range(15)]
    assert a8[:3] == [an.RESINT_AS[14], an.RESINT_AS[13], an.RESINT_AS[12]]
    assert 1 not in a8 and a8[14] == an.RESINT_AS[14]  # indexes 14 .. 1, then again
    assert [ops._a(side2)[0] for _ in range(2)] == [an.RESTINY_AS[6], an.RESTINY_AS[5]]
    a, ainv = ops._a(side2)
    assert (a * ainv) % (1 << 16) == 1
