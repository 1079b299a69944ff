"""The reference's other strategies (ahead_amd.strategies) on the GPU operators against the same
strategies on the oracle's: unhardened chains, early and late check-and-decode, the re-encoding
chains, DMR / TMR with a replica whose values were changed, all thirteen queries."""
import numpy as np
import pytest

import oracle_bat as OB
from ahead_amd import ssb
from ahead_amd import strategies as S

pytestmark = pytest.mark.gpu

SMALL = dict(n_fact=200_000, n_date=2556, n_cust=3_000, n_supp=200, n_part=1_400)
SEL_FLIPS = tuple(c for c, v in ssb.COLUMNS.items() if v[4])


def same_errors(a, b):
    ea, eb = dict(a.items), dict(b.items)
    assert sorted(ea) == sorted(eb)
    for k in eb:
        assert np.array_equal(ea[k], eb[k]), k


@pytest.fixture(scope="module")
def schema():
    spec = ssb.SsbSpec(rate=0.0, seed=61, **SMALL)
    t = ssb.device_tables(spec)
    return spec, t, S.plain_tables(t), OB.tables(spec)


@pytest.mark.parametrize("name", list(ssb.QUERIES))
def test_unhardened_and_modular_strategies(schema, name):
    spec, t, plain, host = schema
    q = ssb.QUERIES[name]
    print(f"strategies {name}: normal, dmr, tmr", flush=True)
    ref = ssb.chain(host, q, OB)
    assert S.same(S.normal(plain, q), ref)
    assert S.same(S.modular(S.replicas(plain, 2), q), ref)
    assert S.same(S.modular(S.replicas(plain, 2), q, concurrent=True), ref)
    assert S.same(S.modular(S.replicas(plain, 3), q), ref)


@pytest.mark.parametrize("name", list(ssb.QUERIES))
def test_early_late_and_reencoding_with_flips(name):
    q = ssb.QUERIES[name]
    attrs = {d.attr for d in q.dims}
    flips = tuple(c for c in SEL_FLIPS if c not in attrs)
    spec = ssb.SsbSpec(rate=2e-3, seed=62, flip_cols=flips, **SMALL)
    t, host = ssb.device_tables(spec), OB.tables(spec)
    print(f"strategies {name}: early, late, re-encoding", flush=True)
    _, eg = S.early_decode(t, q)
    _, eo = S.early_decode(host, q, OB)
    same_errors(eg, eo)
    try:
        ro = S.early(host, q, OB)
    except RuntimeError:
        with pytest.raises(RuntimeError):
            S.early(t, q)
    else:
        rg = S.early(t, q)
        assert S.same(rg, ro)
    lg, lo = S.late(t, q), S.late(host, q, OB)
    assert S.same(lg, lo)
    same_errors(lg["errors"], lo["errors"])
    rg, ro = S.continuous_reenc(t, q), S.continuous_reenc(host, q, OB)
    assert S.same(rg, ro) and S.same(rg, ssb.chain(host, q, OB))
    same_errors(rg["errors"], ro["errors"])


def test_a_changed_replica_is_detected_and_outvoted(schema):
    spec, t, plain, host = schema
    q = ssb.QUERIES["Q2.1"]
    ref = ssb.chain(host, q, OB)
    rows = ssb.textbook(spec, q)["rows"]
    assert rows > 0
    print("strategies: changed replica", flush=True)
    for k, bad in ((2, 1), (3, 0), (3, 2)):
        reps = S.replicas(S.replicas(plain, 2)[1], k)
        col = reps[bad]["lo_revenue"].tail
        col.buf.t[: col.n * col.meta.width].view(-1, col.meta.width)[:, 1] ^= 1  # every revenue + 256
        assert not S.same(ssb.chain(reps[bad], q), ref)
        if k == 2:
            with pytest.raises(S.ModularityError):
                S.modular(reps, q, concurrent=True)
        else:
            assert S.same(S.modular(reps, q), ref)


def test_late_with_a_flipped_measure_fails_its_size_check():
    spec = ssb.SsbSpec(rate=5e-2, seed=63, flip_cols=("lo_revenue",), **SMALL)
    print("strategies: late with a flipped measure", flush=True)
    with pytest.raises(RuntimeError):
        S.late(ssb.device_tables(spec), ssb.QUERIES["Q2.1"])
