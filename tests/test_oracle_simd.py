"""The all-core vectorised CPU restatement (oracle/an_simd.cpp, the CPU baseline) equals the scalar
oracle bit for bit: selectAN + decoded sum over every code width, predicate and combiner, and the
fused star query over all thirteen SSB queries with flips in every fact column and attribute, on
one thread and on several (row ranges merged in order)."""
import numpy as np
import pytest

import oracle_bat as OB
import oracle_lib as O
from ahead_amd import an, datagen, ssb

SMALL = dict(n_fact=120_000, n_date=2556, n_cust=3_000, n_supp=300, n_part=1_400)
KEY_COLS = {"d_datekey", "c_custkey", "s_suppkey", "p_partkey"}
ALL_FLIPS = tuple(c for c in ssb.COLUMNS if c not in KEY_COLS)


def test_simd_level_is_reported():
    assert O.simd_level() in (0, 3, 4)


@pytest.mark.parametrize("width,a,rng", [(2, 29, (0, 255)), (4, 233, (1, 50)), (8, 65117, (90_000, 10_500_000))])
@pytest.mark.parametrize("pred", [(0, 30, 0, 0, 0), (3, 10, 1, 40, 1), (4, 7, 4, 9, 2), (5, 3, 0, 0, 0)])
@pytest.mark.parametrize("threads", [1, 5])
def test_select_simd_equals_oracle(width, a, rng, pred, threads):
    types = {2: an.RESTINY, 4: an.RESSHORT, 8: an.RESINT}
    meta = an.an_meta(types[width], a)
    n = 70_001
    codes = datagen.encode(datagen.uniform(n, 3, width, *rng), a, width)
    datagen.inject_flips(codes, 3, 9, datagen.flip_threshold(3e-3))
    op1, v1, op2, v2, comb = pred
    lo = rng[0] + v1 % (rng[1] - rng[0] + 1)
    hi = rng[0] + v2 % (rng[1] - rng[0] + 1)
    th1, th2 = (lo * a) & an.MASK64, (hi * a) & an.MASK64
    col = O.Col(codes, meta)
    ref = O.select_an(col, op1, th1, op2, th2, comb)
    got = O.select_an_simd(col, op1, th1, op2, th2, comb, threads=threads)
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]) and got[2] == ref[2]
    assert ref[1].size > 0


@pytest.mark.parametrize("name", list(ssb.QUERIES))
def test_star_query_simd_equals_oracle(name):
    spec = ssb.SsbSpec(rate=2e-3, seed=41, flip_cols=ALL_FLIPS, **SMALL)
    q = ssb.QUERIES[name]
    t = OB.tables(spec)
    ref = ssb.fused(t, q, OB, fallback=False)
    for threads in (1, 7):
        OB.SIMD_THREADS = threads
        try:
            got = ssb.fused(t, q, OB, fallback=False)
        finally:
            OB.SIMD_THREADS = None
        assert got["status"] == ref["status"] == 0
        assert got["rows"] == ref["rows"]
        assert np.array_equal(got["first_row"], ref["first_row"])
        if "result" in ref:
            assert got["result"] == ref["result"]
        else:
            assert np.array_equal(got["sums"], ref["sums"])
            for x, y in zip(got["keys"], ref["keys"]):
                assert np.array_equal(x, y)
        ge, re = dict(got["errors"].items), dict(ref["errors"].items)
        assert sorted(ge) == sorted(re)
        for k in re:
            assert np.array_equal(ge[k], re[k]), k
