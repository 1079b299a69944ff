"""The four star-schema pipelines (Q1.1, Q2.1, Q3.1, Q4.1 operator chains of the reference's
continuous drivers) on device BATs against the same chains run on the oracle's operators:
results, qualifying rows and every operator's error vector, bit for bit, with flips injected; and
on fault-free data against the textbook evaluation of each query."""
import numpy as np
import pytest

import oracle_bat as OB
from ahead_amd import ssb

pytestmark = pytest.mark.gpu

SMALL = dict(n_fact=250_000, n_date=2556, n_cust=3_000, n_supp=200, n_part=1_400)


def same(got, ref):
    assert got["rows"] == ref["rows"]
    if "result" in ref:
        assert got["result"] == ref["result"]
    else:
        assert got["groups"] == ref["groups"]
        assert np.array_equal(got["sums"], ref["sums"])
        for a, b in zip(got["keys"], ref["keys"]):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("name", ["P1", "P2", "P3", "P4"])
def test_pipeline_matches_textbook(name):
    spec = ssb.SsbSpec(rate=0.0, seed=5, **SMALL)
    got = ssb.PIPELINES[name](ssb.device_tables(spec))
    print(name, "rows", got["rows"], "groups", got.get("groups"), flush=True)
    assert got["errors"].total() == 0
    same(got, ssb.textbook(spec, name))


@pytest.mark.parametrize("name", ["P1", "P2", "P3", "P4"])
def test_pipeline_matches_oracle_chain_with_flips(name):
    spec = ssb.SsbSpec(rate=2e-3, seed=9, **SMALL)
    got = ssb.PIPELINES[name](ssb.device_tables(spec))
    ref = ssb.PIPELINES[name](OB.tables(spec), ops=OB)
    print(name, "rows", got["rows"], "errors", got["errors"].total(), flush=True)
    same(got, ref)
    ge, re = got["errors"].items, ref["errors"].items
    assert [k for k, _ in ge] == [k for k, _ in re]
    for (k, a), (_, b) in zip(ge, re):
        assert np.array_equal(a, b), k
    assert ref["errors"].total() > 0
