"""The star-schema pipelines (ahead_amd.ssb: the Q1.1 / Q2.1 / Q3.1 / Q4.1 operator chains) run on
the oracle's operators: on fault-free data each equals the textbook evaluation of its query on
decoded values.  This pins the chains and the oracle operators together, on the CPU."""
import numpy as np
import pytest

import oracle_bat as OB
from ahead_amd import ssb

SMALL = dict(n_fact=60_000, n_date=2556, n_cust=1_500, n_supp=100, n_part=700)


@pytest.mark.parametrize("name", ["P1", "P2", "P3", "P4"])
def test_oracle_pipeline_matches_textbook(name):
    spec = ssb.SsbSpec(rate=0.0, seed=3, **SMALL)
    got = ssb.PIPELINES[name](OB.tables(spec), ops=OB)
    ref = ssb.textbook(spec, name)
    assert got["errors"].total() == 0
    assert got["rows"] == ref["rows"] and ref["rows"] > 0
    if name == "P1":
        assert got["result"] == ref["result"]
    else:
        assert got["groups"] == ref["groups"]
        assert np.array_equal(got["sums"], ref["sums"])
        for a, b in zip(got["keys"], ref["keys"]):
            assert np.array_equal(a, b)


def test_oracle_pipeline_reports_flipped_quantities():
    spec = ssb.SsbSpec(rate=1e-3, seed=4, **SMALL)
    host = ssb.host_tables(spec)
    got = ssb.p1(OB.tables(spec), ops=OB)
    e = dict(got["errors"].items)
    flipped = host["lo_quantity"][2]
    assert flipped.size > 0
    assert np.array_equal(e["select_lo_quantity.e1"], flipped.astype(np.uint64) * np.uint64(16089))
