"""Tables in AHEAD's format loaded into device BATs: codes arrive unchanged, re-encoding to a
minimum-bit-flip-weight A while loading (--AN-minbfw) keeps every decoded value and every
detected corruption, and operators run on the loaded columns exactly as on the oracle."""
import numpy as np
import pytest

import oracle_lib as O
from ahead_amd import an, bat, capi, datagen, loader

pytestmark = pytest.mark.gpu


def test_load_codes_and_reencode_on_load(tmp_path):
    n = 200_003
    qv, pv = datagen.uniform(n, 5, 1, 1, 50), datagen.uniform(n, 5, 2, 0, 10**7)
    q = datagen.encode(qv, 233, 2)
    p = datagen.encode(pv, 65117, 8)
    flipped = datagen.inject_flips(q, 5, 3, datagen.flip_threshold(1e-3))
    loader.write_table(tmp_path / "lo", {"quantity": ("RESTINY", q), "price": ("RESINT", p), "tag": ("CHAR", (q & 0x7F).astype(np.uint8))})
    t = loader.load_table(tmp_path / "lo")
    assert set(t) == {"quantity", "price"}
    assert np.array_equal(t["quantity"].tail.to_numpy(), q) and np.array_equal(t["price"].tail.to_numpy(), p)
    # re-encode to AsBFW[1] (= 3) for restiny and AsBFW[3] (= 881) for resint while loading
    r = loader.load_table(tmp_path / "lo", min_bfw={"restiny": 1, "resint": 3})
    mq, mp = r["quantity"].tail.meta, r["price"].tail.meta
    assert mq.a == an.RESTINY_AS_BFW[1] and mp.a == an.RESINT_AS_BFW[3]
    got_q = r["quantity"].tail.to_numpy()
    exp_q = ((q.astype(np.uint64) * np.uint64(an.inverse(233, 16) * mq.a % 65536)) & np.uint64(0xFFFF)).astype(np.uint16)
    assert np.array_equal(got_q, exp_q)
    dec_p = r["price"].tail.to_numpy() * np.uint64(mp.a_inv)
    assert np.array_equal(dec_p, pv)
    # checks of the re-encoded codes agree with the oracle's (a flip need not stay detectable:
    # A=3 has a lower bit-flip weight than 233)
    err = capi.check_an(r["quantity"].tail)
    ref = O.check_an(O.Col(got_q, mq))
    assert np.array_equal(err, ref) and 0 < ref.size <= flipped.size
    # an operator chain on the loaded columns equals the oracle
    sel, e = bat.selectAN(t["quantity"], bat.LT, 25 * 233)
    ro, re, _ = O.select_an(O.Col(q, loader.column_meta(an.RESTINY)), 0, 25 * 233)
    assert np.array_equal(sel.head.to_numpy(), ro) and np.array_equal(e, re)
