"""Scratch data of one operator never reaches the scan state of the next.

A workspace's scratch regions keep whatever the last operator left there: hash-table words, row
counts, first-row arrays.  The decoupled look-back of every later launch trusts a status word
whose epoch equals its own, so its status words must live where no scratch data is ever written.
Here a hash join over a dense build side (keys 0..N-1, so its first-row array holds every small
integer and, read as status words, every small epoch) precedes fetch joins long enough that a
status array placed at the start of the scratch area would span that array; every result is
compared with the oracle.  It runs in a fresh thread, so the workspace is fresh and its epochs
start at 1, the range the first-row array covers.
"""
import threading

import numpy as np
import pytest

import oracle_bat as OB
from ahead_amd import an, bat, datagen

pytestmark = pytest.mark.gpu

N_BUILD = 8000
N_FACT = 17_000_000
RESINT = an.an_meta(an.RESINT, 65117)
RESSHORT = an.an_meta(an.RESSHORT, 233)
RESOID = an.an_meta(an.RESOID, an.A_OID)


def _inputs():
    keys = datagen.encode(np.arange(N_BUILD, dtype=np.uint64), 65117, 8)
    fk = datagen.encode(datagen.uniform(50_000, 31, 1, 0, N_BUILD - 1), 65117, 8)
    col = datagen.encode(datagen.uniform(N_FACT, 31, 2, 0, 0xFFFF), 233, 4)
    datagen.inject_flips(col, 31, 3, datagen.flip_threshold(1e-5))
    with np.errstate(over="ignore"):
        oids = np.arange(N_FACT, dtype=np.uint64)[::-1].copy() * np.uint64(an.A_OID)
    return keys, fk, col, oids


def _flat(res):
    out = []
    for x in res:
        if x is None:
            out.append(None)
        elif isinstance(x, np.ndarray):
            out.append(x.astype(np.uint64))
        else:
            out.append(np.concatenate([x.head_numpy().astype(np.uint64) if x.head.buf is not None
                                       else np.zeros(0, np.uint64), x.tail_numpy().astype(np.uint64)]))
    return out


def _oflat(res):
    out = []
    for x in res:
        if x is None:
            out.append(None)
        elif isinstance(x, np.ndarray):
            out.append(x.astype(np.uint64))
        else:
            h = x.head.arr.astype(np.uint64) if x.head.arr is not None else np.zeros(0, np.uint64)
            out.append(np.concatenate([h, x.tail.arr.astype(np.uint64)]))
    return out


def _same(got, ref, what):
    assert len(got) == len(ref), what
    for i, (g, r) in enumerate(zip(got, ref)):
        if r is None:
            assert g is None, (what, i)
        else:
            assert np.array_equal(g, r), (what, i)


def test_hash_join_scratch_then_long_fetch_joins():
    keys, fk, col, oids = _inputs()
    o_dim = OB.column(keys, RESINT).reverse()
    o_fact = OB.column(fk, RESINT)
    ref_join = _oflat(OB.hashjoinAN(o_fact, o_dim))
    ref_fetch = _oflat(OB.fetchjoinAN(OB.column(oids, RESOID), OB.column(col, RESSHORT)))

    failures = []

    def run():
        try:
            g_dim = bat.column(keys, RESINT).reverse()
            g_fact = bat.column(fk, RESINT)
            g_oids = bat.column(oids, RESOID)
            g_col = bat.column(col, RESSHORT)
            for rep in range(4):
                _same(_flat(bat.hashjoinAN(g_fact, g_dim)), ref_join, f"hashjoin rep {rep}")
                for k in range(3):
                    _same(_flat(bat.fetchjoinAN(g_oids, g_col)), ref_fetch, f"fetchjoin rep {rep}.{k}")
        except BaseException as e:  # reported in the test's own thread
            failures.append(e)

    t = threading.Thread(target=run)
    t.start()
    t.join()
    if failures:
        raise failures[0]
