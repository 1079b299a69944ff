"""Pins the oracle's star-join pipeline (join + fetch + 2-key group-by + grouped sum) on the CPU.

Fault-free data: the pipeline must equal the textbook relational result computed with numpy
(equi-join by array indexing, GROUP BY in first-occurrence order via np.unique).  With injected
single-bit flips: the FK and measure error vectors follow in closed form from the flip positions
(every single flip of these codes is detected), and dropped rows disappear from the sums.
"""
import numpy as np
import pytest

import oracle_lib as O
from ahead_amd import an, datagen, workloads as W

INV_R = an.inverse(an.A_RESULT, 64)
M64 = (1 << 64) - 1


def cols(h, spec):
    return (O.Col(h["fk"], spec.meta_key), O.Col(h["m"], spec.meta_m), O.Col(h["key"], spec.meta_key),
            O.Col(h["a1"], spec.meta_attr), O.Col(h["a2"], spec.meta_attr))


def textbook(h, spec, drop_rows=None):
    """numpy: join on dense keys, first-occurrence group-by over (a1, a2), sum of decoded measures."""
    inv_k, inv_m, inv_a = an.inverse(W.A_KEY, 64), an.inverse(W.A_MEASURE, 64), an.inverse(W.A_ATTR, 16)
    with np.errstate(over="ignore"):
        fk = h["fk"] * np.uint64(inv_k)
        m = h["m"] * np.uint64(inv_m)
    a1 = ((h["a1"].astype(np.uint32) * inv_a) & 0xFFFF).astype(np.uint64)
    a2 = ((h["a2"].astype(np.uint32) * inv_a) & 0xFFFF).astype(np.uint64)
    keep = np.ones(fk.size, bool) if drop_rows is None else ~drop_rows
    j = fk[keep].astype(np.int64)
    g1, g2, mv = a1[j], a2[j], m[keep]
    combo = g1 * np.uint64(256) + g2
    uniq, first, inverse = np.unique(combo, return_index=True, return_inverse=True)
    order = np.argsort(first)
    sums = np.zeros(uniq.size, dtype=object)
    np.add.at(sums, inverse, mv.astype(object))
    return {
        "sums": [int(sums[o]) for o in order],
        "first": first[order].astype(np.uint64),
        "a1": (uniq[order] >> np.uint64(8)).astype(np.uint64),
        "a2": (uniq[order] & np.uint64(255)).astype(np.uint64),
    }


def test_star1_fault_free_equals_textbook():
    spec = W.Star1Spec(n_fact=200_000, n_dim=3_000, rate=0.0, seed=5)
    h = W.star1_host(spec)
    r = O.star1_pipeline(*cols(h, spec))
    t = textbook(h, spec)
    assert len(r["sums"]) == len(t["sums"]) == 125
    dec = [(int(x) * INV_R) & M64 for x in r["sums"]]
    assert dec == t["sums"]
    assert np.array_equal(r["groid"], t["first"] * np.uint64(an.A_OID))
    assert np.array_equal(r["key1"], (t["a1"] * np.uint64(W.A_ATTR)) & np.uint64(0xFFFF))
    assert np.array_equal(r["key2"], (t["a2"] * np.uint64(W.A_ATTR)) & np.uint64(0xFFFF))
    for k in ("e_fk", "e_dk", "e_a1", "e_a2", "e_m", "e_grp"):
        assert r[k].size == 0


def test_star1_flips_closed_form():
    spec = W.Star1Spec(n_fact=300_000, n_dim=2_000, rate=2e-4, seed=9)
    h = W.star1_host(spec)
    r = O.star1_pipeline(*cols(h, spec))
    fl = h["flips"]
    assert fl["fk"].size > 10 and fl["m"].size > 10
    # every single flip of a 64-bit code with A=63877 / 65117 and maxU=2^32-1 is detected
    assert np.array_equal(r["e_fk"], fl["fk"] * np.uint64(an.A_OID_ERR))
    assert np.array_equal(r["e_dk"], fl["key"] * np.uint64(an.A_OID_ERR))
    # joined rows: truncated decoded FK hits a dimension key (corrupt dimension keys keep their row
    # only under their garbage key)
    inv_k = an.inverse(W.A_KEY, 64)
    with np.errstate(over="ignore"):
        key32 = ((h["fk"] * np.uint64(inv_k)) & np.uint64(0xFFFFFFFF)).astype(np.int64)
        dkey32 = ((h["key"] * np.uint64(inv_k)) & np.uint64(0xFFFFFFFF)).astype(np.int64)
    row_of = {int(k): j for j, k in enumerate(dkey32)}
    jrow = np.array([row_of.get(int(k), -1) for k in key32])
    joined = jrow >= 0
    bad_m = np.zeros(spec.n_fact, bool)
    bad_m[fl["m"].astype(np.int64)] = True
    assert np.array_equal(r["e_m"], np.nonzero(joined & bad_m)[0].astype(np.uint64) * np.uint64(an.A_OID_ERR))
    bad_a1 = np.zeros(spec.n_dim, bool)
    bad_a1[fl["a1"].astype(np.int64)] = True
    hits = joined & bad_a1[np.where(joined, jrow, 0)]
    assert np.array_equal(r["e_a1"], jrow[hits].astype(np.uint64) * np.uint64(an.A_OID_ERR))
    # the sums cover exactly the kept rows
    drop = ~joined | bad_m | hits
    bad_a2 = np.zeros(spec.n_dim, bool)
    bad_a2[fl["a2"].astype(np.int64)] = True
    drop |= joined & bad_a2[np.where(joined, jrow, 0)]
    kept_dec = int(((h["m"][~drop] * np.uint64(an.inverse(W.A_MEASURE, 64)))).astype(object).sum())
    total = sum((int(x) * INV_R) & M64 for x in r["sums"])
    assert total == kept_dec


def test_star1_empty_fact():
    spec = W.Star1Spec(n_fact=0, n_dim=100, rate=0.0, seed=1)
    h = W.star1_host(spec)
    r = O.star1_pipeline(*cols(h, spec))
    assert r["sums"].size == 0
