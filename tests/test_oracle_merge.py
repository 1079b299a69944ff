"""Pins of the oracle's merge-form matchjoinAN (matchjoinAN.tcc:46-143).

Three independent checks:
  * closed forms on sorted inputs: with no corruption the merge is the order-preserving
    intersection of the two key lists, pairing the k-th occurrence of a key in list 1 with its
    k-th occurrence in list 2;
  * a small case worked by hand from the reference's loop nest, where every list-2 row is flagged
    so the error vectors show exactly where the loops check rows (including the list-2 row that
    is checked twice, at the top of an outer iteration and again when skipped);
  * an explicit three-state machine (top of loop / skip list 1 / skip list 2), the form the GPU
    kernel walks per chunk, must agree with the oracle's loop nest on random, unsorted and
    corrupted inputs.
"""
from dataclasses import replace

import numpy as np
import pytest

import oracle_lib as O
from ahead_amd import an, datagen

A_OID = an.A_OID_ERR
RESOID = an.an_meta(an.RESOID, an.A_OID)


def enc(vals, a=an.A_OID):
    with np.errstate(over="ignore"):
        return np.asarray(vals, dtype=np.uint64) * np.uint64(a)


def sorted_keys(n, hi, seed, stream):
    return np.sort(datagen.uniform(n, seed, stream, 0, hi))


def expected_intersection(k1, k2):
    """(i, j) index pairs: k-th occurrence of a key in list 1 with its k-th occurrence in list 2."""
    pairs = []
    i = j = 0
    while i < len(k1) and j < len(k2):
        if k1[i] < k2[j]:
            i += 1
        elif k1[i] > k2[j]:
            j += 1
        else:
            pairs.append((i, j))
            i += 1
            j += 1
    return pairs


@pytest.mark.parametrize("n1,n2,hi", [(0, 5, 10), (5, 0, 10), (1, 1, 0), (1000, 700, 5000), (3000, 3000, 400),
                                      (10_000, 50, 100_000)])
def test_merge_sorted_closed_form(n1, n2, hi):
    k1 = sorted_keys(n1, hi, n1 + n2, 1)
    k2 = sorted_keys(n2, hi, n1 + n2, 2)
    h1 = enc(np.arange(n1) + 17)           # BAT<resoid, resoid>: head = row OIDs, tail = keys
    t2 = enc(np.arange(n2) + 1000)
    ref = O.matchjoin_merge_an(O.Col(h1, RESOID), O.Col(enc(k1), RESOID), O.Col(enc(k2), RESOID),
                               O.Col(t2, RESOID))
    pairs = expected_intersection(k1.tolist(), k2.tolist())
    assert np.array_equal(ref[0], np.array([h1[i] for i, _ in pairs], dtype=np.uint64))
    assert np.array_equal(ref[1], np.array([t2[j] for _, j in pairs], dtype=np.uint64))
    for e in ref[2:]:
        assert e.size == 0
    if n1 and n2 and hi > 10 and n1 == n2:
        # strictly increasing lists: the pairs are exactly the common keys
        common = np.intersect1d(np.unique(k1), np.unique(k2))
        assert len(pairs) >= common.size


def test_merge_hand_worked_check_positions():
    # list 1 keys t1 = [3, 7], list 2 keys h2 = [1, 2, 7].  Codes use A = 1 so decoding is the
    # identity, and maxU = 0 flags every nonzero value: the error vectors then list each visit.
    # Walk (matchjoinAN.tcc:95-134): top(j=0) checks row2 0; 3 < 1 no; skip list 2 while 3 > h2:
    # checks row2 0 again, row2 1, stops at j=2 (3 < 7); top(j=2) checks row2 2; skip list 1 while
    # t1 < 7: checks row1 0; 7 == 7 emits (row1 1, row2 2); both lists end.
    flag_all = an.ColMeta(replace(an.RESINT, max_u=0), 1, 1)
    h1 = np.array([11, 12], dtype=np.uint64)
    t1 = np.array([3, 7], dtype=np.uint64)
    h2 = np.array([1, 2, 7], dtype=np.uint64)
    t2 = np.array([21, 22, 23], dtype=np.uint64)
    out_h, out_t, e1, e2, e3, e4 = O.matchjoin_merge_an(O.Col(h1, flag_all), O.Col(t1, flag_all),
                                                        O.Col(h2, flag_all), O.Col(t2, flag_all))
    assert out_h.tolist() == [12] and out_t.tolist() == [23]
    assert e1.tolist() == [0] and e2.tolist() == [0]
    assert e3.tolist() == [0, 0, A_OID, 2 * A_OID]
    assert e4.tolist() == [0, 0, A_OID, 2 * A_OID]


def machine(h1, t1, h2, t2, dec1, dec2, bad_h1, bad_t1, bad_h2, bad_t2, a_oid=A_OID):
    """The walk as an explicit state machine: TOP -> SKIP1 -> SKIP2 -> (emit) -> TOP."""
    n1, n2 = len(t1), len(h2)
    out, e = [], [[], [], [], []]
    i = j = 0
    state = "TOP"
    while True:
        if state == "TOP":
            if i >= n1 or j >= n2:
                break
            if bad_h2[j]:
                e[2].append(j * a_oid)
            if bad_t2[j]:
                e[3].append(j * a_oid)
            state = "SKIP1"
        if state == "SKIP1":
            while i < n1 and dec1[i] < dec2[j]:
                if bad_h1[i]:
                    e[0].append(i * a_oid)
                if bad_t1[i]:
                    e[1].append(i * a_oid)
                i += 1
            if i >= n1:
                break
            state = "SKIP2"
        if state == "SKIP2":
            while j < n2 and dec1[i] > dec2[j]:
                if bad_h2[j]:
                    e[2].append(j * a_oid)
                if bad_t2[j]:
                    e[3].append(j * a_oid)
                j += 1
            if j >= n2:
                break
            if dec1[i] == dec2[j]:
                out.append((h1[i], t2[j]))
                i += 1
                j += 1
            state = "TOP"
    return out, [[x & an.MASK64 for x in v] for v in e]


@pytest.mark.parametrize("seed", range(12))
def test_merge_state_machine_agrees(seed):
    rng = np.random.default_rng(seed)
    n1, n2 = int(rng.integers(0, 300)), int(rng.integers(0, 300))
    mode = seed % 3
    if mode == 0:      # sorted with duplicates
        k1, k2 = np.sort(rng.integers(0, 60, n1)), np.sort(rng.integers(0, 60, n2))
    elif mode == 1:    # unsorted
        k1, k2 = rng.integers(0, 40, n1), rng.integers(0, 40, n2)
    else:              # sorted, then a few codes corrupted
        k1, k2 = np.sort(rng.integers(0, 1000, n1)), np.sort(rng.integers(0, 1000, n2))
    meta = an.an_meta(an.RESSHORT, 233)
    c1, c2 = datagen.encode(k1, 233, 4), datagen.encode(k2, 233, 4)
    if mode == 2:
        for c in (c1, c2):
            for p in rng.integers(0, max(len(c), 1), 3):
                if len(c):
                    c[p] ^= np.uint32(1 << int(rng.integers(0, 32)))
    h1 = datagen.encode(rng.integers(0, 1 << 20, n1), an.A_OID, 8)
    t2 = datagen.encode(rng.integers(0, 1 << 20, n2), an.A_OID, 8)
    if n1:
        h1[0] ^= np.uint64(1 << 40)
    ref = O.matchjoin_merge_an(O.Col(h1, RESOID), O.Col(c1, meta), O.Col(c2, meta), O.Col(t2, RESOID))
    inv = an.inverse(233, 32)
    dec1 = [(int(x) * inv) & 0xFFFFFFFF for x in c1]
    dec2 = [(int(x) * inv) & 0xFFFFFFFF for x in c2]
    inv_o = an.inverse(an.A_OID, 64)
    bad_o = lambda arr: [((int(x) * inv_o) & an.MASK64) > an.RESOID.max_u for x in arr]
    out, e = machine([int(x) for x in h1], dec1, dec2, [int(x) for x in t2], dec1, dec2, bad_o(h1),
                     [d > 0xFFFF for d in dec1], [d > 0xFFFF for d in dec2], bad_o(t2))
    assert ref[0].tolist() == [p[0] for p in out]
    assert ref[1].tolist() == [p[1] for p in out]
    for k in range(4):
        assert ref[2 + k].tolist() == e[k], k
