"""Pins the oracle's join / group / aggregate / arithmetic operators on the CPU.

Checks are independent formulations: numpy gathers for positional joins, a brute-force nested loop
for the hash join, np.unique first-occurrence order for group-by, np.add.at for grouped sums, plain
numpy products for the Σ-product and element-wise arithmetic, plus the closed-form error vectors
of injected single-bit flips.
"""
import numpy as np
import pytest

import oracle_lib as O
from ahead_amd import an, datagen

M64 = (1 << 64) - 1
RESOID = an.an_meta(an.RESOID, an.A_OID)


def enc_oids(pos):
    with np.errstate(over="ignore"):
        return pos.astype(np.uint64) * np.uint64(an.A_OID)


def test_fetchjoin_matches_numpy_gather():
    n_col, n = 5000, 20000
    d = datagen.uniform(n_col, 3, 1, 0, 255)
    col = datagen.encode(d, 29, 2)
    flips = datagen.inject_flips(col, 3, 2, datagen.flip_threshold(2e-2))
    pos = datagen.uniform(n, 3, 5, 0, n_col + 50)  # some positions beyond the column
    out, e1, e2 = O.fetchjoin_an(O.Col(enc_oids(pos), RESOID), O.Col(col, an.an_meta(an.RESTINY, 29)))
    bad = np.zeros(n_col, bool)
    bad[flips.astype(np.int64)] = True
    inr = pos < n_col
    p = pos.astype(np.int64)
    keep = inr & ~bad[np.where(inr, p, 0)]
    assert np.array_equal(out, col[p[keep]].astype(np.uint64))
    assert np.array_equal(e1, np.nonzero(~inr)[0].astype(np.uint64) * np.uint64(an.A_OID_ERR))
    corrupt = inr & bad[np.where(inr, p, 0)]
    assert np.array_equal(e2, pos[corrupt] * np.uint64(an.A_OID_ERR))
    # re-encoding: kept codes move to A=233 (codes of the same decoded values)
    out2, _, _ = O.fetchjoin_an(O.Col(enc_oids(pos), RESOID), O.Col(col, an.an_meta(an.RESTINY, 29)), 1, 233)
    assert np.array_equal(out2, (d[p[keep]] * np.uint64(233)) & np.uint64(0xFFFF))


def test_matchjoin_positional_matches_numpy():
    n_col, n = 3000, 8000
    vals = datagen.encode(datagen.uniform(n_col, 4, 1, 0, 0xFFFF), 233, 4)
    pos = datagen.uniform(n, 4, 2, 0, n_col - 1)
    heads = datagen.encode(datagen.uniform(n, 4, 3, 0, 1 << 20), an.A_OID, 8)
    hflips = datagen.inject_flips(heads, 4, 7, datagen.flip_threshold(1e-2))
    oh, ot, e1, e2, e4 = O.matchjoin_pos_an(O.Col(heads, RESOID), O.Col(enc_oids(pos), RESOID),
                                            O.Col(vals, an.an_meta(an.RESSHORT, 233)))
    assert np.array_equal(oh, heads) and np.array_equal(ot, vals[pos.astype(np.int64)].astype(np.uint64))
    # resoid maxU is 2^56-1: a head flip is detected iff its decoded value exceeds it
    with np.errstate(over="ignore"):
        dec = heads * np.uint64(an.A_OID_INV)
    assert np.array_equal(e1, np.nonzero(dec > np.uint64(an.RESOID.max_u))[0].astype(np.uint64) *
                          np.uint64(an.A_OID_ERR))
    assert e2.size == 0 and e4.size == 0


def brute_hashjoin(k1, p1, k2, p2):
    """nested loop: for each probe row in order, every build row with an equal key, in order"""
    out = []
    for i in range(len(k1)):
        for j in range(len(k2)):
            if k1[i] == k2[j]:
                out.append((p1[i], p2[j]))
    return out


@pytest.mark.parametrize("build_right", [1, 0])
def test_hashjoin_equals_nested_loop_with_duplicates(build_right):
    n1, n2 = 400, 300
    k1 = datagen.uniform(n1, 8, 1, 0, 120)
    k2 = datagen.uniform(n2, 8, 2, 0, 120)  # duplicate keys on both sides
    h1 = datagen.encode(datagen.uniform(n1, 8, 3, 0, 10**6), an.A_OID, 8)
    t2 = datagen.encode(datagen.uniform(n2, 8, 4, 0, 10**6), an.A_OID, 8)
    meta_k = an.an_meta(an.RESINT, 63877)
    c_h1 = O.Col(h1, RESOID)
    c_t1 = O.Col(datagen.encode(k1, 63877, 8), meta_k)
    c_h2 = O.Col(datagen.encode(k2, 63877, 8), meta_k)
    c_t2 = O.Col(t2, RESOID)
    oh, ot, e1, e2, e3, e4 = O.hashjoin_an(c_h1, c_t1, c_h2, c_t2, build_right)
    if build_right:
        ref = brute_hashjoin(list(k1), list(h1), list(k2), list(t2))
        assert list(zip(oh.tolist(), ot.tolist())) == [(int(a), int(b)) for a, b in ref]
    else:
        # left build: probe with the right rows in order, left rows in insertion order
        ref = brute_hashjoin(list(k2), list(t2), list(k1), list(h1))
        assert list(zip(ot.tolist(), oh.tolist())) == [(int(a), int(b)) for a, b in ref]
    assert all(x.size == 0 for x in (e1, e2, e3, e4))


def test_hashjoin_corrupt_probe_key_still_probes_truncated():
    # hashjoinAN.tcc:199-200: a corrupt key is reported and still looked up with its truncated value
    meta_k = an.an_meta(an.RESINT, 63877)
    inv = an.inverse(63877, 64)
    build_keys = np.arange(1000, dtype=np.uint64)
    # a corrupt code whose decoded value truncates to 7: c = (2^32 + 7) * 63877 mod 2^64
    bad = np.uint64((((1 << 32) + 7) * 63877) & M64)
    probe = np.array([datagen.encode(np.array([5], np.uint64), 63877, 8)[0], bad], dtype=np.uint64)
    assert ((int(bad) * inv) & M64) > 0xFFFFFFFF
    oh, ot, e1, e2, e3, e4 = O.hashjoin_an(O.void(2), O.Col(probe, meta_k),
                                           O.Col(datagen.encode(build_keys, 63877, 8), meta_k), O.void(1000))
    assert oh.tolist() == [0, 1] and ot.tolist() == [5, 7]
    assert e2.tolist() == [1 * an.A_OID_ERR]


def test_groupby_unary_first_occurrence_order():
    n = 50_000
    d = datagen.uniform(n, 12, 1, 0, 40)
    col = O.Col(datagen.encode(d, 29, 2), an.an_meta(an.RESTINY, 29))
    vgid, groid, eh, et = O.groupby_an(O.void(n, seqbase=100), col)
    uniq, first, inv = np.unique(d, return_index=True, return_inverse=True)
    order = np.argsort(first)
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    assert np.array_equal(vgid, rank[inv].astype(np.uint64) * np.uint64(an.A_OID))
    assert np.array_equal(groid, (first[order].astype(np.uint64) + np.uint64(100)) * np.uint64(an.A_OID))
    assert eh.size == 0 and et.size == 0


def test_groupby_binary_equals_pair_dictionary():
    n = 30_000
    a = datagen.uniform(n, 13, 1, 0, 9)
    b = datagen.uniform(n, 13, 2, 0, 6)
    g1 = O.groupby_an(O.void(n), O.Col(datagen.encode(a, 29, 2), an.an_meta(an.RESTINY, 29)))
    ng1 = g1[1].size
    vgid, groid, eh, et, eg = O.groupby2_an(O.void(n), O.Col(datagen.encode(b, 233, 2), an.an_meta(an.RESTINY, 233)),
                                            O.Col(g1[0], RESOID), ng1, 32)
    seen, ids, firsts = {}, [], []
    for i, pair in enumerate(zip(a.tolist(), b.tolist())):
        if pair not in seen:
            seen[pair] = len(seen)
            firsts.append(i)
        ids.append(seen[pair])
    assert np.array_equal(vgid, np.array(ids, np.uint64) * np.uint64(an.A_OID))
    assert np.array_equal(groid, np.array(firsts, np.uint64) * np.uint64(an.A_OID))
    assert eh.size == et.size == eg.size == 0


def test_groupby_binary_rgid_truncation_of_repeats():
    # more than 2^32 / 65117 groups: a repeated key returns the oid_t-truncated RGID
    n_groups = 70_000
    keys = np.concatenate([np.arange(n_groups, dtype=np.uint64), np.array([n_groups - 1], np.uint64)])
    tail = O.Col(datagen.encode(keys, 63877, 8), an.an_meta(an.RESINT, 63877))
    zero = O.Col(np.zeros(keys.size, np.uint64), RESOID)
    vgid, groid, *_ = O.groupby2_an(O.void(keys.size), tail, zero, 1, 64)
    g = n_groups - 1
    assert int(vgid[g]) == (g * an.A_OID) & M64
    assert int(vgid[-1]) == ((g * an.A_OID) & M64) & 0xFFFFFFFF
    assert (g * an.A_OID) >= (1 << 32)


def test_aggregate_sum_grouped_matches_add_at():
    n, G = 40_000, 37
    m = datagen.uniform(n, 14, 1, 90_000, 10_500_000)
    codes = datagen.encode(m, 65117, 8)
    gid = datagen.uniform(n, 14, 2, 0, G - 1)
    flips = datagen.inject_flips(codes, 14, 3, datagen.flip_threshold(1e-3))
    sums, eh, et, eg = O.aggregate_sum_grouped_an(O.void(n), O.Col(codes, an.an_meta(an.RESINT, 65117)),
                                                  O.Col(enc_oids(gid), RESOID), G)
    ref = np.zeros(G, dtype=object)
    with np.errstate(over="ignore"):
        contrib = codes * np.uint64((an.A_RESULT * an.inverse(65117, 64)) & M64)
    np.add.at(ref, gid.astype(np.int64), contrib.astype(object))
    assert [int(x) for x in sums] == [int(x) & M64 for x in ref]
    assert np.array_equal(et, flips * np.uint64(an.A_OID))
    assert eh.size == 0 and eg.size == 0  # exact decoded sums: test_oracle_flip_pins.py


def test_aggregate_sum_grouped_out_of_range_group_is_reported():
    codes = datagen.encode(np.array([1, 2, 3], np.uint64), 65117, 8)
    gid = enc_oids(np.array([0, 5, 1], np.uint64))
    sums, eh, et, eg = O.aggregate_sum_grouped_an(O.void(3), O.Col(codes, an.an_meta(an.RESINT, 65117)),
                                                  O.Col(gid, RESOID), 2)
    assert eg.tolist() == [1 * an.A_OID]
    assert ((int(sums[0]) * an.inverse(an.A_RESULT, 64)) & M64) == 1


def test_mul_sum_matches_numpy():
    n = 30_000
    p = datagen.uniform(n, 15, 1, 0, 10**6)
    q = datagen.uniform(n, 15, 2, 0, 10)
    total, e1, e2 = O.aggregate_mul_sum_an(O.Col(datagen.encode(p, 65117, 8), an.an_meta(an.RESINT, 65117)),
                                           O.Col(datagen.encode(q, 233, 2), an.an_meta(an.RESTINY, 233)))
    dec = (total * an.inverse(an.A_RESULT, 64)) & M64
    assert dec == int((p.astype(object) * q.astype(object)).sum()) & M64
    assert e1.size == 0 and e2.size == 0


@pytest.mark.parametrize("op", [0, 1, 2, 3])
def test_arithmetic_matches_numpy(op):
    n = 10_000
    x = datagen.uniform(n, 16, 1, 0, 10**6)
    y = datagen.uniform(n, 16, 2, 0, 10**5)
    cx = O.Col(datagen.encode(x, 65117, 8), an.an_meta(an.RESINT, 65117))
    cy = O.Col(datagen.encode(y, 63877, 8), an.an_meta(an.RESINT, 63877))
    out, *errs = O.arithmetic_an(O.void(n), cx, O.void(n), cy, op, an.A_RESULT, 8)
    A = an.A_RESULT
    xs, ys = [int(v) * A & M64 for v in x], [int(v) * A & M64 for v in y]
    f = {0: lambda a, b: a + b, 1: lambda a, b: a - b, 2: lambda a, b: a * b, 3: lambda a, b: a // b if b else 0}[op]
    assert out.tolist() == [f(a, b) & M64 for a, b in zip(xs, ys)]
    assert all(e.size == 0 for e in errs)


def test_check_and_decode2_flips_and_counts():
    """checkAndDecodeAN<Head, Tail> (encdecAN_scalar.tcc:129-176): every injected single flip is
    reported on its side (closed form I1), and a pair survives iff neither side was flipped."""
    n = 20_000
    hv = datagen.uniform(n, 61, 1, 0, 1 << 20)
    tv = datagen.uniform(n, 61, 2, 0, 200)
    h = datagen.encode(hv, an.A_OID, 8)
    t = datagen.encode(tv, 29, 2)
    fh = datagen.inject_flips(h, 61, 3, datagen.flip_threshold(2e-3))
    ft = datagen.inject_flips(t, 61, 5, datagen.flip_threshold(2e-3))
    dh, dt, eh, et = O.check_and_decode_an2(O.Col(h, RESOID), O.Col(t, an.an_meta(an.RESTINY, 29)))
    assert np.array_equal(eh, fh.astype(np.uint64) * np.uint64(an.A_OID_ERR))
    assert np.array_equal(et, ft.astype(np.uint64) * np.uint64(an.A_OID_ERR))
    keep = np.ones(n, bool)
    keep[fh.astype(np.int64)] = False
    keep[ft.astype(np.int64)] = False
    assert np.array_equal(dh, hv[keep]) and np.array_equal(dt, tv[keep])


def test_matchjoin_seq_is_the_oid_range_intersection():
    """matchjoinAN <H1, void, void, T2> (matchjoinAN.tcc:146-224): the pairs are the rows whose
    virtual OIDs coincide, i.e. the intersection of [sb1, sb1+n1) and [sb2, sb2+n2); the rows the
    reference visits on each side are those below the end of that intersection."""
    rng = np.random.default_rng(3)
    for _ in range(30):
        n1, n2 = int(rng.integers(0, 400)), int(rng.integers(0, 400))
        sb1, sb2 = int(rng.integers(0, 300)), int(rng.integers(0, 300))
        h1 = datagen.encode(rng.integers(0, 1 << 20, n1).astype(np.uint64), an.A_OID, 8)
        t2 = datagen.encode(rng.integers(0, 1 << 20, n2).astype(np.uint64), an.A_OID, 8)
        f1 = datagen.inject_flips(h1, 7 + _, 1, datagen.flip_threshold(0.05))
        f2 = datagen.inject_flips(t2, 7 + _, 3, datagen.flip_threshold(0.05))
        oh, ot, e1, e4 = O.matchjoin_seq_an(O.Col(h1, RESOID), O.void(n1, sb1), O.void(n2, sb2), O.Col(t2, RESOID))
        lo, hi = max(sb1, sb2), min(sb1 + n1, sb2 + n2)
        m = max(0, hi - lo) if n1 and n2 else 0
        assert np.array_equal(oh, h1[lo - sb1: lo - sb1 + m]) if m else oh.size == 0
        assert np.array_equal(ot, t2[lo - sb2: lo - sb2 + m]) if m else ot.size == 0
        if n1 and n2:
            if m:    # skipped prefix of the lower side, then m pairs
                v1, v2 = lo - sb1 + m, lo - sb2 + m
            else:    # no overlap: only the lower side is walked, up to the end of a list
                v1 = min(n1, sb2 - sb1) if sb1 < sb2 else 0
                v2 = min(n2, sb1 - sb2) if sb2 < sb1 else 0
            assert np.array_equal(e1, f1[f1 < v1].astype(np.uint64) * np.uint64(an.A_OID_ERR))
            assert np.array_equal(e4, f2[f2 < v2].astype(np.uint64) * np.uint64(an.A_OID_ERR))
        else:
            assert e1.size == 0 and e4.size == 0
