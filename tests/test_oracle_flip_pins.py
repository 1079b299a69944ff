"""Pins the oracle's error-position branches under injected single-bit flips.

Every column here uses a code type and A for which closed form I1 holds (2^k / A > maxU): every
single-bit flip of a valid code decodes above maxU.  So the expected error vector of a column is
the sorted injected flip positions (datagen.inject_flips returns them) times A_OID, in the scan
position the reference pushes (groupbyAN.tcc:72-79,131-143; aggregateAN_scalar.tcc:91-99;
arithmeticAN_scalar.tcc:90-106; matchjoinAN.tcc:276-297; hashjoinAN.tcc:134-141,150-156,
177-184,193-199).  None of these expectations is computed by the oracle's own formulas: a position
pushed to the wrong vector, `pos` vs `pos2`, `i` vs the column row, a raw code where a decoded
value belongs or a dropped row that should be kept fails one of them.
"""
import numpy as np
import pytest

import oracle_lib as O
from ahead_amd import an, datagen

M64 = (1 << 64) - 1
ERR = an.A_OID_ERR
GRP = an.A_OID  # groupbyAN / aggregate_sum_groupedAN default A_OID (OperatorsAN.hpp:112,121,133)
TH = datagen.flip_threshold(5e-3)


def i1_holds(meta: an.ColMeta) -> bool:
    return (1 << (8 * meta.type.width)) // meta.a > meta.type.max_u


def scaled(pos, a):
    with np.errstate(over="ignore"):
        return np.asarray(pos, dtype=np.uint64) * np.uint64(a)


def flipped_col(values, meta, seed, stream):
    codes = datagen.encode(values, meta.a, meta.width)
    pos = datagen.inject_flips(codes, seed, stream, TH)
    assert pos.size > 0
    return codes, pos


def test_closed_form_preconditions():
    for meta in (an.an_meta(an.RESTINY, 29), an.an_meta(an.RESTINY, 233), an.an_meta(an.RESSHORT, 233),
                 an.an_meta(an.RESINT, 65117), an.an_meta(an.RESINT, 63877), an.an_meta(an.RESOID, 233)):
        assert i1_holds(meta), meta
    assert not i1_holds(an.an_meta(an.RESOID, an.A_OID))  # why the OID-typed columns below use A = 233


def test_groupby_unary_error_vectors_and_raw_head_groid():
    n = 30_000
    mh, mt = an.an_meta(an.RESSHORT, 233), an.an_meta(an.RESTINY, 29)
    hv = datagen.uniform(n, 71, 1, 0, 60_000)
    tv = datagen.uniform(n, 71, 2, 0, 30)
    h, fh = flipped_col(hv, mh, 71, 3)
    t, ft = flipped_col(tv, mt, 71, 5)
    vgid, groid, eh, et = O.groupby_an(O.Col(h, mh), O.Col(t, mt))
    assert np.array_equal(eh, scaled(fh, GRP))
    assert np.array_equal(et, scaled(ft, GRP))
    # the key is the decoded tail truncated to its 8-bit unencoded type (groupbyAN.tcc:80); corrupt
    # codes take part with that truncated value; VGIDtoROID holds the RAW head code * A_OID
    # (groupbyAN.tcc:87), not the decoded head
    key = ((t.astype(np.uint32) * np.uint32(mt.a_inv)) & np.uint32(0xFFFF)) & np.uint32(0xFF)
    _, first, inv = np.unique(key, return_index=True, return_inverse=True)
    order = np.argsort(first)
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    assert np.array_equal(vgid, scaled(rank[inv], GRP))
    assert np.array_equal(groid, scaled(h[first[order]], GRP))
    assert np.unique(key).size > 31  # some corrupt codes formed groups of their own


def test_groupby_binary_error_vectors_and_decoded_head_groid():
    n = 30_000
    mh, mt, mg = an.an_meta(an.RESSHORT, 233), an.an_meta(an.RESTINY, 233), an.an_meta(an.RESOID, 233)
    G = 11
    hv = datagen.uniform(n, 72, 1, 0, 60_000)
    tv = datagen.uniform(n, 72, 2, 0, 6)
    gv = datagen.uniform(n, 72, 3, 0, G - 1)
    h, fh = flipped_col(hv, mh, 72, 4)
    t, ft = flipped_col(tv, mt, 72, 6)
    g, fg = flipped_col(gv, mg, 72, 8)
    vgid, groid, eh, et, eg = O.groupby2_an(O.Col(h, mh), O.Col(t, mt), O.Col(g, mg), G, 32)
    assert np.array_equal(eh, scaled(fh, GRP))
    assert np.array_equal(et, scaled(ft, GRP))
    assert np.array_equal(eg, scaled(fg, GRP))
    # VGIDtoROID holds the DECODED head * A_OID here (groupbyAN.tcc:152), unlike the unary form;
    # on rows without any flip the groups are the first occurrences of (tail, group id) pairs
    clean = np.ones(n, bool)
    for f in (fh, ft, fg):
        clean[f.astype(np.int64)] = False
    first_pair = {}
    for i in np.nonzero(clean)[0][:2000]:
        first_pair.setdefault((int(tv[i]), int(gv[i])), int(i))
    dec_h = (h.astype(np.uint64) * np.uint64(mh.a_inv)) & np.uint64(0xFFFFFFFF)
    got_first = {int(x) for x in groid}
    for (tk, gk), i in first_pair.items():
        earlier = np.nonzero((tv[:i] == tk) & (gv[:i] == gk))[0]
        if earlier.size == 0:  # i is the pair's first row overall
            assert (int(dec_h[i]) * GRP) & M64 in got_first


def test_mul_sum_error_vectors_over_the_shorter_input():
    n1, n2 = 20_000, 19_900
    m1, m2 = an.an_meta(an.RESINT, 65117), an.an_meta(an.RESTINY, 233)
    p = datagen.uniform(n1, 73, 1, 0, 10**6)
    q = datagen.uniform(n2, 73, 2, 0, 10)
    c1, f1 = flipped_col(p, m1, 73, 3)
    c2, f2 = flipped_col(q, m2, 73, 5)
    total, e1, e2 = O.aggregate_mul_sum_an(O.Col(c1, m1), O.Col(c2, m2))
    m = min(n1, n2)
    assert np.array_equal(e1, scaled(f1[f1 < m], ERR))   # aggregateAN_scalar.tcc:91-99, i over the
    assert np.array_equal(e2, scaled(f2[f2 < m], ERR))   # rows both inputs have
    # corrupt codes are still multiplied in (nothing is dropped): the flipped total differs from the
    # clean one by F * sum over flipped rows of (c1' c2' - c1 c2), F = A_R * inv(A1) * inv(A2)
    clean1, clean2 = datagen.encode(p, 65117, 8), datagen.encode(q, 233, 2)
    t_clean, *_ = O.aggregate_mul_sum_an(O.Col(clean1, m1), O.Col(clean2, m2))
    F = (an.A_RESULT * an.inverse(65117, 64) * an.inverse(233, 64)) & M64
    rows = sorted(set(int(x) for x in f1[f1 < m]) | set(int(x) for x in f2[f2 < m]))
    delta = sum(int(c1[i]) * int(c2[i]) - int(clean1[i]) * int(clean2[i]) for i in rows)
    assert (int(total) - int(t_clean)) % (1 << 64) == (F * delta) % (1 << 64)


def test_arithmetic_four_error_vectors_and_no_drop():
    n = 20_000
    mh, mt1, mt2 = an.an_meta(an.RESSHORT, 233), an.an_meta(an.RESINT, 65117), an.an_meta(an.RESINT, 63877)
    h1, fh1 = flipped_col(datagen.uniform(n, 74, 1, 0, 60_000), mh, 74, 2)
    t1, ft1 = flipped_col(datagen.uniform(n, 74, 3, 0, 10**6), mt1, 74, 4)
    h2, fh2 = flipped_col(datagen.uniform(n, 74, 5, 0, 60_000), mh, 74, 6)
    t2, ft2 = flipped_col(datagen.uniform(n, 74, 7, 0, 10**5), mt2, 74, 8)
    out, eh1, et1, eh2, et2 = O.arithmetic_an(O.Col(h1, mh), O.Col(t1, mt1), O.Col(h2, mh), O.Col(t2, mt2), 1,
                                              an.A_RESULT, 8)
    for got, f in ((eh1, fh1), (et1, ft1), (eh2, fh2), (et2, ft2)):
        assert np.array_equal(got, scaled(f, ERR))
    assert out.size == n  # arithmeticAN_scalar.tcc:107: every row is emitted, corrupt or not


def test_matchjoin_positional_err2_and_err4_positions():
    n_col, n = 4000, 15_000
    mt = an.an_meta(an.RESTINY, 29)
    col, fcol = flipped_col(datagen.uniform(n_col, 75, 1, 0, 255), mt, 75, 2)
    pos = datagen.uniform(n, 75, 3, 0, n_col + 40)
    pos[:5] = n_col  # exactly |col|: the reference's `>` test reads past the end here (D6)
    heads = datagen.encode(datagen.uniform(n, 75, 4, 0, 60_000), 233, 4)
    mh = an.an_meta(an.RESSHORT, 233)
    resoid = an.an_meta(an.RESOID, an.A_OID)
    oh, ot, e1, e2, e4 = O.matchjoin_pos_an(O.Col(heads, mh), O.Col(scaled(pos, an.A_OID), resoid), O.Col(col, mt))
    bad = np.zeros(n_col, bool)
    bad[fcol.astype(np.int64)] = True
    inr = pos < n_col
    p = np.where(inr, pos, 0).astype(np.int64)
    corrupt = inr & bad[p]
    keep = inr & ~bad[p]
    assert e1.size == 0
    assert np.array_equal(e2, scaled(np.nonzero(~inr)[0], ERR))  # pos1: the row of list 1
    assert np.array_equal(e4, scaled(pos[corrupt], ERR))          # pos2: the row of the column
    assert np.array_equal(oh, heads[keep].astype(np.uint64))
    assert np.array_equal(ot, col[p[keep]].astype(np.uint64))


@pytest.mark.parametrize("build_right", [1, 0])
def test_hashjoin_four_error_vectors(build_right):
    n1, n2 = 3000, 2000
    mh, mk = an.an_meta(an.RESSHORT, 233), an.an_meta(an.RESINT, 63877)
    h1, fh1 = flipped_col(datagen.uniform(n1, 76, 1, 0, 60_000), mh, 76, 2)
    t1, ft1 = flipped_col(datagen.uniform(n1, 76, 3, 0, 1500), mk, 76, 4)
    h2, fh2 = flipped_col(datagen.uniform(n2, 76, 5, 0, 1500), mk, 76, 6)
    t2, ft2 = flipped_col(datagen.uniform(n2, 76, 7, 0, 60_000), mh, 76, 8)
    oh, ot, e1, e2, e3, e4 = O.hashjoin_an(O.Col(h1, mh), O.Col(t1, mk), O.Col(h2, mk), O.Col(t2, mh), build_right)
    # every row of both sides is checked once, at its own scan position, whichever side is built
    for got, f in ((e1, fh1), (e2, ft1), (e3, fh2), (e4, ft2)):
        assert np.array_equal(got, scaled(f, ERR))
    # corrupt keys still join with their truncated decoded value (D8): a nested loop on those values
    k1 = ((t1 * np.uint64(mk.a_inv)) & np.uint64(0xFFFFFFFF)).tolist()
    k2 = ((h2 * np.uint64(mk.a_inv)) & np.uint64(0xFFFFFFFF)).tolist()
    idx2 = {}
    for j, k in enumerate(k2):
        idx2.setdefault(k, []).append(j)
    if build_right:
        ref = [(int(h1[i]), int(t2[j])) for i, k in enumerate(k1) for j in idx2.get(k, [])]
    else:
        idx1 = {}
        for i, k in enumerate(k1):
            idx1.setdefault(k, []).append(i)
        ref = [(int(h1[i]), int(t2[j])) for j, k in enumerate(k2) for i in idx1.get(k, [])]
    assert list(zip(oh.tolist(), ot.tolist())) == ref


def test_aggregate_sum_grouped_exact_decoded_sums_and_error_vectors():
    n, G = 40_000, 29
    mt, mg = an.an_meta(an.RESINT, 65117), an.an_meta(an.RESOID, 233)
    m = datagen.uniform(n, 77, 1, 90_000, 10_500_000)
    gid = datagen.uniform(n, 77, 2, 0, G - 1)
    clean_codes = datagen.encode(m, 65117, 8)
    gcodes = datagen.encode(gid, 233, 8)
    # fault-free: every decoded group sum is the textbook GROUP BY sum
    sums, eh, et, eg = O.aggregate_sum_grouped_an(O.void(n), O.Col(clean_codes, mt), O.Col(gcodes, mg), G)
    ref = [0] * G
    for v, g in zip(m.tolist(), gid.tolist()):
        ref[g] += v
    assert [(int(s) * an.inverse(an.A_RESULT, 64)) & M64 for s in sums] == ref
    assert eh.size == et.size == eg.size == 0
    # flips in the measure (added anyway, aggregateAN.tcc:91-97) and in the group ids (skipped)
    codes, fm = flipped_col(m, mt, 77, 3)
    g2, fg = flipped_col(gid, mg, 77, 5)
    sums2, eh, et, eg = O.aggregate_sum_grouped_an(O.void(n), O.Col(codes, mt), O.Col(g2, mg), G)
    assert np.array_equal(et, scaled(fm, GRP)) and np.array_equal(eg, scaled(fg, GRP)) and eh.size == 0
    F = (an.A_RESULT * an.inverse(65117, 64)) & M64
    skip = set(int(x) for x in fg)
    ref2 = [0] * G
    for i in range(n):
        if i not in skip:
            ref2[int(gid[i])] = (ref2[int(gid[i])] + int(codes[i]) * F) & M64
    assert [int(s) for s in sums2] == ref2
