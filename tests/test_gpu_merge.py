"""GPU parity of the merge form of matchjoinAN (matchjoinAN.tcc:46-143) against the oracle,
bit-exact, error vectors included: sorted inputs (the merge-path tiles), sorted inputs with long
runs of equal keys, and unsorted or corrupted inputs (the single-thread path).  A list whose head
and tail are one device column (a mirrored selection, as the star-schema queries pass them) is
staged once; the tests run both that and separate columns."""
import numpy as np
import pytest

import oracle_lib as O
from ahead_amd import an, capi, datagen

pytestmark = pytest.mark.gpu

RESOID = an.an_meta(an.RESOID, an.A_OID)


def same(got, ref):
    assert len(got) == len(ref)
    for k, (g, r) in enumerate(zip(got, ref)):
        assert np.array_equal(np.asarray(g).astype(np.uint64), np.asarray(r).astype(np.uint64)), k


def oid_list(keys, seed, stream, flip_rate=0.0):
    codes = datagen.encode(keys, an.A_OID, 8)
    if flip_rate:
        datagen.inject_flips(codes, seed, stream, datagen.flip_threshold(flip_rate))
    return codes


def run_both(h1, t1, h2, t2, m_h1, m_t1, m_h2, m_t2, *args, alias=True):
    """alias: the same array passed for both sides of a list becomes one device column."""
    ref = O.matchjoin_merge_an(O.Col(h1, m_h1), O.Col(t1, m_t1), O.Col(h2, m_h2), O.Col(t2, m_t2), *args)
    c_t1, c_h2 = capi.Column.from_numpy(t1, m_t1), capi.Column.from_numpy(h2, m_h2)
    c_h1 = c_t1 if alias and h1 is t1 and m_h1 == m_t1 else capi.Column.from_numpy(h1, m_h1)
    c_t2 = c_h2 if alias and t2 is h2 and m_t2 == m_h2 else capi.Column.from_numpy(t2, m_t2)
    got = capi.matchjoin_merge_an(c_h1, c_t1, c_h2, c_t2, *args, err_cap=2)
    same(got, ref)
    return ref


@pytest.mark.parametrize("n1,n2,hi", [(0, 10, 50), (10, 0, 50), (1, 1, 0), (15, 40, 60), (16, 16, 20), (17, 100, 300),
                                      (4095, 5000, 20_000), (4097, 300, 20_000), (100_003, 80_001, 400_000),
                                      (200_000, 1000, 10_000_000), (1000, 200_000, 300_000),
                                      (20_000, 140_000, 200_000), (1_000_000, 600_000, 4_000_000)])
@pytest.mark.parametrize("alias", [True, False])
def test_merge_sorted_oid_lists(n1, n2, hi, alias):
    # two selections' OID lists (strictly ascending), the Q1.x shape: BAT<resoid, resoid> both
    seed = n1 + 3 * n2
    k1 = np.unique(datagen.uniform(n1, seed, 1, 0, hi))
    k2 = np.unique(datagen.uniform(n2, seed, 2, 0, hi))
    t1, t2 = oid_list(k1, seed, 3), oid_list(k2, seed, 4)
    ref = run_both(t1, t1, t2, t2, RESOID, RESOID, RESOID, RESOID, alias=alias)
    assert ref[0].size == np.intersect1d(k1, k2).size


@pytest.mark.parametrize("n1,n2,hi", [(5000, 5000, 30), (20_000, 3000, 7), (3000, 20_000, 700), (65_536, 65_536, 1)])
def test_merge_sorted_with_runs(n1, n2, hi):
    seed = n1 + n2 + hi
    meta = an.an_meta(an.RESSHORT, 233)
    k1 = np.sort(datagen.uniform(n1, seed, 1, 0, hi))
    k2 = np.sort(datagen.uniform(n2, seed, 2, 0, hi))
    h1 = oid_list(np.arange(n1, dtype=np.uint64), seed, 3)
    t2 = datagen.encode(datagen.uniform(n2, seed, 4, 0, 200), 29, 2)
    run_both(h1, datagen.encode(k1, 233, 4), datagen.encode(k2, 233, 4), t2, RESOID, meta, meta,
             an.an_meta(an.RESTINY, 29))


@pytest.mark.parametrize("n1,n2", [(300, 500), (5000, 4000), (20_000, 20_000)])
def test_merge_unsorted_and_corrupted(n1, n2):
    seed = n1 * 7 + n2
    meta = an.an_meta(an.RESSHORT, 233)
    # unsorted keys: the walk is the reference's two-cursor walk all the same
    k1 = datagen.uniform(n1, seed, 1, 0, 50)
    k2 = datagen.uniform(n2, seed, 2, 0, 50)
    h1 = oid_list(np.arange(n1, dtype=np.uint64), seed, 3, 1e-3)
    t2 = oid_list(np.arange(n2, dtype=np.uint64), seed, 4, 1e-3)
    run_both(h1, datagen.encode(k1, 233, 4), datagen.encode(k2, 233, 4), t2, RESOID, meta, meta, RESOID)
    # sorted keys with flipped bits in both key lists
    c1 = datagen.encode(np.sort(datagen.uniform(n1, seed, 5, 0, 3 * n1)), 233, 4)
    c2 = datagen.encode(np.sort(datagen.uniform(n2, seed, 6, 0, 3 * n1)), 233, 4)
    datagen.inject_flips(c1, seed, 7, datagen.flip_threshold(2e-3))
    datagen.inject_flips(c2, seed, 8, datagen.flip_threshold(2e-3))
    run_both(h1, c1, c2, t2, RESOID, meta, meta, RESOID)
    # a mirrored list with flipped bits: both sides read from the one staged column
    run_both(c1, c1, c2, c2, meta, meta, meta, meta)


def test_merge_fetched_tail_with_errors():
    # the shape of a fetched attribute matched against a later selection: list 2 (OIDs with an
    # attribute) several times longer than list 1 (mirrored OIDs).  The keys stay sorted (clean),
    # so the tiles run; the errors are in the columns that are not keys.
    n2 = 400_000
    k2 = np.unique(datagen.uniform(n2, 17, 1, 0, 3 * n2))
    k1 = np.unique(k2[datagen.uniform(k2.size // 7, 17, 2, 0, k2.size - 1).astype(np.int64)])
    k1 = np.union1d(k1, np.unique(datagen.uniform(k1.size // 3, 17, 7, 0, 3 * n2)))  # some unmatched
    t1, h2 = oid_list(k1, 17, 3), oid_list(k2, 17, 4)
    meta = an.an_meta(an.RESSHORT, 233)
    t2 = datagen.encode(datagen.uniform(k2.size, 17, 5, 0, 0xFFFF), 233, 4)
    datagen.inject_flips(t2, 17, 6, datagen.flip_threshold(1e-3))
    ref = run_both(t1, t1, h2, t2, RESOID, RESOID, RESOID, meta)
    assert ref[5].size > 0
    h1 = oid_list(np.arange(k1.size, dtype=np.uint64), 17, 8, 1e-3)
    ref = run_both(h1, t1, h2, t2, RESOID, RESOID, RESOID, meta)
    assert ref[2].size > 0 and ref[5].size > 0


def test_merge_reencode_and_widths():
    n1, n2 = 30_000, 25_000
    k1 = np.unique(datagen.uniform(n1, 91, 1, 0, 60_000))
    k2 = np.unique(datagen.uniform(n2, 91, 2, 0, 60_000))
    t1, h2 = oid_list(k1, 91, 3, 1e-4), oid_list(k2, 91, 4, 1e-4)
    h1 = oid_list(np.arange(k1.size, dtype=np.uint64) + 5, 91, 5, 1e-4)
    t2 = datagen.encode(datagen.uniform(k2.size, 91, 6, 0, 0xFFFF), 63877, 4)
    m_t2 = an.an_meta(an.RESSHORT, 63877)
    f_h1 = (16089 * an.A_OID_INV) & an.MASK64
    f_t2 = (233 * an.inverse(63877, 32)) & 0xFFFFFFFF
    run_both(h1, t1, h2, t2, RESOID, RESOID, RESOID, m_t2, 1, f_h1, 8, 1, f_t2, 4)
    # head-only re-encoding zeroes the tail factor, as the reference does
    run_both(h1, t1, h2, t2, RESOID, RESOID, RESOID, m_t2, 1, f_h1, 8, 1, 0, 4)


def _span_list(n, span, seed, stream, base=1000):
    """n strictly ascending keys in [base, base + span), both ends included."""
    if n <= 2 or span <= 2:
        return np.unique(np.array([base, base + span - 1], dtype=np.uint64))[:n]
    inner = np.unique(datagen.uniform(3 * n, seed, stream, base + 1, base + span - 2))[: max(n - 2, 0)]
    return np.union1d(np.array([base, base + span - 1], dtype=np.uint64), inner).astype(np.uint64)


@pytest.mark.parametrize("n1,n2,bits_per_row", [(300_000, 200_000, 15.9), (300_000, 200_000, 16.2),
                                                (70_000, 1_000_000, 1.5), (1_000_000, 3_000, 15.0),
                                                (5, 64, 1.0), (64, 1, 1.0)])
def test_merge_strictly_ascending_clean_lists(n1, n2, bits_per_row):
    """Strictly ascending key lists with valid codes everywhere, whose walk is the intersection:
    list 2's key span just below and above 16 bits per row, lists of one row and of one bitmap
    word, mirrored lists, separate gathered tails and re-encoded heads."""
    seed = n1 + n2
    span = max(int(bits_per_row * n2), n2)
    k2 = _span_list(n2, span, seed, 1)
    n2 = k2.size
    pick = datagen.uniform(n1 // 2 + 1, seed, 2, 0, n2 - 1).astype(np.int64)
    k1 = np.union1d(k2[pick], datagen.uniform(n1 - n1 // 2, seed, 3, 0, 1000 + span + 500)).astype(np.uint64)
    t1, h2 = oid_list(k1, seed, 4), oid_list(k2, seed, 5)
    ref = run_both(t1, t1, h2, h2, RESOID, RESOID, RESOID, RESOID)
    assert ref[0].size == np.intersect1d(k1, k2).size
    h1 = oid_list(np.arange(k1.size, dtype=np.uint64) + 7, seed, 6)
    t2 = datagen.encode(datagen.uniform(n2, seed, 7, 0, 0xFFFF), 63877, 4)
    m_t2 = an.an_meta(an.RESSHORT, 63877)
    f_h1 = (16089 * an.A_OID_INV) & an.MASK64
    f_t2 = (233 * an.inverse(63877, 32)) & 0xFFFFFFFF
    run_both(h1, t1, h2, t2, RESOID, RESOID, RESOID, m_t2, 1, f_h1, 8, 1, f_t2, 4)
    run_both(h1, t1, h2, t2, RESOID, RESOID, RESOID, m_t2, 0, 0, 8, 0, 0, 4)


@pytest.mark.parametrize("where", ["swap2", "dup2", "corrupt_t2", "swap1", "dup1", "corrupt_h1", "corrupt_t1_matched"])
def test_merge_one_violation_in_ascending_lists(where):
    """One row that breaks the intersection path's conditions (order, uniqueness or a valid code,
    in either list; a corrupt head of a matched list-1 row, which the walk never checks): the
    result is the reference walk's all the same."""
    n1, n2 = 400_000, 300_000
    k2 = _span_list(n2, 4 * n2, 5, 1)
    k1 = np.union1d(k2[::3], datagen.uniform(n1 // 2, 5, 2, 0, 4 * n2 + 2000)).astype(np.uint64)
    t1, h2 = oid_list(k1, 5, 3), oid_list(k2, 5, 4)
    h1 = oid_list(np.arange(k1.size, dtype=np.uint64), 5, 5)
    t2 = oid_list(np.arange(k2.size, dtype=np.uint64), 5, 6)
    r2, r1 = k2.size * 2 // 3 + 1, k1.size // 2 + 1  # rows deep inside a list, off any word boundary
    if where == "swap2":
        h2[[r2, r2 + 1]] = h2[[r2 + 1, r2]]
    elif where == "dup2":
        h2[r2 + 1] = h2[r2]
    elif where == "corrupt_t2":
        t2[r2] ^= np.uint64(1 << 17)
    elif where == "swap1":
        t1[[r1, r1 + 1]] = t1[[r1 + 1, r1]]
    elif where == "dup1":
        t1[r1 + 1] = t1[r1]
    elif where == "corrupt_h1":
        h1[r1] ^= np.uint64(1 << 9)
    else:
        r = int(np.searchsorted(k1, k2[r2]))
        assert k1[r] == k2[r2]
        h1[r] ^= np.uint64(1 << 3)
    run_both(h1, t1, h2, t2, RESOID, RESOID, RESOID, RESOID)
    run_both(t1, t1, h2, h2, RESOID, RESOID, RESOID, RESOID)


def test_merge_list2_dense_then_sparse():
    """A list 2 whose density changes: 150,000 consecutive keys, then 2,000 keys 600 apart (a
    stretch of it spans more bitmap words than one step stages), within 16 key values per row in
    all; list 1 meets both stretches."""
    dense = np.arange(10_000, 160_000, dtype=np.uint64)
    sparse = np.uint64(160_000) + np.arange(1, 2_001, dtype=np.uint64) * np.uint64(600)
    k2 = np.concatenate([dense, sparse])
    assert (int(k2[-1]) - int(k2[0])) // 64 + 1 <= 16 * k2.size // 64 + 2
    k1 = np.union1d(k2[::5], datagen.uniform(100_000, 9, 1, 0, int(k2[-1]) + 1000)).astype(np.uint64)
    t1, h2 = oid_list(k1, 9, 2), oid_list(k2, 9, 3)
    ref = run_both(t1, t1, h2, h2, RESOID, RESOID, RESOID, RESOID)
    assert ref[0].size == np.intersect1d(k1, k2).size
    t2 = oid_list(np.arange(k2.size, dtype=np.uint64) * np.uint64(3), 9, 4)
    run_both(t1, t1, h2, t2, RESOID, RESOID, RESOID, RESOID)


@pytest.mark.parametrize("repeat", [False, True])
def test_merge_output_capacity(repeat):
    """Outputs smaller than the pairs: the first `cap` pairs are written, nothing past them, and the
    count is every pair (the C ABI's contract), on the intersection path (strictly ascending keys,
    a void head) and, with one key repeated in list 2, on the sub-tile path."""
    import ctypes

    k2 = _span_list(200_000, 600_000, 21, 1)
    k1 = np.union1d(k2[::2], datagen.uniform(50_000, 21, 2, 0, 601_000)).astype(np.uint64)
    t1, h2 = oid_list(k1, 21, 3), oid_list(k2, 21, 4)
    if repeat:
        h2[1001] = h2[1000]
    t2 = oid_list(np.arange(k2.size, dtype=np.uint64) + 3, 21, 5)
    ref = O.matchjoin_merge_an(O.void(k1.size, 7), O.Col(t1, RESOID), O.Col(h2, RESOID), O.Col(t2, RESOID))
    total = ref[0].size
    cap = total // 2
    lib, ws = capi.load(), capi.workspace()
    cols = [capi.Column.void(k1.size, 7), capi.Column.from_numpy(t1, RESOID), capi.Column.from_numpy(h2, RESOID),
            capi.Column.from_numpy(t2, RESOID)]
    d = [c.desc() for c in cols]
    sentinel = np.full(cap + 64, 0xA5A5A5A5A5A5A5A5, dtype=np.uint64)
    oh, ot = capi.DeviceBuffer.from_numpy(sentinel), capi.DeviceBuffer.from_numpy(sentinel)
    errs = [capi.DeviceBuffer(8 * 16) for _ in range(4)]
    cnt = capi.DeviceBuffer(8 * 8, zero=True)
    capi.check(lib.ahead_hip_matchjoin_merge_an(*[ctypes.byref(x) for x in d], 0, 0, 8, 0, 0, 8, an.A_OID_ERR,
                                                capi.Buf(oh.ptr, cap), capi.Buf(ot.ptr, cap),
                                                *[capi.Buf(e.ptr, 16) for e in errs], cnt.ptr, ws.h,
                                                capi.current_stream()))
    capi.synchronize()
    c = cnt.to_numpy(np.uint64, 5)
    assert int(c[0]) == total and all(int(x) == r.size for x, r in zip(c[1:], ref[2:]))
    got_h, got_t = oh.to_numpy(np.uint64, cap + 64), ot.to_numpy(np.uint64, cap + 64)
    assert np.array_equal(got_h[:cap], ref[0][:cap].astype(np.uint64))
    assert np.array_equal(got_t[:cap], ref[1][:cap].astype(np.uint64))
    assert np.all(got_h[cap:] == sentinel[cap:]) and np.all(got_t[cap:] == sentinel[cap:])


@pytest.mark.parametrize("kind", ["resshort", "restiny", "int", "shortint", "tinyint"])
def test_merge_ascending_lists_of_every_key_width(kind):
    """Strictly ascending key lists with valid codes in 4-, 2- and 1-byte key columns, encoded and
    plain (the intersection path loads keys at the column's width); OID heads and tails."""
    t, a, hi = {"resshort": (an.RESSHORT, 233, 65_535), "restiny": (an.RESTINY, 29, 255),
                "int": (an.INT, 1, 3_000_000), "shortint": (an.SHORTINT, 1, 65_535),
                "tinyint": (an.TINYINT, 1, 255)}[kind]
    meta = an.an_meta(t, a) if t.encoded else an.plain_meta(t)
    n = min(hi // 2, 400_000)
    k2 = np.unique(datagen.uniform(n, 31, 1, 0, hi))
    k1 = np.unique(datagen.uniform(n, 31, 2, 0, hi))
    c1, c2 = datagen.encode(k1, a, t.width), datagen.encode(k2, a, t.width)
    h1 = oid_list(np.arange(k1.size, dtype=np.uint64), 31, 3)
    t2 = oid_list(np.arange(k2.size, dtype=np.uint64) + 11, 31, 4)
    ref = run_both(h1, c1, c2, t2, RESOID, meta, meta, RESOID)
    assert ref[0].size == np.intersect1d(k1, k2).size and ref[0].size > 0
    run_both(c1, c1, c2, c2, meta, meta, meta, meta)
