"""GPU parity of the per-operator kernels (fetch, positional match, hash join, group-by, grouped
sum, Σ-product, arithmetic) against the CPU oracle, bit-exact, including error positions."""
import numpy as np
import pytest

import oracle_lib as O
from ahead_amd import an, capi, datagen

pytestmark = pytest.mark.gpu

RESOID = an.an_meta(an.RESOID, an.A_OID)
SIZES = [0, 1, 255, 2047, 2048, 2049, 100_003]


def enc_oids(pos):
    with np.errstate(over="ignore"):
        return pos.astype(np.uint64) * np.uint64(an.A_OID)


def gcol(arr, meta):
    return capi.Column.from_numpy(arr, meta)


def same(got, ref):
    assert len(got) == len(ref)
    for i, (g, r) in enumerate(zip(got, ref)):
        if isinstance(r, np.ndarray):
            assert np.array_equal(np.asarray(g).astype(np.uint64), r.astype(np.uint64)), i
        else:
            assert g == r, i


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("t,a", [(an.RESTINY, 29), (an.RESSHORT, 233), (an.RESINT, 65117)])
def test_fetchjoin_parity(n, t, a):
    n_col = 7000
    col = datagen.encode(datagen.uniform(n_col, n + 1, 1, 0, t.max_u), a, t.width)
    datagen.inject_flips(col, n + 1, 2, datagen.flip_threshold(5e-3))
    pos = datagen.uniform(n, n + 1, 3, 0, n_col + 20)
    meta = an.an_meta(t, a)
    for reenc, ar in [(0, 0), (1, 233)]:
        ref = O.fetchjoin_an(O.Col(enc_oids(pos), RESOID), O.Col(col, meta), reenc, ar)
        got = capi.fetchjoin_an(gcol(enc_oids(pos), RESOID), gcol(col, meta), reenc, ar, err_cap=4)
        same(got, ref)


@pytest.mark.parametrize("n", SIZES)
def test_matchjoin_positional_parity(n):
    n_col = 5000
    vals = datagen.encode(datagen.uniform(n_col, 21, 1, 0, 0xFFFF), 233, 4)
    datagen.inject_flips(vals, 21, 2, datagen.flip_threshold(3e-3))
    pos = datagen.uniform(n, 21, 3, 0, n_col + 5)
    heads = datagen.encode(datagen.uniform(n, 21, 4, 0, 1 << 20), an.A_OID, 8)
    datagen.inject_flips(heads, 21, 5, datagen.flip_threshold(3e-3))
    meta = an.an_meta(an.RESSHORT, 233)
    f_h = 65117 * an.inverse(an.A_OID, 64) & an.MASK64
    for rh, rt in [(0, 0), (1, 1)]:
        args = (rh, f_h, 8, rt, 233 * an.inverse(233, 32) & 0xFFFFFFFF, 4)
        ref = O.matchjoin_pos_an(O.Col(heads, RESOID), O.Col(enc_oids(pos), RESOID), O.Col(vals, meta), *args)
        got = capi.matchjoin_pos_an(gcol(heads, RESOID), gcol(enc_oids(pos), RESOID), gcol(vals, meta), *args)
        same(got, ref)


# build_flips: with flipped build keys the key range spans 2^32 and the probe uses the hash
# table; without, the keys are dense and the probe uses the direct map (one row or several)
@pytest.mark.parametrize("build_flips", [True, False])
@pytest.mark.parametrize("build_right", [1, 0])
@pytest.mark.parametrize("n1,n2,kmax", [(0, 10, 5), (10, 0, 5), (1, 1, 0), (5000, 300, 120), (100_003, 20_000, 25_000),
                                        (30_000, 3_000, 100), (50_000, 40_000, 1_000_000)])
def test_hashjoin_parity(build_flips, build_right, n1, n2, kmax):
    seed = n1 + 7 * n2 + build_right
    k1 = datagen.uniform(n1, seed, 1, 0, kmax)
    k2 = datagen.uniform(n2, seed, 2, 0, kmax)
    meta_k = an.an_meta(an.RESINT, 63877)
    c1 = datagen.encode(k1, 63877, 8)
    c2 = datagen.encode(k2, 63877, 8)
    datagen.inject_flips(c1, seed, 3, datagen.flip_threshold(2e-3))
    if build_flips:
        datagen.inject_flips(c2, seed, 5, datagen.flip_threshold(2e-3))
    h1 = datagen.encode(datagen.uniform(n1, seed, 7, 0, 10**6), an.A_OID, 8)
    t2 = datagen.encode(datagen.uniform(n2, seed, 8, 0, 10**6), an.A_OID, 8)
    f = (an.A_OID * an.inverse(an.A_OID, 64)) & an.MASK64
    for reenc in (0, 1):
        args = (build_right, reenc, f, 8, reenc, f, 8)
        ref = O.hashjoin_an(O.Col(h1, RESOID), O.Col(c1, meta_k), O.Col(c2, meta_k), O.Col(t2, RESOID), *args)
        got = capi.hashjoin_an(gcol(h1, RESOID), gcol(c1, meta_k), gcol(c2, meta_k), gcol(t2, RESOID), *args,
                               out_cap=16, err_cap=2)
        same(got, ref)


def test_hashjoin_void_sides_and_dense_keys():
    # the star-join shape: hashjoinAN(BAT<void,resint> fact FK, reverse(BAT<void,resint> dim key))
    nf, nd = 200_000, 4_000
    fk = datagen.encode(datagen.uniform(nf, 31, 1, 0, nd - 1), 63877, 8)
    datagen.inject_flips(fk, 31, 2, datagen.flip_threshold(1e-4))
    key = datagen.encode(np.arange(nd, dtype=np.uint64), 63877, 8)
    meta_k = an.an_meta(an.RESINT, 63877)
    args = (1, 1, an.A_OID, 8, 1, an.A_OID, 8)
    ref = O.hashjoin_an(O.void(nf), O.Col(fk, meta_k), O.Col(key, meta_k), O.void(nd), *args)
    got = capi.hashjoin_an(capi.Column.void(nf), gcol(fk, meta_k), gcol(key, meta_k), capi.Column.void(nd), *args)
    same(got, ref)


@pytest.mark.parametrize("nf,nd", [(120_000, 5), (300_000, 13), (40_000, 2), (470_000, 2)])
def test_hashjoin_long_row_lists(nf, nd):
    """A left build over the fact FK against a few dimension keys: every key's row list is far
    longer than the 8192 rows ordered in LDS alone (about 235 000 with 470 000 rows over 2 keys,
    the rows per date key of a 600 M-row fact), so the lists are merged in HBM; pairs in probe order, build rows in insertion order (hashjoinAN.tcc:130-172), as the
    oracle gives them.  Flips in the build keys drop rows from the lists and report them."""
    fk = datagen.encode(datagen.uniform(nf, nd, 1, 0, nd - 1), 63877, 8)
    datagen.inject_flips(fk, nd, 2, datagen.flip_threshold(1e-4))
    key = datagen.encode(np.arange(nd, dtype=np.uint64), 63877, 8)
    meta_k = an.an_meta(an.RESINT, 63877)
    for reenc in (0, 1):
        args = (0, reenc, an.A_OID, 8, reenc, an.A_OID, 8)  # build_right = 0: the fact side is built
        ref = O.hashjoin_an(O.void(nf), O.Col(fk, meta_k), O.Col(key, meta_k), O.void(nd), *args)
        got = capi.hashjoin_an(capi.Column.void(nf), gcol(fk, meta_k), gcol(key, meta_k), capi.Column.void(nd), *args)
        same(got, ref)
        assert got[0].size >= nf - 50


# The probe for the common shapes (8- or 4-byte key codes, void or 8-byte payloads, 4- or 8-byte
# outputs, room for a pair per probe row): all rows joining (one pass), some missing (the variant
# with the key filter in LDS and the second, writing pass), both build sides, re-encoding.
@pytest.mark.parametrize("key_w", [8, 4])
@pytest.mark.parametrize("void_pay", [True, False])
@pytest.mark.parametrize("build_right", [1, 0])
@pytest.mark.parametrize("wout", [8, 4])
@pytest.mark.parametrize("hit", ["all", "some"])
def test_hashjoin_lean_shapes(key_w, void_pay, build_right, wout, hit):
    seed = key_w + 3 * void_pay + 7 * build_right + wout + (11 if hit == "all" else 0)
    np_, nb = 150_000, 20_000
    a_k = 63877 if key_w == 8 else 233
    meta_k = an.an_meta(an.RESINT if key_w == 8 else an.RESSHORT, a_k)
    bkeys = np.arange(nb, dtype=np.uint64) * (1 if hit == "all" else 3)
    pk = datagen.uniform(np_, seed, 1, 0, nb - 1) if hit == "all" else datagen.uniform(np_, seed, 1, 0, 3 * nb)
    kc, bc = datagen.encode(pk, a_k, key_w), datagen.encode(bkeys, a_k, key_w)
    datagen.inject_flips(kc, seed, 2, datagen.flip_threshold(1e-3))
    pay_p = None if void_pay else datagen.encode(datagen.uniform(np_, seed, 3, 0, 10**6), an.A_OID, 8)
    pay_b = None if void_pay else datagen.encode(datagen.uniform(nb, seed, 4, 0, 10**6), an.A_OID, 8)
    if pay_p is not None:
        datagen.inject_flips(pay_p, seed, 5, datagen.flip_threshold(1e-3))
    f = (an.A_OID * an.inverse(an.A_OID, 64)) & an.MASK64
    # the output width applies to re-encoded values (hashjoinAN.tcc): 8-byte payload codes go to
    # 4-byte outputs only re-encoded
    for reenc in ((0, 1) if wout == 8 or void_pay else (1,)):
        args = (build_right, reenc, f, wout, reenc, f, wout)
        if build_right:  # probe <pay_p, key>, build <key, pay_b>
            o = (O.void(np_) if void_pay else O.Col(pay_p, RESOID), O.Col(kc, meta_k), O.Col(bc, meta_k),
                 O.void(nb) if void_pay else O.Col(pay_b, RESOID))
            g = (capi.Column.void(np_) if void_pay else gcol(pay_p, RESOID), gcol(kc, meta_k), gcol(bc, meta_k),
                 capi.Column.void(nb) if void_pay else gcol(pay_b, RESOID))
        else:  # build <pay_b, key>, probe <key, pay_p>
            o = (O.void(nb) if void_pay else O.Col(pay_b, RESOID), O.Col(bc, meta_k), O.Col(kc, meta_k),
                 O.void(np_) if void_pay else O.Col(pay_p, RESOID))
            g = (capi.Column.void(nb) if void_pay else gcol(pay_b, RESOID), gcol(bc, meta_k), gcol(kc, meta_k),
                 capi.Column.void(np_) if void_pay else gcol(pay_p, RESOID))
        ref = O.hashjoin_an(*o, *args)
        got = capi.hashjoin_an(*g, *args, err_cap=2)
        same(got, ref)
        assert ref[0].size > 0 and (hit == "all") == (ref[0].size >= 0.99 * np_)


@pytest.mark.parametrize("n,card", [(0, 5), (1, 5), (2049, 3), (100_003, 25), (300_000, 70_000)])
def test_groupby_parity(n, card):
    d = datagen.uniform(n, n + card, 1, 0, card - 1)
    t = an.RESTINY if card <= 256 else an.RESINT
    a = 29 if t is an.RESTINY else 63877
    codes = datagen.encode(d, a, t.width)
    datagen.inject_flips(codes, n + card, 2, datagen.flip_threshold(1e-3))
    meta = an.an_meta(t, a)
    ref = O.groupby_an(O.void(n, 3), O.Col(codes, meta))
    got = capi.groupby_an(capi.Column.void(n, 3), gcol(codes, meta), err_cap=2)
    same(got, ref)
    # binary on top of it
    d2 = datagen.uniform(n, n + card, 3, 0, 4)
    codes2 = datagen.encode(d2, 233, 2)
    datagen.inject_flips(codes2, n + card, 4, datagen.flip_threshold(1e-3))
    m2 = an.an_meta(an.RESTINY, 233)
    ng = ref[1].size
    ref2 = O.groupby2_an(O.void(n), O.Col(codes2, m2), O.Col(ref[0], RESOID), ng, 32)
    got2 = capi.groupby2_an(capi.Column.void(n), gcol(codes2, m2), gcol(ref[0], RESOID), ng, 32, err_cap=2)
    same(got2, ref2)


def test_groupby_binary_many_groups_truncation_parity():
    n_groups = 70_000
    keys = np.concatenate([np.arange(n_groups, dtype=np.uint64), np.arange(100, dtype=np.uint64) + n_groups - 100])
    tail = datagen.encode(keys, 63877, 8)
    meta = an.an_meta(an.RESINT, 63877)
    zero = np.zeros(keys.size, np.uint64)
    ref = O.groupby2_an(O.void(keys.size), O.Col(tail, meta), O.Col(zero, RESOID), 1, 64)
    got = capi.groupby2_an(capi.Column.void(keys.size), gcol(tail, meta), gcol(zero, RESOID), 1, 64)
    same(got, ref)


@pytest.mark.parametrize("n,G", [(0, 3), (1, 1), (2049, 7), (200_003, 125), (100_000, 5000)])
def test_aggregate_sum_grouped_parity(n, G):
    m = datagen.uniform(n, G + n, 1, 90_000, 10_500_000)
    codes = datagen.encode(m, 65117, 8)
    datagen.inject_flips(codes, G + n, 2, datagen.flip_threshold(1e-3))
    gid = datagen.uniform(n, G + n, 3, 0, G)  # G itself is out of range: reported, skipped
    gcodes = enc_oids(gid)
    datagen.inject_flips(gcodes, G + n, 4, datagen.flip_threshold(1e-3))
    meta = an.an_meta(an.RESINT, 65117)
    ref = O.aggregate_sum_grouped_an(O.void(n), O.Col(codes, meta), O.Col(gcodes, RESOID), G)
    got = capi.aggregate_sum_grouped_an(capi.Column.void(n), gcol(codes, meta), gcol(gcodes, RESOID), G, err_cap=2)
    same(got, ref)


@pytest.mark.parametrize("n", SIZES)
def test_mul_sum_parity(n):
    p = datagen.encode(datagen.uniform(n, 41, 1, 0, 10**6), 65117, 8)
    q = datagen.encode(datagen.uniform(n + 3, 41, 2, 0, 10), 233, 2)
    datagen.inject_flips(p, 41, 3, datagen.flip_threshold(1e-3))
    datagen.inject_flips(q, 41, 4, datagen.flip_threshold(1e-3))
    mp, mq = an.an_meta(an.RESINT, 65117), an.an_meta(an.RESTINY, 233)
    ref = O.aggregate_mul_sum_an(O.Col(p, mp), O.Col(q, mq), 5)
    got = capi.aggregate_mul_sum_an(gcol(p, mp), gcol(q, mq), 5)
    same(got, ref)


@pytest.mark.parametrize("op", [0, 1, 2, 3])
@pytest.mark.parametrize("n", [0, 2049, 50_001])
def test_arithmetic_parity(op, n):
    x = datagen.encode(datagen.uniform(n, 51, 1, 0, 10**6), 65117, 8)
    y = datagen.encode(datagen.uniform(n, 51, 2, 0, 10**5), 63877, 8)
    datagen.inject_flips(x, 51, 3, datagen.flip_threshold(1e-3))
    mx, my = an.an_meta(an.RESINT, 65117), an.an_meta(an.RESINT, 63877)
    ref = O.arithmetic_an(O.void(n), O.Col(x, mx), O.void(n), O.Col(y, my), op, an.A_RESULT, 8)
    got = capi.arithmetic_an(capi.Column.void(n), gcol(x, mx), capi.Column.void(n), gcol(y, my), op, an.A_RESULT, 8)
    same(got, ref)


@pytest.mark.parametrize("n", [0, 1, 2049, 100_003])
def test_check_and_decode2_parity(n):
    h = datagen.encode(datagen.uniform(n, 71, 1, 0, 1 << 20), an.A_OID, 8)
    t = datagen.encode(datagen.uniform(n, 71, 2, 0, 0xFFFF), 233, 4)
    datagen.inject_flips(h, 71, 3, datagen.flip_threshold(2e-3))
    datagen.inject_flips(t, 71, 5, datagen.flip_threshold(2e-3))
    mt = an.an_meta(an.RESSHORT, 233)
    ref = O.check_and_decode_an2(O.Col(h, RESOID), O.Col(t, mt))
    got = capi.check_and_decode2_an(gcol(h, RESOID), gcol(t, mt), err_cap=2)
    same(got, ref)


@pytest.mark.parametrize("n1,n2,sb1,sb2", [(0, 10, 0, 0), (10, 0, 3, 0), (2049, 3000, 0, 0), (5000, 4000, 0, 700),
                                           (4000, 5000, 900, 0), (100_003, 90_000, 12, 5), (300, 200, 1000, 0)])
def test_matchjoin_seq_parity(n1, n2, sb1, sb2):
    seed = n1 + n2 + sb1
    h1 = datagen.encode(datagen.uniform(n1, seed, 1, 0, 1 << 20), an.A_OID, 8)
    t2 = datagen.encode(datagen.uniform(n2, seed, 2, 0, 0xFFFF), 233, 4)
    datagen.inject_flips(h1, seed, 3, datagen.flip_threshold(2e-3))
    datagen.inject_flips(t2, seed, 5, datagen.flip_threshold(2e-3))
    mt = an.an_meta(an.RESSHORT, 233)
    f_h = (16089 * an.A_OID_INV) & an.MASK64
    for args in [(0, 0, 8, 0, 0, 4), (1, f_h, 8, 1, (65117 * an.inverse(233, 32)) & 0xFFFFFFFF, 4)]:
        ref = O.matchjoin_seq_an(O.Col(h1, RESOID), O.void(n1, sb1), O.void(n2, sb2), O.Col(t2, mt), *args)
        got = capi.matchjoin_seq_an(gcol(h1, RESOID), capi.Column.void(n1, sb1), capi.Column.void(n2, sb2),
                                    gcol(t2, mt), *args, err_cap=2)
        same(got, ref)
