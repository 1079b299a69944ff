"""Pins the CPU oracle's AN codec, check/decode and selectAN (CPU only).

The reference ships no golden vectors for these operators, so the oracle is pinned by
  * the reference's own constant tables (A / A^-1, include/util/resilience.hpp:56-109),
  * closed forms: every single-bit flip of a valid code is detected when 2^k / A > maxU,
  * textbook results computed independently with numpy on the decoded values,
  * the A = 1 special case (the unhardened operators).
"""
import numpy as np
import pytest

import oracle_lib as O
from ahead_amd import an, datagen

TABLES = [
    (an.RESTINY_AS, an.RESTINY_AINVS, 16),
    (an.RESTINY32_AS, an.RESTINY32_AINVS, 32),
    (an.RESSHORT_AS, an.RESSHORT_AINVS, 32),
    (an.RESINT_AS, an.RESINT_AINVS, 64),
]


@pytest.mark.parametrize("As,Ainvs,bits", TABLES)
def test_inverse_reproduces_reference_tables(As, Ainvs, bits):
    for a, ainv in zip(As, Ainvs):
        assert O.inverse(a, bits) == ainv
        assert (a * ainv) % (1 << bits) == 1
        assert pow(a, -1, 1 << bits) == ainv  # independent: Python's modular inverse
        assert an.inverse(a, bits) == ainv


def test_inverse_of_config_constants():
    # A values of the benchmark workloads, cross-checked against Python's pow(a, -1, m)
    for a, bits, expect in [(233, 32, 0x1FDCD759), (29, 16, 0xC235), (63877, 64, 0x82038D02D142174D),
                            (65117, 64, 0x3E36D3BC8D9BC5F5), (16089, 64, 0x227F016A87066169)]:
        assert O.inverse(a, bits) == expect == pow(a, -1, 1 << bits)
    assert O.inverse(2, 32) == 0  # even A has no inverse (ext_euclidean returns 0)


SINGLE_FLIP_CASES = [
    (an.RESTINY, 29), (an.RESTINY, 233), (an.RESSHORT, 233), (an.RESSHORT, 63877),
    (an.RESINT, 63877), (an.RESINT, 65117),
]


@pytest.mark.parametrize("t,a", SINGLE_FLIP_CASES)
def test_every_single_bit_flip_is_detected(t, a):
    # closed form: c' = A*d +- 2^j; A odd > 1 cannot divide 2^j, so c' * A^-1 >= 2^k / A > maxU
    k = 8 * t.width
    assert (1 << k) // a > t.max_u
    d = datagen.uniform(2000, 7, 1, 0, t.max_u)
    codes = datagen.encode(d, a, t.width)
    flipped = np.concatenate([codes ^ np.array(1 << j, dtype=codes.dtype) for j in range(k)])
    meta = an.an_meta(t, a)
    assert O.check_an(O.Col(codes, meta)).size == 0
    err = O.check_an(O.Col(flipped, meta))
    assert err.size == flipped.size  # every flipped code reported, in order
    assert np.array_equal(err, np.arange(flipped.size, dtype=np.uint64) * np.uint64(an.A_OID_ERR))


def _config1(n=50_000, seed=11, rate=0.0):
    d = datagen.uniform(n, seed, 1, 1, 50)
    codes = datagen.encode(d, 233, 4)
    flips = datagen.inject_flips(codes, seed, 2, datagen.flip_threshold(rate))
    return d, codes, flips


def _np_pred(op, x, th):
    return {0: x < th, 1: x <= th, 2: x > th, 3: x >= th, 4: x == th, 5: x != th}[op]


@pytest.mark.parametrize("op", range(6))
def test_select_single_predicate_matches_numpy(op):
    d, codes, _ = _config1()
    th = 24 * 233
    oids, err, s = O.select_an(O.Col(codes, an.an_meta(an.RESSHORT, 233)), op, th)
    mask = _np_pred(op, d, 24)  # on fault-free data, comparing codes == comparing values
    expect = np.nonzero(mask)[0].astype(np.uint64) * np.uint64(an.A_OID)
    assert np.array_equal(oids, expect)
    assert err.size == 0
    assert s == int(d[mask].sum())


@pytest.mark.parametrize("comb", [1, 2])
@pytest.mark.parametrize("ops", [(3, 1), (2, 0), (4, 5), (1, 3)])
def test_select_two_predicates_matches_numpy(comb, ops):
    d, codes, _ = _config1(seed=5)
    t1, t2 = 10, 30
    oids, err, s = O.select_an(O.Col(codes, an.an_meta(an.RESSHORT, 233)), ops[0], t1 * 233, ops[1], t2 * 233,
                               comb)
    m1, m2 = _np_pred(ops[0], d, t1), _np_pred(ops[1], d, t2)
    mask = (m1 & m2) if comb == 1 else (m1 | m2)
    assert np.array_equal(oids, np.nonzero(mask)[0].astype(np.uint64) * np.uint64(an.A_OID))
    assert s == int(d[mask].sum())


def test_select_with_flips_reports_sorted_positions_and_keeps_corrupt_rows():
    d, codes, flips = _config1(n=200_000, seed=3, rate=1e-4)
    assert flips.size > 5
    meta = an.an_meta(an.RESSHORT, 233)
    oids, err, s = O.select_an(O.Col(codes, meta), 3, 1 * 233, 1, 24 * 233, 1)
    assert np.array_equal(err, flips * np.uint64(an.A_OID_ERR))
    # predicate evaluated on the raw (possibly corrupt) codes, as the reference does
    mask = (codes >= 233) & (codes <= 24 * 233)
    assert np.array_equal(oids, np.nonzero(mask)[0].astype(np.uint64) * np.uint64(an.A_OID))
    clean = mask.copy()
    clean[flips.astype(np.int64)] = False
    assert s == int(d[clean].sum())


def test_select_seqbase_and_oid_wrap():
    # head() is an oid_t: (uint32)(seqbase + pos) before the 64-bit multiply
    d = np.arange(8, dtype=np.uint64)
    codes = datagen.encode(d, 233, 4)
    meta = an.an_meta(an.RESSHORT, 233).with_seqbase(0xFFFFFFFE)
    oids, _, _ = O.select_an(O.Col(codes, meta), 1, 7 * 233)
    heads = (np.arange(8, dtype=np.uint64) + np.uint64(0xFFFFFFFE)) & np.uint64(0xFFFFFFFF)
    assert np.array_equal(oids, heads * np.uint64(an.A_OID))


def test_select_with_a_equal_one_is_the_plain_filter():
    d = datagen.uniform(10_000, 9, 1, 0, 0xFFFF)
    meta = an.ColMeta(an.RESSHORT, 1, 1)
    oids, err, s = O.select_an(O.Col(d.astype(np.uint32), meta), 0, 1000)
    assert err.size == 0
    assert np.array_equal(oids, np.nonzero(d < 1000)[0].astype(np.uint64) * np.uint64(an.A_OID))


@pytest.mark.parametrize("t,a", SINGLE_FLIP_CASES)
def test_check_and_decode_matches_numpy(t, a):
    d = datagen.uniform(30_000, 4, 1, 0, t.max_u)
    codes = datagen.encode(d, a, t.width)
    flips = datagen.inject_flips(codes, 4, 2, datagen.flip_threshold(1e-3))
    meta = an.an_meta(t, a)
    dec, err = O.check_and_decode_an(O.Col(codes, meta))
    keep = np.ones(d.size, bool)
    keep[flips.astype(np.int64)] = False
    assert np.array_equal(dec, d[keep])
    assert np.array_equal(err, flips * np.uint64(an.A_OID_ERR))
    dec_all, err2 = O.decode_an(O.Col(codes, meta))
    assert np.array_equal(err2, err)
    assert np.array_equal(dec_all[keep], d[keep])
    # a decoded corrupt code is the modular product, truncated to the unencoded width
    inv = an.inverse(a, 8 * t.width)
    m = (1 << (8 * t.width)) - 1
    for p in flips[:5].astype(np.int64):
        assert int(dec_all[p]) == ((int(codes[p]) * inv) & m) & ((1 << t.unenc_bits) - 1)


def test_encode_matches_numpy():
    d = datagen.uniform(1000, 2, 1, 0, 255).astype(np.uint8)
    out = O.encode_an(O.Col(d, an.plain_meta(an.TINYINT)), 233, 2)
    assert np.array_equal(out, (d.astype(np.uint64) * 233) & 0xFFFF)
