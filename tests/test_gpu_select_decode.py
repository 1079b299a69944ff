"""GPU parity of selectAN and the check/decode family against the CPU oracle (bit-exact)."""
import numpy as np
import pytest

import oracle_lib as O
from ahead_amd import an, capi, datagen

pytestmark = pytest.mark.gpu

TYPES = [(an.RESTINY, 29), (an.RESTINY, 233), (an.RESSHORT, 233), (an.RESSHORT, 63877), (an.RESINT, 63877),
         (an.RESINT, 65117)]
# sizes around the tile boundaries (8192 / 4096 / 2048 rows per tile for 2 / 4 / 8-byte codes)
SIZES = [0, 1, 7, 2047, 2048, 2049, 4095, 4097, 8191, 8193, 65_537, 300_001]


def make_codes(t, a, n, seed, rate, hi=None):
    hi = t.max_u if hi is None else hi
    d = datagen.uniform(n, seed, 1, 0, hi)
    codes = datagen.encode(d, a, t.width)
    flips = datagen.inject_flips(codes, seed, 2, datagen.flip_threshold(rate))
    return d, codes, flips


def test_device_generator_matches_host_twin():
    n = 100_003
    for width, lo, hi, a in [(2, 0, 255, 29), (4, 1, 50, 233), (8, 90_000, 10_500_000, 65117), (1, 0, 24, 0)]:
        buf = capi.DeviceBuffer(width * n)
        capi.gen_uniform(buf, width, n, 1234, 7, lo, hi, a)
        capi.synchronize()
        host = datagen.uniform(n, 1234, 7, lo, hi)
        host = datagen.encode(host, a, width) if a else host.astype(datagen.DTYPES[width])
        assert np.array_equal(buf.to_numpy(host.dtype, n), host)
        if width > 1:
            capi.inject_flips(buf, width, n, 99, 3, datagen.flip_threshold(1e-3))
            capi.synchronize()
            pos = datagen.inject_flips(host, 99, 3, datagen.flip_threshold(1e-3))
            assert pos.size > 10
            assert np.array_equal(buf.to_numpy(host.dtype, n), host)


@pytest.mark.parametrize("t,a", TYPES)
@pytest.mark.parametrize("n", SIZES)
def test_select_range_parity(t, a, n):
    d, codes, flips = make_codes(t, a, n, seed=n + a, rate=2e-3)
    meta = an.an_meta(t, a)
    lo, hi = (t.max_u // 5) * a, (t.max_u // 2) * a
    ref = O.select_an(O.Col(codes, meta), capi.GE, lo, capi.LE, hi, capi.COMB_AND)
    got = capi.select_an(capi.Column.from_numpy(codes, meta), capi.GE, lo, capi.LE, hi, capi.COMB_AND,
                         with_sum=True)
    assert np.array_equal(got[0], ref[0])
    assert np.array_equal(got[1], ref[1])
    assert got[2] == ref[2]


@pytest.mark.parametrize("op", range(6))
@pytest.mark.parametrize("comb", [0, 1, 2])
@pytest.mark.parametrize("t,a", [(an.RESTINY, 233), (an.RESSHORT, 233), (an.RESINT, 65117)])
def test_select_all_predicates(op, comb, t, a):
    n = 20_011
    _, codes, _ = make_codes(t, a, n, seed=op * 3 + comb, rate=1e-3, hi=100)
    meta = an.an_meta(t, a).with_seqbase(17)
    th1, th2 = 40 * a, 60 * a
    ref = O.select_an(O.Col(codes, meta), op, th1, (op + 2) % 6, th2, comb)
    got = capi.select_an(capi.Column.from_numpy(codes, meta), op, th1, (op + 2) % 6, th2, comb, with_sum=True)
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)


def test_select_threshold_extremes():
    # LT 0 and GT maxcode are empty; NE of a value absent from the column keeps everything
    t, a = an.RESSHORT, 233
    _, codes, _ = make_codes(t, a, 10_000, seed=3, rate=0)
    meta = an.an_meta(t, a)
    col = capi.Column.from_numpy(codes, meta)
    for op, th in [(capi.LT, 0), (capi.GT, 0xFFFFFFFF), (capi.NE, 1), (capi.LE, 0xFFFFFFFF), (capi.GE, 0)]:
        ref = O.select_an(O.Col(codes, meta), op, th)
        got = capi.select_an(col, op, th, with_sum=True)
        for g, r in zip(got, ref):
            assert np.array_equal(g, r)


def test_select_misaligned_input_and_error_overflow_retry():
    t, a = an.RESSHORT, 233
    _, codes, flips = make_codes(t, a, 50_001, seed=8, rate=2e-2)
    meta = an.an_meta(t, a)
    full = capi.Column.from_numpy(codes, meta)
    # a column view starting one element into the buffer: not 16-byte aligned
    view = capi.Column(full.buf, codes.size - 1, meta)
    import ctypes

    class Off(capi.DeviceBuffer):
        pass

    off = Off.__new__(Off)
    off.t = full.buf.t[4:]
    off.nbytes = full.buf.nbytes - 4
    view.buf = off
    ref = O.select_an(O.Col(codes[1:], meta), capi.LT, 25 * a)
    got = capi.select_an(view, capi.LT, 25 * a, with_sum=True, err_cap=3)  # forces the re-run path
    assert len(ref[1]) > 3
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)


def test_select_config1_size():
    # the select+sum workload: 6M 16-bit values in [1,50], A=233 in uint32, 1e-6 flips
    n = 6_000_000
    d = datagen.uniform(n, 42, 1, 1, 50)
    codes = datagen.encode(d, 233, 4)
    flips = datagen.inject_flips(codes, 42, 2, datagen.flip_threshold(1e-6))
    meta = an.an_meta(an.RESSHORT, 233)
    got = capi.select_an(capi.Column.from_numpy(codes, meta), capi.GE, 233, capi.LE, 24 * 233, capi.COMB_AND,
                         with_sum=True)
    ref = O.select_an(O.Col(codes, meta), capi.GE, 233, capi.LE, 24 * 233, capi.COMB_AND)
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)
    assert np.array_equal(got[1], flips * np.uint64(an.A_OID_ERR))


@pytest.mark.parametrize("t,a", TYPES)
@pytest.mark.parametrize("n", [0, 1, 4097, 100_003])
def test_check_decode_family_parity(t, a, n):
    _, codes, flips = make_codes(t, a, n, seed=n + 1, rate=3e-3)
    meta = an.an_meta(t, a)
    col = capi.Column.from_numpy(codes, meta)
    oc = O.Col(codes, meta)
    assert np.array_equal(capi.check_an(col), O.check_an(oc))
    for got, ref in zip(capi.decode_an(col), O.decode_an(oc)):
        assert np.array_equal(got.astype(np.uint64), ref)
    for got, ref in zip(capi.check_and_decode_an(col), O.check_and_decode_an(oc)):
        assert np.array_equal(got.astype(np.uint64), ref)


@pytest.mark.parametrize("src", [an.TINYINT, an.SHORTINT, an.INT])
def test_encode_parity(src):
    n = 12_345
    d = datagen.uniform(n, 5, 1, 0, src.max_u).astype(datagen.DTYPES[src.width])
    meta = an.plain_meta(src)
    for ow, a in [(2, 29), (4, 233), (8, 65117)]:
        got = capi.encode_an(capi.Column.from_numpy(d, meta), a, ow)
        ref = O.encode_an(O.Col(d, meta), a, ow)
        assert np.array_equal(got.astype(np.uint64), ref)


@pytest.mark.parametrize("max_sub", [2, 4, 16])
@pytest.mark.parametrize("t,a", [(an.RESTINY, 233), (an.RESSHORT, 233), (an.RESINT, 65117)])
def test_select_multi_subtile_tiles(max_sub, t, a):
    # force look-back tiles of several 16 KiB sub-tiles at a small size (min_tiles = 3)
    n = 1_000_003
    _, codes, _ = make_codes(t, a, n, seed=max_sub, rate=5e-4, hi=100)
    meta = an.an_meta(t, a)
    ref = O.select_an(O.Col(codes, meta), capi.LT, 37 * a, capi.GT, 80 * a, capi.COMB_OR)
    with capi.tiling(max_sub, 3):
        got = capi.select_an(capi.Column.from_numpy(codes, meta), capi.LT, 37 * a, capi.GT, 80 * a, capi.COMB_OR,
                             with_sum=True, err_cap=7)
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)
