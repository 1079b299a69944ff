"""GPU parity of groupbyAN (unary and binary) against the oracle, bit-exact, over the shapes its
kernels branch on: the tail's element width (1, 2, 4, 8 bytes; plain and AN-encoded), the
number of groups (the LDS dictionary up to 512 groups, the HBM table beyond, the ordered
compaction beyond 8192), keys whose key + 1 wraps (all-ones values), row counts that end inside
a batch, and columns that start off a 16-byte boundary (element loads instead of vector loads)."""
import numpy as np
import pytest

import oracle_lib as O
from ahead_amd import an, capi, datagen

pytestmark = pytest.mark.gpu

RESOID = an.an_meta(an.RESOID, an.A_OID)


class _At:
    """A device buffer seen from a byte offset (a column that does not start 16-byte aligned)."""

    def __init__(self, buf, off):
        self.buf, self.ptr = buf, buf.ptr + off


def gcol(arr, meta, offset_elems=0):
    if offset_elems == 0:
        return capi.Column.from_numpy(arr, meta)
    pad = np.zeros(arr.size + offset_elems, arr.dtype)
    pad[offset_elems:] = arr
    buf = capi.DeviceBuffer.from_numpy(pad)
    return capi.Column(_At(buf, offset_elems * arr.dtype.itemsize), arr.size, meta)


def same(got, ref):
    assert len(got) == len(ref)
    for i, (g, r) in enumerate(zip(got, ref)):
        assert np.array_equal(np.asarray(g).astype(np.uint64), np.asarray(r).astype(np.uint64)), i


def _values(n, card, seed, width, ones):
    v = datagen.uniform(n, seed, 1, 0, card - 1)
    if ones and n:
        # the all-ones value of the element width: its key + 1 wraps to 0
        top = np.uint64((1 << (8 * width)) - 1)
        idx = datagen.uniform(max(1, n // 5000), seed, 2, 0, n - 1)
        v[idx] = top
    return v


PLAIN = [(an.TINYINT, np.uint8), (an.SHORTINT, np.uint16), (an.INT, np.uint32), (an.BIGINT, np.uint64)]


@pytest.mark.parametrize("card", [3, 400, 3000, 20_000])
@pytest.mark.parametrize("ti", range(4))
def test_groupby_plain_widths(ti, card):
    t, dt = PLAIN[ti]
    if t.width == 1 and card > 256:
        pytest.skip("at most 256 one-byte values")
    n = 200_003
    vals = _values(n, card, 101 + ti + card, t.width, ones=True).astype(dt)
    meta = an.plain_meta(t)
    ref = O.groupby_an(O.void(n, 7), O.Col(vals, meta))
    got = capi.groupby_an(capi.Column.void(n, 7), gcol(vals, meta), err_cap=2)
    same(got, ref)
    # binary on top: the second key's width sets the larger type
    ng = ref[1].size
    vals2 = _values(n, 9, 201 + ti + card, t.width, ones=False).astype(dt)
    ref2 = O.groupby2_an(O.void(n), O.Col(vals2, meta), O.Col(ref[0], RESOID), ng, t.larger_bits)
    got2 = capi.groupby2_an(capi.Column.void(n), gcol(vals2, meta), gcol(ref[0], RESOID), ng, t.larger_bits,
                            err_cap=2)
    same(got2, ref2)


ENCODED = [(an.RESTINY, 29), (an.RESSHORT, 233), (an.RESINT, 65117), (an.RESBIGINT, 65117)]


@pytest.mark.parametrize("offset", [0, 1])
@pytest.mark.parametrize("card", [25, 500, 513, 9000])
@pytest.mark.parametrize("ei", range(4))
def test_groupby_encoded_widths(ei, card, offset):
    t, a = ENCODED[ei]
    if t.unenc_bits == 8 and card > 256:
        pytest.skip("at most 256 one-byte values")
    n = 150_001 + offset
    vals = datagen.uniform(n, 300 + ei + card, 1, 0, card - 1)
    codes = datagen.encode(vals, a, t.width)
    datagen.inject_flips(codes, 300 + ei + card, 2, datagen.flip_threshold(1e-3))
    meta = an.an_meta(t, a)
    ref = O.groupby_an(O.void(n, 3), O.Col(codes, meta))
    got = capi.groupby_an(capi.Column.void(n, 3), gcol(codes, meta, offset), err_cap=4)
    same(got, ref)
    ng = ref[1].size
    grp = ref[0].copy()
    datagen.inject_flips(grp, 400 + ei + card, 3, datagen.flip_threshold(1e-4))
    ref2 = O.groupby2_an(O.void(n), O.Col(codes, meta), O.Col(grp, RESOID), ng, t.larger_bits)
    got2 = capi.groupby2_an(capi.Column.void(n), gcol(codes, meta, offset), gcol(grp, RESOID, offset), ng,
                            t.larger_bits, err_cap=4)
    same(got2, ref2)


@pytest.mark.parametrize("n", [1, 63, 64, 1023, 1024, 1025, 4097])
def test_groupby_batch_edges(n):
    """Row counts around one lane run and one wave batch; keys first seen in the last rows."""
    vals = (np.arange(n, dtype=np.uint64) * np.uint64(7)) % np.uint64(97)
    vals[-1] = 200
    codes = datagen.encode(vals, 29, 2)
    meta = an.an_meta(an.RESTINY, 29)
    ref = O.groupby_an(O.void(n), O.Col(codes, meta))
    got = capi.groupby_an(capi.Column.void(n), gcol(codes, meta), err_cap=2)
    same(got, ref)


def test_groupby_many_new_keys_per_wave():
    """More distinct keys than a wave's set holds: the set starts again and keys are listed
    twice; ascending, descending and interleaved first occurrences."""
    n = 300_000
    base = np.arange(n, dtype=np.uint64)
    for vals in (base % np.uint64(40_000), (np.uint64(n) - base) % np.uint64(5_000), (base * np.uint64(7919)) % np.uint64(1500)):
        codes = datagen.encode(vals, 65117, 8)
        meta = an.an_meta(an.RESINT, 65117)
        ref = O.groupby_an(O.void(n), O.Col(codes, meta))
        got = capi.groupby_an(capi.Column.void(n), gcol(codes, meta), err_cap=2)
        same(got, ref)


@pytest.mark.parametrize("card", [25, 300])
def test_groupby_dictionary_overflow_retry(card):
    """A dictionary of 16 slots cannot hold the keys: the call reports status 1 (overflow) and the
    wrapper calls again with a larger table; the second result is the reference's."""
    n = 50_003
    vals = datagen.uniform(n, 500 + card, 1, 0, card - 1)
    codes = datagen.encode(vals, 233, 4)
    datagen.inject_flips(codes, 500 + card, 2, datagen.flip_threshold(1e-3))
    meta = an.an_meta(an.RESSHORT, 233)
    ref = O.groupby_an(O.void(n), O.Col(codes, meta))
    got = capi.groupby_an(capi.Column.void(n), gcol(codes, meta), err_cap=4, table_bits=4)
    same(got, ref)
    ng = ref[1].size
    ref2 = O.groupby2_an(O.void(n), O.Col(codes, meta), O.Col(ref[0], RESOID), ng, an.RESSHORT.larger_bits)
    got2 = capi.groupby2_an(capi.Column.void(n), gcol(codes, meta), gcol(ref[0], RESOID), ng,
                            an.RESSHORT.larger_bits, err_cap=4, table_bits=4)
    same(got2, ref2)
