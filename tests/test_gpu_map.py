"""The streaming engine (maptile.hpp) on fetchjoinAN, where a corrupt fetched code or an OID past
the column drops the row: no drops, few drops (entries moved into place), many drops (every wave
writes again), an output smaller than the input (the tile engine runs instead), all bit-exact
against the oracle, error positions included."""
import ctypes

import numpy as np
import pytest

import oracle_lib as O
from ahead_amd import an, capi, datagen

pytestmark = pytest.mark.gpu

RESOID = an.an_meta(an.RESOID, an.A_OID)
META = an.an_meta(an.RESSHORT, 233)


def enc_oids(pos):
    with np.errstate(over="ignore"):
        return pos.astype(np.uint64) * np.uint64(an.A_OID)


def inputs(n, drops, past_end, seed):
    n_col, bad = 5000, 17
    rng = np.random.default_rng(seed)
    col = datagen.encode(rng.integers(0, 0xFFFF, n_col).astype(np.uint64), 233, 4)
    col[bad] ^= np.uint32(1 << 5)  # the one corrupt code
    pos = rng.integers(0, n_col, n).astype(np.uint64)
    pos[pos == bad] = bad + 1
    hit = rng.choice(n, drops + past_end, replace=False)
    pos[hit[:drops]] = bad                            # rows that fetch the corrupt code
    pos[hit[drops:]] = n_col + rng.integers(0, 50, past_end).astype(np.uint64)  # OIDs past the column
    return enc_oids(pos), col


@pytest.mark.parametrize("n,drops,past_end", [
    (1_000_003, 0, 0), (1_000_003, 1, 0), (1_000_003, 10, 3), (1_000_003, 64, 0), (1_000_003, 60, 5),
    (1_000_003, 400, 20), (70_001, 0, 1), (513, 2, 1),
])
def test_fetchjoin_drop_paths(n, drops, past_end):
    oids, col = inputs(n, drops, past_end, seed=n + drops)
    print(f"fetchjoin n={n} drops={drops} past_end={past_end}", flush=True)
    for reenc, ar in [(0, 0), (1, 63877)]:
        ref = O.fetchjoin_an(O.Col(oids, RESOID), O.Col(col, META), reenc, ar)
        got = capi.fetchjoin_an(capi.Column.from_numpy(oids, RESOID), capi.Column.from_numpy(col, META), reenc, ar,
                                err_cap=1 << 12)
        assert len(got[0]) == n - drops - past_end
        for i, (g, r) in enumerate(zip(got, ref)):
            assert np.array_equal(np.asarray(g).astype(np.uint64), np.asarray(r).astype(np.uint64)), i


def test_fetchjoin_output_smaller_than_input_uses_tile_engine():
    n = 300_007
    oids, col = inputs(n, 5, 2, seed=3)
    ref = O.fetchjoin_an(O.Col(oids, RESOID), O.Col(col, META), 0, 0)
    lib, ws = capi.load(), capi.workspace()
    do, dc = capi.Column.from_numpy(oids, RESOID), capi.Column.from_numpy(col, META)
    cap = n // 2
    out = capi.DeviceBuffer(4 * cap)
    e = [capi.DeviceBuffer(8 * 64) for _ in range(2)]
    cnt = capi.DeviceBuffer(8 * 3)
    d1, d2 = do.desc(), dc.desc()
    capi.check(lib.ahead_hip_fetchjoin_an(ctypes.byref(d1), ctypes.byref(d2), 0, 0, an.A_OID_ERR,
                                          capi.Buf(out.ptr, cap), capi.Buf(e[0].ptr, 64), capi.Buf(e[1].ptr, 64),
                                          cnt.ptr, ws.h, capi.current_stream()))
    capi.synchronize()
    c = cnt.to_numpy(np.uint64, 3)
    assert [int(x) for x in c] == [len(ref[0]), len(ref[1]), len(ref[2])]
    assert np.array_equal(out.to_numpy(np.uint32, cap), ref[0][:cap].astype(np.uint32))
    assert np.array_equal(e[0].to_numpy(np.uint64, int(c[1])), ref[1])
    assert np.array_equal(e[1].to_numpy(np.uint64, int(c[2])), ref[2])
