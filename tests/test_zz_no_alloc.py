"""The allocation-free select (ahead_hip_select_an_noalloc) returns exactly what the workspace select
(ahead_hip_select_an) returns: OIDs, error positions, counts and the fused decoded sum, on two calls
in a row with the second on a smaller input, for every code width, with and without the sum."""
import ctypes

import numpy as np
import pytest

from ahead_amd import an, capi, datagen

pytestmark = pytest.mark.gpu

CASES = [  # (code width, A, value range, predicate)
    (4, 233, (1, 50), (capi.GE, 1, capi.LE, 24, capi.COMB_AND)),
    (2, 29, (0, 255), (capi.LT, 100, capi.LT, 0, capi.COMB_NONE)),
    (8, 65117, (90_000, 10_500_000), (capi.LE, 2_000_000, capi.GE, 9_000_000, capi.COMB_OR)),
]


def _run(entry, col, pred, with_sum, lib, ws):
    op1, v1, op2, v2, comb = pred
    a = col.meta.a
    n = col.n
    oids = capi.DeviceBuffer(8 * max(n, 1))
    err_cap = 1 << 14
    err = capi.DeviceBuffer(8 * err_cap)
    cnt = capi.DeviceBuffer(16)
    dsum = capi.DeviceBuffer(8) if with_sum else None
    d = col.desc()
    args = [ctypes.byref(d), op1, (v1 * a) & an.MASK64, op2, (v2 * a) & an.MASK64, comb, an.A_OID, an.A_OID_ERR,
            oids.ptr, n, err.ptr, err_cap, cnt.ptr, dsum.ptr if dsum else None]
    if entry == "ws":
        capi.check(lib.ahead_hip_select_an(*args, ws.h, capi.current_stream()))
    else:
        capi.check(lib.ahead_hip_select_an_noalloc(*args, capi.current_stream()))
    capi.synchronize()
    n_oids, n_err = (int(x) for x in cnt.to_numpy(np.uint64, 2))
    assert n_err <= err_cap
    total = int(dsum.to_numpy(np.uint64, 1)[0]) if dsum else None
    return oids.to_numpy(np.uint64, n_oids), err.to_numpy(np.uint64, n_err), total


@pytest.mark.parametrize("width,a,rng,pred", CASES)
def test_noalloc_select_equals_workspace_select(width, a, rng, pred):
    lib = capi.load()
    ws = capi.workspace()
    types = {2: an.RESTINY, 4: an.RESSHORT, 8: an.RESINT}
    meta = an.an_meta(types[width], a)
    for call, n in enumerate((3_000_017, 250_003)):  # two calls in a row, the second smaller
        print(f"noalloc: width {width} call {call} n {n}", flush=True)
        codes = datagen.encode(datagen.uniform(n, 5 + call, width, *rng), a, width)
        datagen.inject_flips(codes, 5 + call, 40 + width, datagen.flip_threshold(1e-4))
        col = capi.Column.from_numpy(codes, meta)
        for with_sum in (True, False):
            ref = _run("ws", col, pred, with_sum, lib, ws)
            got = _run("noalloc", col, pred, with_sum, lib, ws)
            assert np.array_equal(got[0], ref[0]), ("oids", width, n, with_sum)
            assert np.array_equal(got[1], ref[1]), ("errors", width, n, with_sum)
            assert got[2] == ref[2], ("sum", width, n, with_sum)
            assert ref[1].size > 0 and ref[0].size > 0


def test_noalloc_select_empty_and_tiny_columns():
    lib = capi.load()
    ws = capi.workspace()
    meta = an.an_meta(an.RESSHORT, 233)
    for n in (0, 1, 4097):
        codes = datagen.encode(datagen.uniform(n, 9, 4, 1, 50), 233, 4)
        col = capi.Column.from_numpy(codes, meta)
        pred = (capi.GE, 1, capi.LE, 24, capi.COMB_AND)
        ref = _run("ws", col, pred, True, lib, ws)
        got = _run("noalloc", col, pred, True, lib, ws)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]) and got[2] == ref[2], n
