"""Device-resident BATs and the hardened operators on them, for Python callers.

The Python counterpart of include/ahead/ (the C++ templates): a BAT is a pair of device columns
(capi.Column; a void side is virtual), `mirror_head` / `reverse` / `clear_head` share columns as
the reference's TempBAT does (include/column_storage/TempBat.h:157-178), and every operator keeps
its results in HBM and returns error-position vectors as numpy arrays (None for a side that is not
AN-encoded, lib/column_operators/AN/ANhelper.tcc:107-109).  The operators are named after
include/column_operators/OperatorsAN.hpp:33-298 and take the same arguments; each one runs its
C-ABI entry point on the current stream and waits for the counts, as the reference's operators
are synchronous.  Output buffers are sized from the inputs and grown (the call repeated) when a
count exceeds them.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import an, capi
from .capi import Buf, Column, DeviceBuffer, check, current_stream, load, workspace

LT, LE, GT, GE, EQ, NE = capi.LT, capi.LE, capi.GT, capi.GE, capi.EQ, capi.NE
NONE_, AND, OR = capi.COMB_NONE, capi.COMB_AND, capi.COMB_OR
ADD, SUB, MUL, DIV = 0, 1, 2, 3

A_OID_ERR = an.A_OID_ERR     # error positions (AsBFW.back())
A_OID = an.A_OID             # OID heads, group ids (As.back())
RESOID_META = an.an_meta(an.RESOID, an.A_OID)

# TypeMap<T>::v2_encoded_t and T::v2_select_t (include/util/resilience.hpp:278-432, v2types.hpp)
_ENCODED = {"tinyint": an.RESTINY, "shortint": an.RESSHORT, "int": an.RESINT, "bigint": an.RESBIGINT,
            "oid": an.RESOID, "void": an.RESOID}


def encoded_type(t: an.ColType) -> an.ColType:
    return t if t.encoded else _ENCODED[t.name]


def select_type(t: an.ColType) -> an.ColType:
    return an.OID if t.name == "void" else t


@dataclass
class BAT:
    head: Column
    tail: Column | None

    def size(self) -> int:
        return self.head.n if self.head.buf is not None or self.tail is None else self.tail.n

    def mirror_head(self) -> "BAT":
        return BAT(self.head, self.head)

    def reverse(self) -> "BAT":
        return BAT(self.tail, self.head)

    def clear_head(self) -> "BAT":
        return BAT(Column.void(self.tail.n, self.head.meta.seqbase if self.head.buf is None else 0), self.tail)

    def tail_numpy(self) -> np.ndarray:
        return self.tail.to_numpy()

    def head_numpy(self) -> np.ndarray:
        return self.head.to_numpy()


def column(values: np.ndarray, meta: an.ColMeta) -> BAT:
    """A <void, T> BAT over host values (codes for AN-encoded types), copied to the device."""
    return BAT(Column.void(values.size), Column.from_numpy(values, meta))


def _errs(bufs, counts, encoded) -> list:
    return [b.to_numpy(np.uint64, c) if e else None for b, c, e in zip(bufs, counts, encoded)]


class _Runner:
    """Allocates outputs, runs an entry point, reads counts, grows what overflowed, repeats."""

    def __init__(self, ncounts: int):
        self.lib, self.ws = load(), workspace()
        self.cnt = DeviceBuffer(8 * ncounts)
        self.ncounts = ncounts

    def run(self, call, caps: list[list[int]], count_of: list[int], shrink: tuple = ()):
        """caps: [elem_bytes, cap] per output; count_of[i]: counts index bounding output i.
        Outputs listed in `shrink` (sized by an upper bound, e.g. every row of a join's probe
        side) are copied into buffers of their actual size when they came out at most half full,
        so that results held by a caller occupy what they use and the bound-sized buffer goes
        back to the allocator for the next operator."""
        while True:
            bufs = [DeviceBuffer(max(8, eb * cap)) for eb, cap in caps]
            check(call(bufs, self.cnt.ptr))
            c = [int(x) for x in self.cnt.to_numpy(np.uint64, self.ncounts)]
            grow = False
            for i, (eb, cap) in enumerate(caps):
                if c[count_of[i]] > cap:
                    caps[i][1] = c[count_of[i]]
                    grow = True
            if not grow:
                break
        for i in shrink:
            need = max(8, caps[i][0] * c[count_of[i]])
            if 2 * need <= bufs[i].nbytes:
                bufs[i] = bufs[i].prefix_copy(need)
        return bufs, c


def _b(buf: DeviceBuffer, cap: int) -> Buf:
    return Buf(buf.ptr, cap)


ERR_CAP = 1024


def selectAN(bat: BAT, op1: int, th1: int, op2: int = LT, th2: int = 0, comb: int = NONE_,
             a_oid: int = A_OID_ERR) -> tuple[BAT, np.ndarray]:
    """selectAN<Op> / selectAN<Op1, Op2, OpCombine> (selectAN_scalar.tcc:66-95, 149-181) on a
    <void, T> BAT with encoded thresholds: (BAT<resoid, T> with only the OIDs, error positions)."""
    r = _Runner(2)
    n = bat.tail.n
    d = bat.tail.desc()

    caps = [[8, max(n, 1)], [8, ERR_CAP]]

    def call2(b, cnt):
        return r.lib.ahead_hip_select_an(ctypes.byref(d), op1, th1 & an.MASK64, op2, th2 & an.MASK64, comb, A_OID,
                                         a_oid, b[0].ptr, caps[0][1], b[1].ptr, caps[1][1], cnt, None, r.ws.h,
                                         current_stream())

    bufs, c = r.run(call2, caps, [0, 1], shrink=(0,))
    oids = Column(bufs[0], c[0], RESOID_META)
    return BAT(oids, None), bufs[1].to_numpy(np.uint64, c[1])


def fetchjoinAN(oids: BAT, col: BAT, reenc: tuple[int, int] | None = None, a_oid: int = A_OID_ERR):
    """fetchjoinAN (matchjoinAN.tcc:312-353): (BAT<void, TypeMap<T2>::v2_encoded_t>, e1, e2)."""
    r = _Runner(3)
    t_out = encoded_type(col.tail.meta.type)
    do, dc = oids.tail.desc(), col.tail.desc()
    n = oids.tail.n
    caps = [[t_out.width, max(n, 1)], [8, ERR_CAP], [8, ERR_CAP]]
    a_re = reenc[0] if reenc else 0

    def call(b, cnt):
        return r.lib.ahead_hip_fetchjoin_an(ctypes.byref(do), ctypes.byref(dc), 1 if reenc else 0, a_re & an.MASK64,
                                            a_oid, _b(b[0], caps[0][1]), _b(b[1], caps[1][1]), _b(b[2], caps[2][1]),
                                            cnt, r.ws.h, current_stream())

    bufs, c = r.run(call, caps, [0, 1, 2], shrink=(0,))
    m = col.tail.meta
    meta = an.ColMeta(t_out, reenc[0], reenc[1]) if reenc else an.ColMeta(t_out, m.a, m.a_inv)
    out = BAT(Column.void(c[0]), Column(bufs[0], c[0], meta))
    e = _errs(bufs[1:], c[1:], [True, col.tail.meta.type.encoded])
    return out, e[0], e[1]


def matchjoinAN(b1: BAT, b2: BAT, reenc_h: tuple[int, int] | None = None, reenc_t: tuple[int, int] | None = None,
                a_oid: int = A_OID_ERR):
    """matchjoinAN (matchjoinAN.tcc:46-143 merge form, :240-301 positional form):
    (BAT<H1::v2_select_t, T2::v2_select_t>, e1, e2, e3, e4)."""
    h1, t1, h2, t2 = b1.head, b1.tail, b2.head, b2.tail
    th, tt = select_type(h1.meta.type), select_type(t2.meta.type)
    reencode = reenc_h is not None
    a_h = reenc_h[0] if reencode else 0
    a_t = reenc_t[0] if reenc_t is not None else 0  # head-only re-encoding passes A_T2R = 0
    f_h = (a_h * h1.meta.a_inv) & ((1 << (8 * th.width)) - 1) if h1.meta.type.encoded else 0
    f_t = (a_t * t2.meta.a_inv) & ((1 << (8 * tt.width)) - 1) if t2.meta.type.encoded else 0
    rh = reencode and h1.meta.type.encoded
    rt = reencode and t2.meta.type.encoded
    d = [x.desc() for x in (h1, t1, h2, t2)]
    positional = t1.meta.type.name == "resoid" and h2.buf is None
    if t1.buf is None and h2.buf is None:  # <H1, void, void, T2>: seqbase alignment
        r = _Runner(3)
        cap = max(min(t1.n, h2.n), 1)
        caps = [[th.width, cap], [tt.width, cap], [8, ERR_CAP], [8, ERR_CAP]]

        def call(b, cnt):
            return r.lib.ahead_hip_matchjoin_seq_an(*[ctypes.byref(x) for x in d], int(rh), f_h, th.width, int(rt),
                                                    f_t, tt.width, a_oid, *[_b(x, c[1]) for x, c in zip(b, caps)],
                                                    cnt, r.ws.h, current_stream())

        bufs, c = r.run(call, caps, [0, 0, 1, 2], shrink=(0, 1))
        mh = an.ColMeta(th, reenc_h[0], reenc_h[1]) if reencode else an.ColMeta(th, h1.meta.a, h1.meta.a_inv)
        mt = an.ColMeta(tt, a_t, reenc_t[1] if reenc_t else 0) if reencode else an.ColMeta(tt, t2.meta.a, t2.meta.a_inv)
        e = _errs(bufs[2:], c[1:3], [h1.meta.type.encoded, t2.meta.type.encoded])
        return BAT(Column(bufs[0], c[0], mh), Column(bufs[1], c[0], mt)), e[0], None, None, e[1]
    r = _Runner(5 if not positional else 4)
    cap = max(t1.n if positional else min(t1.n, h2.n), 1)
    caps = [[th.width, cap], [tt.width, cap], [8, ERR_CAP], [8, ERR_CAP], [8, ERR_CAP], [8, ERR_CAP]]
    if positional:
        caps = caps[:5]

        def call(b, cnt):
            return r.lib.ahead_hip_matchjoin_pos_an(ctypes.byref(d[0]), ctypes.byref(d[1]), ctypes.byref(d[3]),
                                                    int(rh), f_h, th.width, int(rt), f_t, tt.width, a_oid,
                                                    *[_b(x, c[1]) for x, c in zip(b, caps)], cnt, r.ws.h,
                                                    current_stream())

        bufs, c = r.run(call, caps, [0, 0, 1, 2, 3], shrink=(0, 1))
        e = _errs(bufs[2:], c[1:], [h1.meta.type.encoded, True, t2.meta.type.encoded])
        errs = (e[0], e[1], None, e[2])
    else:
        def call(b, cnt):
            return r.lib.ahead_hip_matchjoin_merge_an(*[ctypes.byref(x) for x in d], int(rh), f_h, th.width, int(rt),
                                                      f_t, tt.width, a_oid, *[_b(x, c[1]) for x, c in zip(b, caps)],
                                                      cnt, r.ws.h, current_stream())

        bufs, c = r.run(call, caps, [0, 0, 1, 2, 3, 4], shrink=(0, 1))
        errs = tuple(_errs(bufs[2:], c[1:5], [h1.meta.type.encoded, t1.meta.type.encoded, h2.meta.type.encoded,
                                              t2.meta.type.encoded]))
    mh = an.ColMeta(th, reenc_h[0], reenc_h[1]) if reencode else an.ColMeta(th, h1.meta.a, h1.meta.a_inv)
    if reencode:
        mt = an.ColMeta(tt, a_t, reenc_t[1] if reenc_t is not None else 0)
    else:
        mt = an.ColMeta(tt, t2.meta.a, t2.meta.a_inv)
    return (BAT(Column(bufs[0], c[0], mh), Column(bufs[1], c[0], mt)),) + errs


def hashjoinAN(b1: BAT, b2: BAT, reenc_h: tuple[int, int] | None = None, reenc_t: tuple[int, int] | None = None,
               build_right: bool = True, a_oid: int = A_OID_ERR):
    """hashjoinAN with the unencoded-key hash map (hashjoinAN.tcc:54-223, entries :552-585):
    (TempBAT<..>, e1, e2, e3, e4); pairs in probe order, build rows in insertion order."""
    h1, t1, h2, t2 = b1.head, b1.tail, b2.head, b2.tail
    th = encoded_type(h1.meta.type) if reenc_h else select_type(h1.meta.type)
    tt = encoded_type(t2.meta.type) if reenc_t else select_type(t2.meta.type)
    mask_h, mask_t = (1 << (8 * encoded_type(h1.meta.type).width)) - 1, (1 << (8 * encoded_type(t2.meta.type).width)) - 1
    f_h = ((reenc_h[0] * h1.meta.a_inv) & mask_h if h1.meta.type.encoded else reenc_h[0] & mask_h) if reenc_h else 0
    f_t = ((reenc_t[0] * t2.meta.a_inv) & mask_t if t2.meta.type.encoded else reenc_t[0] & mask_t) if reenc_t else 0
    d = [x.desc() for x in (h1, t1, h2, t2)]
    r = _Runner(6)
    cap = max(t1.n, h2.n, 1)
    caps = [[th.width, cap], [tt.width, cap], [8, ERR_CAP], [8, ERR_CAP], [8, ERR_CAP], [8, ERR_CAP]]

    def call(b, cnt):
        return r.lib.ahead_hip_hashjoin_an(*[ctypes.byref(x) for x in d], int(build_right), int(bool(reenc_h)), f_h,
                                           th.width, int(bool(reenc_t)), f_t, tt.width, a_oid,
                                           *[_b(x, c[1]) for x, c in zip(b, caps)], cnt, r.ws.h, current_stream())

    bufs, c = r.run(call, caps, [0, 0, 1, 2, 3, 4], shrink=(0, 1))
    if c[5] != 0:
        raise capi.AheadError(capi.AHEAD_ERR_DEVICE, "hashjoinAN: the build table overflowed (internal limit)")
    mh = an.ColMeta(th, reenc_h[0], reenc_h[1]) if reenc_h else an.ColMeta(th, h1.meta.a, h1.meta.a_inv)
    mt = an.ColMeta(tt, reenc_t[0], reenc_t[1]) if reenc_t else an.ColMeta(tt, t2.meta.a, t2.meta.a_inv)
    errs = _errs(bufs[2:], c[1:5], [h1.meta.type.encoded, t1.meta.type.encoded, h2.meta.type.encoded,
                                    t2.meta.type.encoded])
    return (BAT(Column(bufs[0], c[0], mh), Column(bufs[1], c[0], mt)),) + tuple(errs)


def groupbyAN(bat: BAT, grouping: BAT | None = None, num_groups: int = 0, a_oid: int = A_OID):
    """groupbyAN unary (groupbyAN.tcc:51-94): (VOIDtoRGID, VGIDtoROID, e_head, e_tail);
    binary (:99-165) with grouping / num_groups: (..., e_grouping)."""
    h, t = bat.head, bat.tail
    n = t.n
    r = _Runner(6)
    meta = an.ColMeta(an.RESOID, a_oid, an.inverse(a_oid, 64))
    bits = 0
    while True:
        caps = [[8, max(n, 1)], [8, max(n, 1)], [8, ERR_CAP], [8, ERR_CAP], [8, ERR_CAP]]
        if grouping is None:
            caps = caps[:4]
            dh, dt = h.desc(), t.desc()

            def call(b, cnt):
                return r.lib.ahead_hip_groupby_an(ctypes.byref(dh), ctypes.byref(dt), a_oid, bits, b[0].ptr,
                                                  _b(b[1], caps[1][1]), _b(b[2], caps[2][1]), _b(b[3], caps[3][1]),
                                                  cnt, r.ws.h, current_stream())

            bufs, c = r.run(call, caps, [0, 1, 2, 3])
        else:
            dh, dt, dg = h.desc(), t.desc(), grouping.tail.desc()
            larger = t.meta.type.larger_bits

            def call(b, cnt):
                return r.lib.ahead_hip_groupby2_an(ctypes.byref(dh), ctypes.byref(dt), ctypes.byref(dg), num_groups,
                                                   larger, a_oid, bits, b[0].ptr, *[_b(x, cc[1]) for x, cc in
                                                                                   zip(b[1:], caps[1:])],
                                                   cnt, r.ws.h, current_stream())

            bufs, c = r.run(call, caps, [0, 1, 2, 3, 4])
        if c[5] == 0:
            break
        bits = max(bits, 22) + 2
        if bits > 28:
            raise capi.AheadError(capi.AHEAD_ERR_DEVICE, "groupbyAN: too many groups for the device dictionary")
    v2r = BAT(Column.void(c[0]), Column(bufs[0], c[0], meta))
    g2r = BAT(Column.void(c[1]), Column(bufs[1], c[1], meta))
    e = [bufs[2].to_numpy(np.uint64, c[2]) if h.meta.type.encoded and h.buf is not None else None,
         bufs[3].to_numpy(np.uint64, c[3]) if t.meta.type.encoded else None]
    if grouping is not None:
        e.append(bufs[4].to_numpy(np.uint64, c[4]))
    return (v2r, g2r) + tuple(e)


def aggregate_sum_groupedAN(bat: BAT, grouping: BAT, num_groups: int, a_result: int = an.A_RESULT,
                            a_oid: int = A_OID, result_type: an.ColType = an.RESBIGINT):
    """aggregate_sum_groupedAN<Result> (aggregateAN.tcc:41-119): (BAT<void, Result>, e_h, e_t, e_g)."""
    h, t, g = bat.head, bat.tail, grouping.tail
    if t.n != g.n:
        raise capi.AheadError(capi.AHEAD_ERR_SIZE_MISMATCH, "aggregate_sum_groupedAN: sizes differ")
    r = _Runner(4)
    caps = [[8, max(num_groups, 1)], [8, ERR_CAP], [8, ERR_CAP], [8, ERR_CAP]]
    dh, dt, dg = h.desc(), t.desc(), g.desc()

    def call(b, cnt):
        return r.lib.ahead_hip_aggregate_sum_grouped_an(ctypes.byref(dh), ctypes.byref(dt), ctypes.byref(dg),
                                                        num_groups, a_result, a_oid, b[0].ptr,
                                                        *[_b(x, cc[1]) for x, cc in zip(b[1:], caps[1:])], cnt,
                                                        r.ws.h, current_stream())

    bufs, c = r.run(call, caps, [0, 1, 2, 3])
    meta = an.ColMeta(result_type, a_result, an.inverse(a_result, 64))
    e = _errs(bufs[1:], c[1:], [h.meta.type.encoded and h.buf is not None, t.meta.type.encoded, True])
    return (BAT(Column.void(num_groups), Column(bufs[0], num_groups, meta)),) + tuple(e)


def aggregate_mul_sumAN(b1: BAT, b2: BAT, init: int = 0, ra: int = an.A_RESULT, a_oid: int = A_OID_ERR,
                        result_type: an.ColType = an.RESBIGINT):
    """aggregate_mul_sumAN<Result> (aggregateAN_scalar.tcc:58-131): (1-row BAT, e1, e2)."""
    t1, t2 = b1.tail, b2.tail
    r = _Runner(3)
    caps = [[8, 1], [8, ERR_CAP], [8, ERR_CAP]]
    d1, d2 = t1.desc(), t2.desc()

    def call(b, cnt):
        return r.lib.ahead_hip_aggregate_mul_sum_an(ctypes.byref(d1), ctypes.byref(d2), init, int(result_type.encoded),
                                                    ra, a_oid, b[0].ptr, _b(b[1], caps[1][1]), _b(b[2], caps[2][1]),
                                                    cnt, r.ws.h, current_stream())

    bufs, c = r.run(call, caps, [0, 1, 2])
    meta = an.ColMeta(result_type, ra, an.inverse(ra, 64))
    e = _errs(bufs[1:], c[1:], [t1.meta.type.encoded, t2.meta.type.encoded])
    return (BAT(Column.void(1), Column(bufs[0], 1, meta)),) + tuple(e)


def arithmeticAN(op: int, b1: BAT, b2: BAT, a_result: int = an.A_RESULT, a_oid: int = A_OID_ERR,
                 result_type: an.ColType = an.RESBIGINT):
    """arithmeticAN<Op, Result> (arithmeticAN_scalar.tcc:60-112): (BAT<void, Result>, e1..e4)."""
    h1, t1, h2, t2 = b1.head, b1.tail, b2.head, b2.tail
    if t1.n != t2.n:
        raise capi.AheadError(capi.AHEAD_ERR_SIZE_MISMATCH, "arithmeticAN: sizes differ")
    n = t1.n
    r = _Runner(5)
    caps = [[result_type.width, max(n, 1)], [8, ERR_CAP], [8, ERR_CAP], [8, ERR_CAP], [8, ERR_CAP]]
    d = [x.desc() for x in (h1, t1, h2, t2)]

    def call(b, cnt):
        return r.lib.ahead_hip_arithmetic_an(*[ctypes.byref(x) for x in d], op, a_result, result_type.width, a_oid,
                                             b[0].ptr, *[_b(x, cc[1]) for x, cc in zip(b[1:], caps[1:])], cnt,
                                             r.ws.h, current_stream())

    bufs, c = r.run(call, caps, [0, 1, 2, 3, 4])
    meta = an.ColMeta(result_type, a_result, an.inverse(a_result, 8 * result_type.width))
    e = _errs(bufs[1:], c[1:], [h1.meta.type.encoded and h1.buf is not None, t1.meta.type.encoded,
                                h2.meta.type.encoded and h2.buf is not None, t2.meta.type.encoded])
    return (BAT(Column.void(n), Column(bufs[0], n, meta)),) + tuple(e)


def _decoded(col: Column, call, a_oid: int) -> tuple:
    """Runs a decode entry point over a column into device memory, then widens the values back to
    the code width (an.unchecked_type, A = 1).  Returns (column, error positions)."""
    lib, ws = capi.load(), capi.workspace()
    wd = col.meta.type.unenc_bits // 8
    dec = DeviceBuffer(wd * max(col.n, 1))
    cnt = DeviceBuffer(16)
    d = col.desc()
    err_cap = 4096
    while True:
        err = DeviceBuffer(8 * err_cap)
        capi.check(call(lib, d, a_oid, dec, err, err_cap, cnt, ws))
        capi.synchronize()
        n_dec, n_err = capi._counts(cnt, 2)
        if n_err <= err_cap:
            break
        err_cap = n_err
    ut = an.unchecked_type(col.meta.type)
    out = DeviceBuffer(ut.width * max(n_dec, 1))
    plain = {1: an.TINYINT, 2: an.SHORTINT, 4: an.INT, 8: an.BIGINT}[wd]
    if n_dec:
        capi.encode_an_into(Column(dec, n_dec, an.plain_meta(plain)), 1, ut.width, out)
    return Column(out, n_dec, an.ColMeta(ut, 1, 1)), err.to_numpy(np.uint64, n_err)


def checkAndDecodeAN(bat: BAT, a_oid: int = A_OID_ERR):
    """checkAndDecodeAN of a BAT's tail (OperatorsAN.hpp:157-161; encdecAN_scalar.tcc:185-210):
    corrupt rows are dropped and reported (i * a_oid).  The decoded values come back at the code
    width with A = 1 and no check, the form the other operators take.  Returns (BAT<void, values>,
    errors)."""
    def call(lib, d, a, dec, err, cap, cnt, ws):
        return lib.ahead_hip_check_and_decode_an(ctypes.byref(d), a, dec.ptr, d.n, err.ptr, cap, cnt.ptr, ws.h,
                                                 capi.current_stream())

    c, e = _decoded(bat.tail, call, a_oid)
    return BAT(Column.void(c.n), c), e


def decodeAN(bat: BAT, a_oid: int = A_OID_ERR):
    """decodeAN of a BAT's tail (encdecAN_scalar.tcc:96-126): every row decoded (corrupt ones
    too, and reported), values at the code width with A = 1 and no check."""
    def call(lib, d, a, dec, err, cap, cnt, ws):
        return lib.ahead_hip_decode_an(ctypes.byref(d), a, dec.ptr, err.ptr, cap, cnt.ptr, ws.h, capi.current_stream())

    c, e = _decoded(bat.tail, call, a_oid)
    return BAT(Column.void(c.n), c), e


def copy(bat: BAT) -> BAT:
    """A device copy of a <void, T> BAT (the reference's copy(), miscellaneous.tcc): a replica."""
    t = bat.tail
    buf = t.buf.prefix_copy(t.buf.nbytes) if t.buf is not None else None
    return BAT(bat.head, Column(buf, t.n, t.meta))


def _star_query_obj(n_fact: int, dims: list, preds: list, meas: list, kind: int, pre_mask: int = 0,
                    a_result: int = an.A_RESULT, a_oid: int = A_OID_ERR, ev_scan: tuple | None = None):
    cdims = []
    for d in dims:
        x = dict(d)
        x["key_index"] = key_index(d["key"])
        sel = d.get("sel")
        if sel is not None:
            x["sel"] = sel.buf if sel.buf is not None else DeviceBuffer(8)
            x["sel_count"] = DeviceBuffer.from_numpy(np.array([sel.n], dtype=np.uint64))
            x["sel_a_inv"] = sel.meta.a_inv
        cdims.append(x)
    return capi.StarQuery(n_fact, cdims, preds, meas, kind, pre_mask, a_result, a_oid, ev_scan=ev_scan)


def star_query(n_fact: int, dims: list, preds: list, meas: list, kind: int, pre_mask: int = 0,
               a_result: int = an.A_RESULT, a_oid: int = A_OID_ERR, ev_scan: tuple | None = None) -> dict:
    """The fused star query (ahead_hip_star_query) over device columns.  dims: dicts in the
    chain's order with fk, key, attr (Column or None), sel (Column of encoded OIDs or None), flags;
    preds: (Column, op1, th1, op2, th2, comb) with encoded thresholds; meas: Columns.  Returns the
    entry point's outputs as numpy (status, groups, kept, sums, first_row, keys by dimension, the
    16 error vectors by kind)."""
    q = _star_query_obj(n_fact, dims, preds, meas, kind, pre_mask, a_result, a_oid, ev_scan)
    q.launch()
    return q.results()


def star_query_batch(calls: list, ev_scan: tuple | None = None) -> list:
    """Several fused star queries over one fact table with shared fact scans
    (ahead_hip_star_query_batch): `calls` holds star_query's keyword arguments per query; returns
    star_query's result per query.  ev_scan brackets all of the batch's scans."""
    qs = [_star_query_obj(**c) for c in calls]
    if ev_scan is not None and qs:
        qs[0].args.ev_scan[0], qs[0].args.ev_scan[1] = ev_scan
    capi.star_query_batch(qs)
    return [q.results() for q in qs]


def star_query_ws_bytes(n_fact: int, dims: list, meas: list, kind: int) -> int:
    """Device workspace bytes of a fused star query over these columns
    (ahead_hip_star_query_ws_bytes): it depends on the dimensions' row counts and on which of them
    have a key index, not on the selections or predicates."""
    cdims = [dict(d, key_index=key_index(d["key"]), sel=None) for d in dims]
    return capi.StarQuery(n_fact, cdims, [], meas, kind).ws_bytes()


def key_index(key: Column) -> "capi.StarKeyIndex":
    """The star-query key index of a resident dimension key column, built at its first use and kept
    with the column (the column's codes must not change afterwards)."""
    ix = getattr(key, "_star_key_index", None)
    if ix is None:
        ix = capi.StarKeyIndex(key)
        key._star_key_index = ix
    return ix
