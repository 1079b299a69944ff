"""Host BATs over numpy arrays with the operator API of ahead_amd.bat, computed by the CPU oracle.

TEST INFRASTRUCTURE: lets the star-schema pipelines of ahead_amd.ssb run on the oracle (the
checker), so a test can compare every intermediate error vector and the final results of a whole
operator chain with the GPU's.  Also the bench's CPU baseline of those pipelines.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

import oracle_lib as O
from ahead_amd import an

LT, LE, GT, GE, EQ, NE = 0, 1, 2, 3, 4, 5
NONE_, AND, OR = 0, 1, 2
ADD, SUB, MUL, DIV = 0, 1, 2, 3
A_OID_ERR = an.A_OID_ERR
A_OID = an.A_OID
RESOID_META = an.an_meta(an.RESOID, an.A_OID)
_ENCODED = {"tinyint": an.RESTINY, "shortint": an.RESSHORT, "int": an.RESINT, "bigint": an.RESBIGINT,
            "oid": an.RESOID, "void": an.RESOID}


def encoded_type(t):
    return t if t.encoded else _ENCODED[t.name]


def select_type(t):
    return an.OID if t.name == "void" else t


@dataclass
class HCol:
    arr: np.ndarray | None
    n: int
    meta: an.ColMeta

    @property
    def buf(self):
        return self.arr

    def oc(self) -> O.Col:
        return O.Col(self.arr, self.meta, self.n)

    def to_numpy(self) -> np.ndarray:
        return self.arr


def void(n: int, seqbase: int = 0) -> HCol:
    return HCol(None, n, an.plain_meta(an.VOID, seqbase))


def _col(arr: np.ndarray, t: an.ColType, a: int, a_inv: int) -> HCol:
    arr = np.ascontiguousarray(arr.astype(np.dtype(f"u{t.width}")))
    return HCol(arr, arr.size, an.ColMeta(t, a, a_inv))


@dataclass
class BAT:
    head: HCol
    tail: HCol | None

    def size(self) -> int:
        return self.head.n if self.head.arr is not None or self.tail is None else self.tail.n

    def mirror_head(self):
        return BAT(self.head, self.head)

    def reverse(self):
        return BAT(self.tail, self.head)

    def clear_head(self):
        return BAT(void(self.tail.n, self.head.meta.seqbase if self.head.arr is None else 0), self.tail)


def column(values: np.ndarray, meta: an.ColMeta) -> BAT:
    return BAT(void(values.size), HCol(np.ascontiguousarray(values), values.size, meta))


def _e(v, encoded):
    return v if encoded else None


def selectAN(bat, op1, th1, op2=LT, th2=0, comb=NONE_, a_oid=A_OID_ERR):
    oids, err, _ = O.select_an(bat.tail.oc(), op1, th1, op2, th2, comb, a_oid=a_oid)
    return BAT(_col(oids, an.RESOID, an.A_OID, an.A_OID_INV), None), err


def fetchjoinAN(oids, col, reenc=None, a_oid=A_OID_ERR):
    t_out = encoded_type(col.tail.meta.type)
    out, e1, e2 = O.fetchjoin_an(oids.tail.oc(), col.tail.oc(), 1 if reenc else 0, reenc[0] if reenc else 0, a_oid)
    m = col.tail.meta
    c = _col(out, t_out, reenc[0], reenc[1]) if reenc else _col(out, t_out, m.a, m.a_inv)
    return BAT(void(c.n), c), e1, _e(e2, m.type.encoded)


def matchjoinAN(b1, b2, reenc_h=None, reenc_t=None, a_oid=A_OID_ERR):
    h1, t1, h2, t2 = b1.head, b1.tail, b2.head, b2.tail
    th, tt = select_type(h1.meta.type), select_type(t2.meta.type)
    reencode = reenc_h is not None
    a_h = reenc_h[0] if reencode else 0
    a_t = reenc_t[0] if reenc_t is not None else 0
    f_h = (a_h * h1.meta.a_inv) & ((1 << (8 * th.width)) - 1) if h1.meta.type.encoded else 0
    f_t = (a_t * t2.meta.a_inv) & ((1 << (8 * tt.width)) - 1) if t2.meta.type.encoded else 0
    rh, rt = int(reencode and h1.meta.type.encoded), int(reencode and t2.meta.type.encoded)
    positional = t1.meta.type.name == "resoid" and h2.arr is None
    if positional:
        oh, ot, e1, e2, e4 = O.matchjoin_pos_an(h1.oc(), t1.oc(), t2.oc(), rh, f_h, th.width, rt, f_t, tt.width, a_oid)
        errs = (_e(e1, h1.meta.type.encoded), e2, None, _e(e4, t2.meta.type.encoded))
    else:
        oh, ot, e1, e2, e3, e4 = O.matchjoin_merge_an(h1.oc(), t1.oc(), h2.oc(), t2.oc(), rh, f_h, th.width, rt, f_t,
                                                      tt.width, a_oid)
        errs = (_e(e1, h1.meta.type.encoded), _e(e2, t1.meta.type.encoded), _e(e3, h2.meta.type.encoded),
                _e(e4, t2.meta.type.encoded))
    mh = (reenc_h[0], reenc_h[1]) if reencode else (h1.meta.a, h1.meta.a_inv)
    mt = (a_t, reenc_t[1] if reenc_t is not None else 0) if reencode else (t2.meta.a, t2.meta.a_inv)
    return (BAT(_col(oh, th, *mh), _col(ot, tt, *mt)),) + errs


def hashjoinAN(b1, b2, reenc_h=None, reenc_t=None, build_right=True, a_oid=A_OID_ERR):
    h1, t1, h2, t2 = b1.head, b1.tail, b2.head, b2.tail
    th = encoded_type(h1.meta.type) if reenc_h else select_type(h1.meta.type)
    tt = encoded_type(t2.meta.type) if reenc_t else select_type(t2.meta.type)
    mask_h = (1 << (8 * encoded_type(h1.meta.type).width)) - 1
    mask_t = (1 << (8 * encoded_type(t2.meta.type).width)) - 1
    f_h = ((reenc_h[0] * h1.meta.a_inv) & mask_h if h1.meta.type.encoded else reenc_h[0] & mask_h) if reenc_h else 0
    f_t = ((reenc_t[0] * t2.meta.a_inv) & mask_t if t2.meta.type.encoded else reenc_t[0] & mask_t) if reenc_t else 0
    oh, ot, e1, e2, e3, e4 = O.hashjoin_an(h1.oc(), t1.oc(), h2.oc(), t2.oc(), int(build_right), int(bool(reenc_h)),
                                           f_h, th.width, int(bool(reenc_t)), f_t, tt.width, a_oid)
    mh = (reenc_h[0], reenc_h[1]) if reenc_h else (h1.meta.a, h1.meta.a_inv)
    mt = (reenc_t[0], reenc_t[1]) if reenc_t else (t2.meta.a, t2.meta.a_inv)
    return (BAT(_col(oh, th, *mh), _col(ot, tt, *mt)), _e(e1, h1.meta.type.encoded), _e(e2, t1.meta.type.encoded),
            _e(e3, h2.meta.type.encoded), _e(e4, t2.meta.type.encoded))


def groupbyAN(bat, grouping=None, num_groups=0, a_oid=A_OID):
    h, t = bat.head, bat.tail
    inv = an.inverse(a_oid, 64)
    if grouping is None:
        v2r, g2r, e1, e2 = O.groupby_an(h.oc(), t.oc(), a_oid)
        extra = ()
    else:
        v2r, g2r, e1, e2, e3 = O.groupby2_an(h.oc(), t.oc(), grouping.tail.oc(), num_groups, t.meta.type.larger_bits,
                                              a_oid)
        extra = (e3,)
    a = _col(v2r, an.RESOID, a_oid, inv)
    b = _col(g2r, an.RESOID, a_oid, inv)
    return (BAT(void(a.n), a), BAT(void(b.n), b), _e(e1, h.meta.type.encoded and h.arr is not None),
            _e(e2, t.meta.type.encoded)) + extra


def aggregate_sum_groupedAN(bat, grouping, num_groups, a_result=an.A_RESULT, a_oid=A_OID, result_type=an.RESBIGINT):
    sums, e1, e2, e3 = O.aggregate_sum_grouped_an(bat.head.oc(), bat.tail.oc(), grouping.tail.oc(), num_groups,
                                                  a_result, a_oid)
    c = _col(sums, result_type, a_result, an.inverse(a_result, 64))
    return (BAT(void(num_groups), c), _e(e1, bat.head.meta.type.encoded and bat.head.arr is not None),
            _e(e2, bat.tail.meta.type.encoded), e3)


def aggregate_mul_sumAN(b1, b2, init=0, ra=an.A_RESULT, a_oid=A_OID_ERR, result_type=an.RESBIGINT):
    total, e1, e2 = O.aggregate_mul_sum_an(b1.tail.oc(), b2.tail.oc(), init, int(result_type.encoded), ra, a_oid)
    c = _col(np.array([total], dtype=np.uint64), result_type, ra, an.inverse(ra, 64))
    return BAT(void(1), c), _e(e1, b1.tail.meta.type.encoded), _e(e2, b2.tail.meta.type.encoded)


def arithmeticAN(op, b1, b2, a_result=an.A_RESULT, a_oid=A_OID_ERR, result_type=an.RESBIGINT):
    out, e1, e2, e3, e4 = O.arithmetic_an(b1.head.oc(), b1.tail.oc(), b2.head.oc(), b2.tail.oc(), op, a_result,
                                          result_type.width, a_oid)
    c = _col(out, result_type, a_result, an.inverse(a_result, 8 * result_type.width))
    return (BAT(void(c.n), c), _e(e1, b1.head.meta.type.encoded and b1.head.arr is not None),
            _e(e2, b1.tail.meta.type.encoded), _e(e3, b2.head.meta.type.encoded and b2.head.arr is not None),
            _e(e4, b2.tail.meta.type.encoded))


SIMD_THREADS = None  # set to a thread count (0: all) to run star queries on the SIMD restatement


def _decoded(bat, fn, a_oid):
    dec, err = fn(bat.tail.oc(), a_oid)
    ut = an.unchecked_type(bat.tail.meta.type)
    c = _col(np.asarray(dec), ut, 1, 1)
    return BAT(void(c.n), c), err


def checkAndDecodeAN(bat, a_oid=A_OID_ERR):
    """bat.checkAndDecodeAN on the oracle: corrupt rows dropped, values at the code width."""
    return _decoded(bat, O.check_and_decode_an, a_oid)


def decodeAN(bat, a_oid=A_OID_ERR):
    return _decoded(bat, O.decode_an, a_oid)


def copy(bat):
    t = bat.tail
    return BAT(bat.head, HCol(None if t.arr is None else t.arr.copy(), t.n, t.meta))


def star_query(n_fact, dims, preds, meas, kind, pre_mask=0, a_result=an.A_RESULT, a_oid=A_OID_ERR, ev_scan=None):
    """The oracle's fused star query (or_star_query, or or_star_query_simd when SIMD_THREADS is set)
    with the argument form of bat.star_query."""
    odims = []
    for d in dims:
        sel = d.get("sel")
        odims.append({"fk": d["fk"].oc(), "key": d["key"].oc(), "attr": d["attr"].oc() if d.get("attr") else None,
                      "sel": sel.arr if sel is not None else None,
                      "sel_a_inv": sel.meta.a_inv if sel is not None else an.A_OID_INV, "flags": d.get("flags", 0)})
    opreds = [(c.oc(), op1, th1, op2, th2, comb) for c, op1, th1, op2, th2, comb in preds]
    return O.star_query(n_fact, odims, opreds, [m.oc() for m in meas], kind, pre_mask, a_result, a_oid,
                        simd_threads=SIMD_THREADS)


def tables(spec) -> dict:
    """The star schema of ahead_amd.ssb as host BATs (same streams as the device tables)."""
    from ahead_amd import ssb

    return {name: column(codes, meta) for name, (codes, meta, _) in ssb.host_tables(spec).items()}
