"""ctypes binding of the CPU oracle (oracle/an_oracle.c) for the tests.

The oracle is the checker: tests compare the HIP operators against it, and pin the oracle itself
against closed forms and textbook results computed independently with numpy.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

from ahead_amd import an

ROOT = Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"
ORACLE_SO = ORACLE_DIR / "_build" / "liboracle.so"
SIMD_SO = ORACLE_DIR / "_build" / "liboracle_simd.so"


class OrCol(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("n", ctypes.c_uint64),
        ("width", ctypes.c_int),
        ("encoded", ctypes.c_int),
        ("unenc_bits", ctypes.c_int),
        ("a", ctypes.c_uint64),
        ("a_inv", ctypes.c_uint64),
        ("max_u", ctypes.c_uint64),
        ("seqbase", ctypes.c_uint64),
    ]


class OrVec(ctypes.Structure):
    _fields_ = [("v", ctypes.POINTER(ctypes.c_uint64)), ("n", ctypes.c_uint64), ("cap", ctypes.c_uint64)]

    def numpy(self) -> np.ndarray:
        if self.n == 0:
            return np.zeros(0, dtype=np.uint64)
        return np.ctypeslib.as_array(self.v, shape=(self.n,)).copy()


class OrSqDim(ctypes.Structure):
    _fields_ = [("fk", OrCol), ("key", OrCol), ("attr", OrCol), ("sel", ctypes.c_void_p), ("n_sel", ctypes.c_uint64),
                ("sel_a_inv", ctypes.c_uint64), ("flags", ctypes.c_int)]


class OrSqPred(ctypes.Structure):
    _fields_ = [("col", OrCol), ("op1", ctypes.c_int), ("op2", ctypes.c_int), ("comb", ctypes.c_int),
                ("th1", ctypes.c_uint64), ("th2", ctypes.c_uint64)]


class OrSqArgs(ctypes.Structure):
    _fields_ = [("n_fact", ctypes.c_uint64), ("n_dims", ctypes.c_int), ("n_preds", ctypes.c_int),
                ("n_meas", ctypes.c_int), ("meas_kind", ctypes.c_int), ("meas_pre_mask", ctypes.c_int),
                ("dims", OrSqDim * 4), ("preds", OrSqPred * 2), ("meas", OrCol * 2), ("a_result", ctypes.c_uint64),
                ("a_oid_err", ctypes.c_uint64)]


class OrSqOut(ctypes.Structure):
    _fields_ = [("sums", OrVec), ("first_row", OrVec), ("keys", OrVec * 4), ("err", OrVec * 16),
                ("kept", ctypes.c_uint64), ("status", ctypes.c_int)]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not ORACLE_SO.exists():
            build()
        L = ctypes.CDLL(str(ORACLE_SO))
        P = ctypes.POINTER
        c = ctypes.POINTER(OrCol)
        v = ctypes.POINTER(OrVec)
        u64, i = ctypes.c_uint64, ctypes.c_int
        sig = {
            "or_inverse": ([u64, i], u64),
            "or_select_an": ([c, i, u64, i, u64, i, u64, u64, v, v], u64),
            "or_check_an": ([c, u64, v], None),
            "or_decode_an": ([c, u64, v, v], None),
            "or_check_and_decode_an": ([c, u64, v, v], None),
            "or_check_and_decode_an2": ([c, c, u64, v, v, v, v], None),
            "or_encode_an": ([c, u64, i, v], None),
            "or_fetchjoin_an": ([c, c, i, u64, u64, v, v, v], None),
            "or_matchjoin_pos_an": ([c, c, c, i, u64, i, i, u64, i, u64, v, v, v, v, v], None),
            "or_matchjoin_merge_an": ([c, c, c, c, i, u64, i, i, u64, i, u64, v, v, v, v, v, v], None),
            "or_matchjoin_seq_an": ([c, c, c, c, i, u64, i, i, u64, i, u64, v, v, v, v], None),
            "or_hashjoin_an": ([c, c, c, c, i, i, u64, i, i, u64, i, u64, v, v, v, v, v, v], None),
            "or_groupby_an": ([c, c, u64, v, v, v, v], None),
            "or_groupby2_an": ([c, c, c, u64, i, u64, v, v, v, v, v], i),
            "or_aggregate_sum_grouped_an": ([c, c, c, u64, u64, u64, P(u64), v, v, v], i),
            "or_aggregate_mul_sum_an": ([c, c, u64, i, u64, u64, v, v], u64),
            "or_arithmetic_an": ([c, c, c, c, i, u64, i, u64, v, v, v, v, v], i),
            "or_star1_pipeline": ([c, c, c, c, c, u64, u64, u64, u64] + [v] * 10, ctypes.c_int64),
            "or_vec_free": ([v], None),
            "or_star_query": ([P(OrSqArgs), P(OrSqOut)], i),
            "or_sq_out_free": ([P(OrSqOut)], None),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


class Col:
    """Keeps the numpy array alive next to its OrCol descriptor."""

    def __init__(self, arr: np.ndarray | None, meta: an.ColMeta, n: int | None = None):
        self.arr = None if arr is None else np.ascontiguousarray(arr)
        self.meta = meta
        t = meta.type
        if self.arr is not None:
            assert self.arr.dtype.itemsize == t.width, (self.arr.dtype, t)
        nn = self.arr.size if self.arr is not None else int(n)
        self.c = OrCol(
            self.arr.ctypes.data if self.arr is not None else None,
            nn, t.width, int(t.encoded), t.unenc_bits, meta.a & an.MASK64,
            meta.a_inv & an.MASK64, t.max_u & an.MASK64, meta.seqbase,
        )

    @property
    def ref(self):
        return ctypes.byref(self.c)


def void(n: int, seqbase: int = 0) -> Col:
    return Col(None, an.plain_meta(an.VOID, seqbase), n)


class Vecs:
    """A bundle of output vectors that frees itself."""

    def __init__(self, k: int):
        self.v = [OrVec() for _ in range(k)]

    def refs(self):
        return [ctypes.byref(x) for x in self.v]

    def numpy(self):
        out = [x.numpy() for x in self.v]
        for x in self.v:
            lib().or_vec_free(ctypes.byref(x))
        return out


def inverse(a: int, bits: int) -> int:
    return int(lib().or_inverse(a, bits))


def select_an(col: Col, op1, th1, op2=0, th2=0, comb=0, a_head=an.A_OID, a_oid=an.A_OID_ERR):
    vs = Vecs(2)
    s = lib().or_select_an(col.ref, op1, th1 & an.MASK64, op2, th2 & an.MASK64, comb, a_head, a_oid, *vs.refs())
    oids, err = vs.numpy()
    return oids, err, int(s)


def check_an(col: Col, a_oid=an.A_OID_ERR):
    vs = Vecs(1)
    lib().or_check_an(col.ref, a_oid, *vs.refs())
    return vs.numpy()[0]


def decode_an(col: Col, a_oid=an.A_OID_ERR):
    vs = Vecs(2)
    lib().or_decode_an(col.ref, a_oid, *vs.refs())
    return tuple(vs.numpy())


def check_and_decode_an(col: Col, a_oid=an.A_OID_ERR):
    vs = Vecs(2)
    lib().or_check_and_decode_an(col.ref, a_oid, *vs.refs())
    return tuple(vs.numpy())


def check_and_decode_an2(head: Col, tail: Col, a_oid=an.A_OID_ERR):
    vs = Vecs(4)
    lib().or_check_and_decode_an2(head.ref, tail.ref, a_oid, *vs.refs())
    return tuple(vs.numpy())


def encode_an(col: Col, a: int, out_width: int):
    vs = Vecs(1)
    lib().or_encode_an(col.ref, a, out_width, *vs.refs())
    return vs.numpy()[0]


def fetchjoin_an(oids: Col, col: Col, reenc=0, a_reenc=0, a_oid=an.A_OID_ERR):
    vs = Vecs(3)
    lib().or_fetchjoin_an(oids.ref, col.ref, reenc, a_reenc, a_oid, *vs.refs())
    return tuple(vs.numpy())


def matchjoin_pos_an(h1: Col, t1: Col, t2: Col, reenc_h=0, f_h1=0, wout_h=8, reenc_t=0, f_t2=0, wout_t=8,
                     a_oid=an.A_OID_ERR):
    vs = Vecs(5)
    lib().or_matchjoin_pos_an(h1.ref, t1.ref, t2.ref, reenc_h, f_h1, wout_h, reenc_t, f_t2, wout_t, a_oid,
                              *vs.refs())
    return tuple(vs.numpy())


def matchjoin_merge_an(h1: Col, t1: Col, h2: Col, t2: Col, reenc_h=0, f_h1=0, wout_h=8, reenc_t=0, f_t2=0,
                       wout_t=8, a_oid=an.A_OID_ERR):
    vs = Vecs(6)
    lib().or_matchjoin_merge_an(h1.ref, t1.ref, h2.ref, t2.ref, reenc_h, f_h1, wout_h, reenc_t, f_t2, wout_t, a_oid,
                                *vs.refs())
    return tuple(vs.numpy())


def matchjoin_seq_an(h1: Col, t1: Col, h2: Col, t2: Col, reenc_h=0, f_h1=0, wout_h=8, reenc_t=0, f_t2=0,
                     wout_t=8, a_oid=an.A_OID_ERR):
    vs = Vecs(4)
    lib().or_matchjoin_seq_an(h1.ref, t1.ref, h2.ref, t2.ref, reenc_h, f_h1, wout_h, reenc_t, f_t2, wout_t, a_oid,
                              *vs.refs())
    return tuple(vs.numpy())


def hashjoin_an(h1: Col, t1: Col, h2: Col, t2: Col, build_right=1, reenc_h=0, f_h1=0, wout_h=8, reenc_t=0,
                f_t2=0, wout_t=8, a_oid=an.A_OID_ERR):
    vs = Vecs(6)
    lib().or_hashjoin_an(h1.ref, t1.ref, h2.ref, t2.ref, build_right, reenc_h, f_h1, wout_h, reenc_t, f_t2,
                         wout_t, a_oid, *vs.refs())
    return tuple(vs.numpy())


def groupby_an(head: Col, tail: Col, a_oid=an.A_OID):
    vs = Vecs(4)
    lib().or_groupby_an(head.ref, tail.ref, a_oid, *vs.refs())
    return tuple(vs.numpy())


def groupby2_an(head: Col, tail: Col, grouping: Col, num_groups: int, larger_bits: int, a_oid=an.A_OID):
    vs = Vecs(5)
    r = lib().or_groupby2_an(head.ref, tail.ref, grouping.ref, num_groups, larger_bits, a_oid, *vs.refs())
    out = vs.numpy()
    if r != 0:
        raise RuntimeError("size mismatch")
    return tuple(out)


def aggregate_sum_grouped_an(head: Col, tail: Col, grouping: Col, num_groups: int, a_result=an.A_RESULT,
                             a_oid=an.A_OID):
    vs = Vecs(3)
    sums = np.zeros(max(num_groups, 1), dtype=np.uint64)
    r = lib().or_aggregate_sum_grouped_an(head.ref, tail.ref, grouping.ref, num_groups, a_result, a_oid,
                                          sums.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), *vs.refs())
    out = vs.numpy()
    if r != 0:
        raise RuntimeError("size mismatch")
    return (sums[:num_groups], *out)


def aggregate_mul_sum_an(t1: Col, t2: Col, init=0, result_encoded=1, ra=an.A_RESULT, a_oid=an.A_OID_ERR):
    vs = Vecs(2)
    total = lib().or_aggregate_mul_sum_an(t1.ref, t2.ref, init, result_encoded, ra, a_oid, *vs.refs())
    return (int(total), *vs.numpy())


def arithmetic_an(h1: Col, t1: Col, h2: Col, t2: Col, op: int, a_result: int, result_width: int,
                  a_oid=an.A_OID_ERR):
    vs = Vecs(5)
    r = lib().or_arithmetic_an(h1.ref, t1.ref, h2.ref, t2.ref, op, a_result, result_width, a_oid, *vs.refs())
    out = vs.numpy()
    if r != 0:
        raise RuntimeError("size mismatch")
    return tuple(out)


def star1_pipeline(fact_fk: Col, fact_m: Col, dim_key: Col, dim_a1: Col, dim_a2: Col,
                   a_oid_join=an.A_OID, a_oid_err=an.A_OID_ERR, a_oid_grp=an.A_OID, a_result=an.A_RESULT):
    vs = Vecs(10)
    ng = lib().or_star1_pipeline(fact_fk.ref, fact_m.ref, dim_key.ref, dim_a1.ref, dim_a2.ref, a_oid_join,
                                 a_oid_err, a_oid_grp, a_result, *vs.refs())
    out = vs.numpy()
    if ng < 0:
        raise RuntimeError("pipeline failed")
    names = ["sums", "groid", "key1", "key2", "e_fk", "e_dk", "e_a1", "e_a2", "e_m", "e_grp"]
    return dict(zip(names, out))


SQ_SUM, SQ_SUB, SQ_MULSUM = 0, 1, 2
SQ_PROBE, SQ_REFETCH = 1, 2


_simd = None


def simd_lib():
    """The all-core vectorised restatement (oracle/an_simd.cpp), the CPU baseline."""
    global _simd
    if _simd is None:
        if not SIMD_SO.exists():
            build()
        L = ctypes.CDLL(str(SIMD_SO))
        P = ctypes.POINTER
        c = P(OrCol)
        v = P(OrVec)
        u64, i = ctypes.c_uint64, ctypes.c_int
        sig = {"or_simd_level": ([], i),
               "or_select_an_simd": ([c, i, u64, i, u64, i, u64, u64, v, v, i], u64),
               "or_star_query_simd": ([P(OrSqArgs), P(OrSqOut), i], i)}
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _simd = L
    return _simd


def simd_level() -> int:
    return int(simd_lib().or_simd_level())


def select_an_simd(col: Col, op1, th1, op2=0, th2=0, comb=0, a_head=an.A_OID, a_oid=an.A_OID_ERR, threads=0):
    vs = Vecs(2)
    s = simd_lib().or_select_an_simd(col.ref, op1, th1, op2, th2, comb, a_head, a_oid, *vs.refs(), threads)
    out = vs.numpy()
    return out[0], out[1], int(s)


def star_query(n_fact: int, dims: list, preds: list, meas: list, kind: int, pre_mask: int = 0,
               a_result=an.A_RESULT, a_oid_err=an.A_OID_ERR, simd_threads: int | None = None) -> dict:
    """or_star_query (or, with simd_threads set, or_star_query_simd on that many threads, 0: all).
    dims: dicts with fk, key (Col), attr (Col or None), sel (uint64 numpy array of encoded OIDs or
    None), sel_a_inv, flags; preds: (Col, op1, th1, op2, th2, comb); meas: Cols."""
    a = OrSqArgs()
    a.n_fact = n_fact
    a.n_dims = len(dims)
    a.n_preds = len(preds)
    a.n_meas = len(meas)
    a.meas_kind = kind
    a.meas_pre_mask = pre_mask
    keep = []
    for i, d in enumerate(dims):
        x = a.dims[i]
        x.fk = d["fk"].c
        x.key = d["key"].c
        if d.get("attr") is not None:
            x.attr = d["attr"].c
        sel = d.get("sel")
        if sel is not None:
            sel = np.ascontiguousarray(sel, dtype=np.uint64)
            keep.append(sel)
            x.sel = sel.ctypes.data
            x.n_sel = sel.size
        x.sel_a_inv = d.get("sel_a_inv", an.A_OID_INV) & an.MASK64
        x.flags = d.get("flags", 0)
    for i, (col, op1, th1, op2, th2, comb) in enumerate(preds):
        a.preds[i] = OrSqPred(col.c, op1, op2, comb, th1 & an.MASK64, th2 & an.MASK64)
    for i, m in enumerate(meas):
        a.meas[i] = m.c
    a.a_result = a_result
    a.a_oid_err = a_oid_err
    o = OrSqOut()
    if simd_threads is None:
        st = lib().or_star_query(ctypes.byref(a), ctypes.byref(o))
    else:
        st = simd_lib().or_star_query_simd(ctypes.byref(a), ctypes.byref(o), simd_threads)
    res = {"status": st, "kept": int(o.kept), "sums": o.sums.numpy(), "first_row": o.first_row.numpy(),
           "keys": {i: o.keys[i].numpy() for i, d in enumerate(dims) if d.get("attr") is not None},
           "err": [o.err[k].numpy() for k in range(16)]}
    res["groups"] = res["sums"].size
    lib().or_sq_out_free(ctypes.byref(o))
    return res
