"""ctypes binding of libahead_hip.so (include/ahead_hip.h) for tests, smoke and bench.

PyTorch only supplies device memory and the stream: every operator runs in the HIP library.
If the library is missing, loading raises instead of falling back to anything.
"""
from __future__ import annotations

import ctypes
import threading
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import an

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libahead_hip.so"

AHEAD_OK = 0
(AHEAD_ERR_INVALID, AHEAD_ERR_SIZE_MISMATCH, AHEAD_ERR_HIP, AHEAD_ERR_NOMEM, AHEAD_ERR_TOO_LARGE,
 AHEAD_ERR_DEVICE) = range(1, 7)
STATUS_NAMES = {0: "OK", 1: "INVALID", 2: "SIZE_MISMATCH", 3: "HIP", 4: "NOMEM", 5: "TOO_LARGE", 6: "DEVICE"}

LT, LE, GT, GE, EQ, NE = range(6)
COMB_NONE, COMB_AND, COMB_OR = range(3)


class AheadError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"ahead_hip error {STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class AheadCol(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("n", ctypes.c_uint64),
        ("width", ctypes.c_uint32),
        ("encoded", ctypes.c_uint32),
        ("unenc_bits", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("a", ctypes.c_uint64),
        ("a_inv", ctypes.c_uint64),
        ("max_u", ctypes.c_uint64),
        ("seqbase", ctypes.c_uint64),
    ]


class Buf(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("cap", ctypes.c_uint64)]


class SelectJob(ctypes.Structure):
    _fields_ = [("col", AheadCol), ("op1", ctypes.c_int32), ("op2", ctypes.c_int32), ("comb", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("th1", ctypes.c_uint64), ("th2", ctypes.c_uint64),
                ("a_head", ctypes.c_uint64), ("a_oid", ctypes.c_uint64), ("oids", ctypes.c_void_p),
                ("oids_cap", ctypes.c_uint64), ("err", ctypes.c_void_p), ("err_cap", ctypes.c_uint64),
                ("d_counts", ctypes.c_void_p), ("d_sum", ctypes.c_void_p)]


class Star1Args(ctypes.Structure):
    _fields_ = [("fact_fk", AheadCol), ("fact_m", AheadCol), ("dim_key", AheadCol), ("dim_a1", AheadCol),
                ("dim_a2", AheadCol), ("a_oid_join", ctypes.c_uint64), ("a_oid_err", ctypes.c_uint64),
                ("a_oid_grp", ctypes.c_uint64), ("a_result", ctypes.c_uint64)]


class Star1Out(ctypes.Structure):
    _fields_ = [("sums", ctypes.c_void_p), ("groid", ctypes.c_void_p), ("key1", ctypes.c_void_p),
                ("key2", ctypes.c_void_p), ("groups_cap", ctypes.c_uint64),
                ("e_fk", ctypes.c_void_p), ("e_dk", ctypes.c_void_p), ("e_a1", ctypes.c_void_p),
                ("e_a2", ctypes.c_void_p), ("e_m", ctypes.c_void_p),
                ("e_fk_cap", ctypes.c_uint64), ("e_dk_cap", ctypes.c_uint64), ("e_a1_cap", ctypes.c_uint64),
                ("e_a2_cap", ctypes.c_uint64), ("e_m_cap", ctypes.c_uint64), ("d_counts", ctypes.c_void_p)]


SQ_MAX_DIMS, SQ_MAX_PREDS, SQ_MAX_MEAS, SQ_KINDS = 4, 2, 2, 16
SQ_SUM, SQ_SUB, SQ_MULSUM = range(3)
SQ_PROBE, SQ_REFETCH = 1, 2
SQ_E_PRED, SQ_E_MEAS, SQ_E_FK, SQ_E_REFETCH, SQ_E_ATTR = 0, 2, 4, 8, 12
SQ_STATUS = {0: "OK", 1: "BAD_OID", 2: "BAD_KEY", 3: "DUP_KEY", 4: "EVENTS", 5: "KEY_SPAN", 6: "GROUPS"}


class SqDim(ctypes.Structure):
    _fields_ = [("fk", AheadCol), ("key", AheadCol), ("attr", AheadCol), ("sel", ctypes.c_void_p),
                ("sel_count", ctypes.c_void_p), ("sel_a_inv", ctypes.c_uint64), ("flags", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("key_index", ctypes.c_void_p)]


class SqPred(ctypes.Structure):
    _fields_ = [("col", AheadCol), ("op1", ctypes.c_int32), ("op2", ctypes.c_int32), ("comb", ctypes.c_int32),
                ("reserved", ctypes.c_uint32), ("th1", ctypes.c_uint64), ("th2", ctypes.c_uint64)]


class SqArgs(ctypes.Structure):
    _fields_ = [("n_fact", ctypes.c_uint64), ("n_dims", ctypes.c_uint32), ("n_preds", ctypes.c_uint32),
                ("meas_kind", ctypes.c_uint32), ("meas_pre_mask", ctypes.c_uint32), ("dims", SqDim * SQ_MAX_DIMS),
                ("preds", SqPred * SQ_MAX_PREDS), ("meas", AheadCol * SQ_MAX_MEAS), ("a_result", ctypes.c_uint64),
                ("a_oid_err", ctypes.c_uint64), ("ev_scan", ctypes.c_void_p * 2)]


class SqOut(ctypes.Structure):
    _fields_ = [("sums", ctypes.c_void_p), ("first_row", ctypes.c_void_p), ("keys", ctypes.c_void_p * SQ_MAX_DIMS),
                ("groups_cap", ctypes.c_uint64), ("err", ctypes.c_void_p * SQ_KINDS),
                ("err_cap", ctypes.c_uint64 * SQ_KINDS), ("d_counts", ctypes.c_void_p)]


class ServeSel(ctypes.Structure):
    _fields_ = [("col", AheadCol), ("op1", ctypes.c_int32), ("op2", ctypes.c_int32), ("comb", ctypes.c_int32),
                ("reserved", ctypes.c_uint32), ("th1", ctypes.c_uint64), ("th2", ctypes.c_uint64),
                ("a_oid", ctypes.c_uint64), ("a_oid_inv", ctypes.c_uint64)]


class ServeQuery(ctypes.Structure):
    _fields_ = [("q", SqArgs), ("has_sel", ctypes.c_uint32 * SQ_MAX_DIMS), ("sel", ServeSel * SQ_MAX_DIMS)]


class ServeResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_uint64), ("groups", ctypes.c_uint64), ("kept", ctypes.c_uint64),
                ("events", ctypes.c_uint64), ("sums", ctypes.c_void_p), ("first_row", ctypes.c_void_p),
                ("keys", ctypes.c_void_p * SQ_MAX_DIMS), ("err", ctypes.c_void_p * SQ_KINDS),
                ("n_err", ctypes.c_uint64 * SQ_KINDS), ("sel_err", ctypes.c_void_p * SQ_MAX_DIMS),
                ("n_sel_err", ctypes.c_uint64 * SQ_MAX_DIMS), ("t_arrival", ctypes.c_double),
                ("t_start", ctypes.c_double), ("t_finish", ctypes.c_double)]


_u64 = ctypes.c_uint64
_u32 = ctypes.c_uint32
_int = ctypes.c_int
_vp = ctypes.c_void_p
_colp = ctypes.POINTER(AheadCol)
_SIGS = {
    "ahead_hip_last_error": ([], ctypes.c_char_p),
    "ahead_hip_version": ([], ctypes.c_char_p),
    "ahead_hip_ws_create": ([_int, ctypes.POINTER(_vp)], _int),
    "ahead_hip_ws_destroy": ([_vp], _int),
    "ahead_hip_ws_reserve": ([_vp, ctypes.c_size_t], _int),
    "ahead_hip_ws_bytes": ([_vp], ctypes.c_size_t),
    "ahead_hip_set_tiling": ([_u32, _u32], _int),
    "ahead_hip_get_tiling": ([ctypes.POINTER(_u32), ctypes.POINTER(_u32)], _int),
    "ahead_hip_select_an": ([_colp, _int, _u64, _int, _u64, _int, _u64, _u64, _vp, _u64, _vp, _u64,
                             _vp, _vp, _vp, _vp], _int),
    "ahead_hip_select_an_noalloc": ([_colp, _int, _u64, _int, _u64, _int, _u64, _u64, _vp, _u64, _vp, _u64,
                                     _vp, _vp, _vp], _int),
    "ahead_hip_encode_an": ([_colp, _u64, _u32, _vp, _vp], _int),
    "ahead_hip_check_an": ([_colp, _u64, _vp, _u64, _vp, _vp, _vp], _int),
    "ahead_hip_decode_an": ([_colp, _u64, _vp, _vp, _u64, _vp, _vp, _vp], _int),
    "ahead_hip_check_and_decode_an": ([_colp, _u64, _vp, _u64, _vp, _u64, _vp, _vp, _vp], _int),
    "ahead_hip_star1_pipeline": ([ctypes.POINTER(Star1Args), ctypes.POINTER(Star1Out), _vp, _vp], _int),
    "ahead_hip_star1_build": ([ctypes.POINTER(Star1Args), ctypes.POINTER(Star1Out), _vp, _vp], _int),
    "ahead_hip_star1_probe": ([ctypes.POINTER(Star1Args), ctypes.POINTER(Star1Out), _vp, _vp], _int),
    "ahead_hip_fetchjoin_an": ([_colp, _colp, _int, _u64, _u64, Buf, Buf, Buf, _vp, _vp, _vp], _int),
    "ahead_hip_matchjoin_pos_an": ([_colp, _colp, _colp, _int, _u64, _u32, _int, _u64, _u32, _u64, Buf, Buf, Buf,
                                    Buf, Buf, _vp, _vp, _vp], _int),
    "ahead_hip_matchjoin_merge_an": ([_colp, _colp, _colp, _colp, _int, _u64, _u32, _int, _u64, _u32, _u64, Buf, Buf,
                                      Buf, Buf, Buf, Buf, _vp, _vp, _vp], _int),
    "ahead_hip_matchjoin_seq_an": ([_colp, _colp, _colp, _colp, _int, _u64, _u32, _int, _u64, _u32, _u64, Buf, Buf,
                                    Buf, Buf, _vp, _vp, _vp], _int),
    "ahead_hip_check_and_decode2_an": ([_colp, _colp, _u64, _u32, _u32, _vp, _vp, _u64, Buf, Buf, _vp, _vp, _vp],
                                       _int),
    "ahead_hip_hashjoin_an": ([_colp, _colp, _colp, _colp, _int, _int, _u64, _u32, _int, _u64, _u32, _u64, Buf, Buf,
                               Buf, Buf, Buf, Buf, _vp, _vp, _vp], _int),
    "ahead_hip_groupby_an": ([_colp, _colp, _u64, _u32, _vp, Buf, Buf, Buf, _vp, _vp, _vp], _int),
    "ahead_hip_groupby2_an": ([_colp, _colp, _colp, _u64, _u32, _u64, _u32, _vp, Buf, Buf, Buf, Buf, _vp, _vp, _vp],
                              _int),
    "ahead_hip_aggregate_sum_grouped_an": ([_colp, _colp, _colp, _u64, _u64, _u64, _vp, Buf, Buf, Buf, _vp, _vp, _vp],
                                           _int),
    "ahead_hip_aggregate_mul_sum_an": ([_colp, _colp, _u64, _int, _u64, _u64, _vp, Buf, Buf, _vp, _vp, _vp], _int),
    "ahead_hip_arithmetic_an": ([_colp, _colp, _colp, _colp, _int, _u64, _u32, _u64, _vp, Buf, Buf, Buf, Buf, _vp, _vp,
                                 _vp], _int),
    "ahead_hip_select_batch_create": ([ctypes.POINTER(SelectJob), _u32, _u32, _u32, ctypes.POINTER(_vp)], _int),
    "ahead_hip_select_batch_launch": ([_vp, _vp], _int),
    "ahead_hip_select_batch_destroy": ([_vp], _int),
    "ahead_hip_select_batch_info": ([_vp, ctypes.POINTER(_u32), ctypes.POINTER(_u32)], _int),
    "ahead_hip_star_query": ([ctypes.POINTER(SqArgs), ctypes.POINTER(SqOut), _vp, _vp], _int),
    "ahead_hip_star_query_ws_bytes": ([ctypes.POINTER(SqArgs)], ctypes.c_size_t),
    "ahead_hip_star_query_batch": ([ctypes.POINTER(ctypes.POINTER(SqArgs)), ctypes.POINTER(ctypes.POINTER(SqOut)), _u32,
                                    _vp, _vp], _int),
    "ahead_hip_star_query_batch_ws_bytes": ([ctypes.POINTER(ctypes.POINTER(SqArgs)), _u32], ctypes.c_size_t),
    "ahead_serve_create": ([_int, _u32, _u32, _u64, ctypes.POINTER(_vp)], _int),
    "ahead_serve_submit": ([_vp, ctypes.POINTER(ServeQuery), ctypes.POINTER(_u64)], _int),
    "ahead_serve_wait": ([_vp, _u64, ctypes.POINTER(ServeResult)], _int),
    "ahead_serve_release": ([_vp, _u64], _int),
    "ahead_serve_clock": ([_vp], ctypes.c_double),
    "ahead_serve_destroy": ([_vp], _int),
    "ahead_hip_star_key_index_create": ([_colp, ctypes.POINTER(_vp), _vp], _int),
    "ahead_hip_star_key_index_destroy": ([_vp], _int),
    "ahead_hip_star_key_index_status": ([_vp], _int),
    "ahead_hip_gen_uniform": ([_vp, _u32, _u64, _u64, _u64, _u64, _u64, _u64, _vp], _int),
    "ahead_hip_inject_flips": ([_vp, _u32, _u64, _u64, _u64, _u64, _vp], _int),
}

_lib = None


def load(path: str | os.PathLike | None = None) -> ctypes.CDLL:
    """Loads libahead_hip.so once; raises if it is missing (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is not built; run `python -m ahead_amd.build` first")
    lib = ctypes.CDLL(str(p))
    for name, (args, res) in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != AHEAD_OK:
        raise AheadError(status, load().ahead_hip_last_error().decode())


# ------------------------------------------------------------------------------------------------
# device memory (torch supplies allocations and the stream)
# ------------------------------------------------------------------------------------------------

def _torch():
    import torch  # imported lazily so CPU-only users of this module need no GPU

    return torch


class DeviceBuffer:
    """Raw device bytes held by a torch uint8 tensor."""

    def __init__(self, nbytes: int, device: int = 0, zero: bool = False):
        torch = _torch()
        n = max(int(nbytes), 1)
        self.t = (torch.zeros if zero else torch.empty)(n, dtype=torch.uint8, device=f"cuda:{device}")
        self.nbytes = int(nbytes)

    @property
    def ptr(self) -> int:
        return self.t.data_ptr()

    def view(self, offset: int, nbytes: int) -> "DeviceBuffer":
        """nbytes of this buffer from offset on, sharing its memory."""
        b = DeviceBuffer.__new__(DeviceBuffer)
        b.t = self.t[int(offset): int(offset) + max(int(nbytes), 1)]
        b.nbytes = int(nbytes)
        return b

    def prefix_copy(self, nbytes: int) -> "DeviceBuffer":
        """A new buffer holding the first nbytes (a device copy on the current stream)."""
        b = DeviceBuffer.__new__(DeviceBuffer)
        b.t = self.t[: max(int(nbytes), 1)].clone()
        b.nbytes = int(nbytes)
        return b

    @staticmethod
    def from_numpy(arr: np.ndarray, device: int = 0) -> "DeviceBuffer":
        torch = _torch()
        arr = np.ascontiguousarray(arr)
        b = DeviceBuffer(arr.nbytes, device)
        if arr.nbytes:
            b.t[: arr.nbytes].copy_(torch.from_numpy(arr.view(np.uint8).reshape(-1)))
        return b

    def to_numpy(self, dtype, count: int | None = None) -> np.ndarray:
        dtype = np.dtype(dtype)
        count = self.nbytes // dtype.itemsize if count is None else int(count)
        if count == 0:
            return np.zeros(0, dtype=dtype)
        return self.t[: count * dtype.itemsize].cpu().numpy().view(dtype).copy()


def current_stream() -> int:
    torch = _torch()
    return torch.cuda.current_stream().cuda_stream


def synchronize() -> None:
    """Waits for the calling thread's current stream (every wrapper issues its work there), so
    threads serving queries on their own streams do not wait for each other."""
    _torch().cuda.current_stream().synchronize()


@dataclass
class Column:
    """A device-resident BAT side: buffer (None for void) + metadata."""

    buf: DeviceBuffer | None
    n: int
    meta: an.ColMeta

    def desc(self) -> AheadCol:
        m = self.meta
        return AheadCol(
            self.buf.ptr if self.buf is not None else None,
            self.n, m.type.width, int(m.type.encoded), m.type.unenc_bits, 0,
            m.a & an.MASK64, m.a_inv & an.MASK64, m.type.max_u & an.MASK64, m.seqbase,
        )

    @staticmethod
    def from_numpy(arr: np.ndarray, meta: an.ColMeta, device: int = 0) -> "Column":
        assert arr.dtype.itemsize == meta.width, (arr.dtype, meta)
        return Column(DeviceBuffer.from_numpy(arr, device), arr.size, meta)

    @staticmethod
    def void(n: int, seqbase: int = 0) -> "Column":
        return Column(None, n, an.plain_meta(an.VOID, seqbase))

    def to_numpy(self) -> np.ndarray:
        assert self.buf is not None
        return self.buf.to_numpy(np.dtype(f"u{self.meta.width}"), self.n)


class Workspace:
    def __init__(self, device: int = 0):
        self.lib = load()
        h = ctypes.c_void_p()
        check(self.lib.ahead_hip_ws_create(device, ctypes.byref(h)))
        self.h = h
        self.device = device

    def close(self) -> None:
        if self.h:
            check(self.lib.ahead_hip_ws_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_ws_local = threading.local()


def workspace(device: int = 0) -> Workspace:
    """The calling thread's workspace on `device` (a workspace serves one stream at a time, and each
    thread issues its work on its own current stream)."""
    ws = getattr(_ws_local, "ws", None)
    if ws is None:
        ws = _ws_local.ws = {}
    if device not in ws:
        ws[device] = Workspace(device)
    return ws[device]


def set_tiling(max_subtiles: int, min_tiles: int) -> None:
    check(load().ahead_hip_set_tiling(max_subtiles, min_tiles))


def get_tiling() -> tuple[int, int]:
    a, b = _u32(), _u32()
    check(load().ahead_hip_get_tiling(ctypes.byref(a), ctypes.byref(b)))
    return a.value, b.value


class tiling:
    """Context manager that sets the operator tiling and restores the previous one."""

    def __init__(self, max_subtiles: int, min_tiles: int):
        self.new = (max_subtiles, min_tiles)

    def __enter__(self):
        self.old = get_tiling()
        set_tiling(*self.new)

    def __exit__(self, *exc):
        set_tiling(*self.old)


def _counts(buf: DeviceBuffer, k: int) -> list[int]:
    return [int(x) for x in buf.to_numpy(np.uint64, k)]


# ------------------------------------------------------------------------------------------------
# operators (synchronous convenience wrappers: run, read counts, re-run if an output overflowed)
# ------------------------------------------------------------------------------------------------

def select_an(col: Column, op1: int, th1: int, op2: int = LT, th2: int = 0, comb: int = COMB_NONE,
              a_head: int = an.A_OID, a_oid: int = an.A_OID_ERR, with_sum: bool = False,
              err_cap: int = 4096, stream: int | None = None):
    """selectAN -> (oids uint64[], err uint64[], decoded-sum or None)."""
    lib = load()
    ws = workspace()
    s = current_stream() if stream is None else stream
    oids = DeviceBuffer(8 * col.n)
    cnt = DeviceBuffer(16)
    dsum = DeviceBuffer(8) if with_sum else None
    d = col.desc()
    while True:
        err = DeviceBuffer(8 * err_cap)
        check(lib.ahead_hip_select_an(ctypes.byref(d), op1, th1 & an.MASK64, op2, th2 & an.MASK64, comb,
                                      a_head, a_oid, oids.ptr, col.n, err.ptr, err_cap, cnt.ptr,
                                      dsum.ptr if dsum else None, ws.h, s))
        synchronize()
        n_oids, n_err = _counts(cnt, 2)
        if n_err <= err_cap:
            break
        err_cap = n_err
    total = int(dsum.to_numpy(np.uint64, 1)[0]) if dsum else None
    return oids.to_numpy(np.uint64, n_oids), err.to_numpy(np.uint64, n_err), total


class SelectBatch:
    """Many selectAN + decoded-sum instances in one persistent launch (ahead_hip_select_batch_*).

    jobs: list of dicts with keys col (Column), op1, th1, and optionally op2, th2, comb, err_cap.
    Outputs are allocated per job (OIDs: one per row; errors: err_cap, grown on overflow)."""

    def __init__(self, jobs: list[dict], max_subtiles: int = 16, blocks: int = 0,
                 a_head: int = an.A_OID, a_oid: int = an.A_OID_ERR):
        self.lib = load()
        self.jobs = jobs
        self.max_subtiles, self.blocks = max_subtiles, blocks
        self.a_head, self.a_oid = a_head, a_oid
        self.err_caps = [int(j.get("err_cap", 1024)) for j in jobs]
        self.h = None
        self._create()

    def _create(self):
        n = len(self.jobs)
        self.oids = [DeviceBuffer(max(8, 8 * j["col"].n)) for j in self.jobs]
        self.errs = [DeviceBuffer(max(8, 8 * c)) for c in self.err_caps]
        self.counts = DeviceBuffer(16 * n, zero=True)
        self.sums = DeviceBuffer(8 * n, zero=True)
        arr = (SelectJob * n)()
        self._descs = []
        for i, j in enumerate(self.jobs):
            d = j["col"].desc()
            self._descs.append(d)
            arr[i] = SelectJob(d, j["op1"], j.get("op2", LT), j.get("comb", COMB_NONE), 0, j["th1"] & an.MASK64,
                               j.get("th2", 0) & an.MASK64, self.a_head, self.a_oid, self.oids[i].ptr, j["col"].n,
                               self.errs[i].ptr, self.err_caps[i], self.counts.ptr + 16 * i, self.sums.ptr + 8 * i)
        h = ctypes.c_void_p()
        check(self.lib.ahead_hip_select_batch_create(arr, n, self.max_subtiles, self.blocks, ctypes.byref(h)))
        self.h = h

    def info(self) -> tuple[int, int]:
        t, b = ctypes.c_uint32(), ctypes.c_uint32()
        check(self.lib.ahead_hip_select_batch_info(self.h, ctypes.byref(t), ctypes.byref(b)))
        return t.value, b.value

    def launch(self, stream: int | None = None) -> None:
        check(self.lib.ahead_hip_select_batch_launch(self.h, current_stream() if stream is None else stream))

    def results(self) -> list[tuple[np.ndarray, np.ndarray, int]]:
        """Synchronises; re-creates and re-launches with larger error buffers if any overflowed."""
        while True:
            synchronize()
            c = self.counts.to_numpy(np.uint64, 2 * len(self.jobs)).reshape(-1, 2)
            over = [i for i in range(len(self.jobs)) if int(c[i, 1]) > self.err_caps[i]]
            if not over:
                break
            for i in over:
                self.err_caps[i] = int(c[i, 1])
            self.close()
            self._create()
            self.launch()
        sums = self.sums.to_numpy(np.uint64, len(self.jobs))
        return [(self.oids[i].to_numpy(np.uint64, int(c[i, 0])), self.errs[i].to_numpy(np.uint64, int(c[i, 1])),
                 int(sums[i])) for i in range(len(self.jobs))]

    def close(self) -> None:
        if self.h is not None and self.h.value:
            check(self.lib.ahead_hip_select_batch_destroy(self.h))
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def check_an(col: Column, a_oid: int = an.A_OID_ERR, err_cap: int = 4096) -> np.ndarray:
    lib, ws = load(), workspace()
    cnt = DeviceBuffer(16)
    d = col.desc()
    while True:
        err = DeviceBuffer(8 * err_cap)
        check(lib.ahead_hip_check_an(ctypes.byref(d), a_oid, err.ptr, err_cap, cnt.ptr, ws.h, current_stream()))
        synchronize()
        (n_err,) = _counts(cnt, 1)
        if n_err <= err_cap:
            return err.to_numpy(np.uint64, n_err)
        err_cap = n_err


def _unenc_dtype(meta: an.ColMeta):
    return np.dtype(f"u{meta.type.unenc_bits // 8}")


def decode_an(col: Column, a_oid: int = an.A_OID_ERR, err_cap: int = 4096):
    lib, ws = load(), workspace()
    dt = _unenc_dtype(col.meta)
    dec = DeviceBuffer(dt.itemsize * col.n)
    cnt = DeviceBuffer(16)
    d = col.desc()
    while True:
        err = DeviceBuffer(8 * err_cap)
        check(lib.ahead_hip_decode_an(ctypes.byref(d), a_oid, dec.ptr, err.ptr, err_cap, cnt.ptr, ws.h,
                                      current_stream()))
        synchronize()
        n_dec, n_err = _counts(cnt, 2)
        if n_err <= err_cap:
            return dec.to_numpy(dt, n_dec), err.to_numpy(np.uint64, n_err)
        err_cap = n_err


def check_and_decode_an(col: Column, a_oid: int = an.A_OID_ERR, err_cap: int = 4096):
    lib, ws = load(), workspace()
    dt = _unenc_dtype(col.meta)
    dec = DeviceBuffer(dt.itemsize * col.n)
    cnt = DeviceBuffer(16)
    d = col.desc()
    while True:
        err = DeviceBuffer(8 * err_cap)
        check(lib.ahead_hip_check_and_decode_an(ctypes.byref(d), a_oid, dec.ptr, col.n, err.ptr, err_cap,
                                                cnt.ptr, ws.h, current_stream()))
        synchronize()
        n_dec, n_err = _counts(cnt, 2)
        if n_err <= err_cap:
            return dec.to_numpy(dt, n_dec), err.to_numpy(np.uint64, n_err)
        err_cap = n_err


def encode_an(col: Column, a: int, out_width: int) -> np.ndarray:
    lib = load()
    out = DeviceBuffer(out_width * col.n)
    d = col.desc()
    check(lib.ahead_hip_encode_an(ctypes.byref(d), a, out_width, out.ptr, current_stream()))
    synchronize()
    return out.to_numpy(np.dtype(f"u{out_width}"), col.n)


def encode_an_into(col: Column, a: int, out_width: int, out: DeviceBuffer) -> None:
    """out[i] = (code)col[i] * a in out_width bytes, asynchronously on the current stream."""
    d = col.desc()
    check(load().ahead_hip_encode_an(ctypes.byref(d), a & an.MASK64, out_width, out.ptr, current_stream()))


def encode_void_into(buf: DeviceBuffer, n: int, a: int, out_width: int, seqbase: int = 0) -> None:
    """buf[i] = (seqbase + i) * a: the codes of a dense key / OID column."""
    d = Column.void(n, seqbase).desc()
    check(load().ahead_hip_encode_an(ctypes.byref(d), a, out_width, buf.ptr, current_stream()))


def gen_uniform(buf: DeviceBuffer, width: int, n: int, seed: int, stream_id: int, lo: int, hi: int,
                a: int = 0) -> None:
    check(load().ahead_hip_gen_uniform(buf.ptr, width, n, seed, stream_id, lo, hi, a, current_stream()))


def inject_flips(buf: DeviceBuffer, width: int, n: int, seed: int, stream_id: int, threshold: int) -> None:
    check(load().ahead_hip_inject_flips(buf.ptr, width, n, seed, stream_id, threshold, current_stream()))


STAR1_FIELDS = ["sums", "groid", "key1", "key2", "e_fk", "e_dk", "e_a1", "e_a2", "e_m", "e_grp"]


class Star1:
    """Device buffers + argument blocks of one fused star-join configuration (reusable)."""

    def __init__(self, fact_fk: Column, fact_m: Column, dim_key: Column, dim_a1: Column, dim_a2: Column,
                 a_oid_join: int = an.A_OID, a_oid_err: int = an.A_OID_ERR, a_oid_grp: int = an.A_OID,
                 a_result: int = an.A_RESULT, groups_cap: int = 8192, err_cap: int = 4096):
        self.cols = (fact_fk, fact_m, dim_key, dim_a1, dim_a2)
        self.args = Star1Args(fact_fk.desc(), fact_m.desc(), dim_key.desc(), dim_a1.desc(), dim_a2.desc(),
                              a_oid_join, a_oid_err, a_oid_grp, a_result)
        self.groups_cap = groups_cap
        self.err_cap = err_cap
        self._alloc()

    def _alloc(self):
        g, e = self.groups_cap, self.err_cap
        self.bufs = {k: DeviceBuffer(8 * g) for k in ("sums", "groid", "key1", "key2")}
        self.bufs.update({k: DeviceBuffer(8 * e) for k in ("e_fk", "e_dk", "e_a1", "e_a2", "e_m")})
        self.counts = DeviceBuffer(8 * 9)
        b = self.bufs
        self.out = Star1Out(b["sums"].ptr, b["groid"].ptr, b["key1"].ptr, b["key2"].ptr, g,
                            b["e_fk"].ptr, b["e_dk"].ptr, b["e_a1"].ptr, b["e_a2"].ptr, b["e_m"].ptr,
                            e, e, e, e, e, self.counts.ptr)

    def launch(self, stream: int | None = None) -> None:
        """Build + probe (one full pipeline)."""
        s = current_stream() if stream is None else stream
        st = load().ahead_hip_star1_pipeline(ctypes.byref(self.args), ctypes.byref(self.out), workspace().h, s)
        if st:
            check(st)

    def build(self, stream: int | None = None) -> None:
        s = current_stream() if stream is None else stream
        check(load().ahead_hip_star1_build(ctypes.byref(self.args), ctypes.byref(self.out), workspace().h, s))

    def probe(self, stream: int | None = None) -> None:
        s = current_stream() if stream is None else stream
        st = load().ahead_hip_star1_probe(ctypes.byref(self.args), ctypes.byref(self.out), workspace().h, s)
        if st:
            check(st)

    def results(self) -> dict:
        """Synchronises, re-runs with larger buffers if an output was truncated, returns numpy."""
        while True:
            synchronize()
            c = [int(x) for x in self.counts.to_numpy(np.uint64, 9)]
            need_e = max(c[1:6])
            if c[0] <= self.groups_cap and need_e <= self.err_cap:
                break
            self.groups_cap = max(self.groups_cap, c[0])
            self.err_cap = max(self.err_cap, need_e)
            self._alloc()
            self.launch()
        out = {"status": c[7], "kept": c[8]}
        for k, i in (("sums", 0), ("groid", 0), ("key1", 0), ("key2", 0)):
            out[k] = self.bufs[k].to_numpy(np.uint64, c[i])
        for k, i in (("e_fk", 1), ("e_dk", 2), ("e_a1", 3), ("e_a2", 4), ("e_m", 5)):
            out[k] = self.bufs[k].to_numpy(np.uint64, c[i])
        out["e_grp"] = np.zeros(c[6], dtype=np.uint64)
        return out


def star1_pipeline(fact_fk: Column, fact_m: Column, dim_key: Column, dim_a1: Column, dim_a2: Column,
                   **kw) -> dict:
    st = Star1(fact_fk, fact_m, dim_key, dim_a1, dim_a2, **kw)
    st.launch()
    return st.results()


def star_args(n_fact: int, dims: list, preds: list, meas: list, kind: int, pre_mask: int = 0,
              a_result: int = an.A_RESULT, a_oid_err: int = an.A_OID_ERR) -> SqArgs:
    """The ahead_sq_args block of a star query (dims, preds, meas as StarQuery takes them)."""
    a = SqArgs()
    a.n_fact = n_fact
    a.n_dims = len(dims)
    a.n_preds = len(preds)
    a.meas_kind = kind
    a.meas_pre_mask = pre_mask
    for i, d in enumerate(dims):
        x = a.dims[i]
        x.fk = d["fk"].desc()
        x.key = d["key"].desc()
        x.attr = d["attr"].desc() if d.get("attr") is not None else AheadCol()
        x.sel = d["sel"].ptr if d.get("sel") is not None else None
        x.sel_count = d["sel_count"].ptr if d.get("sel") is not None else None
        x.sel_a_inv = d.get("sel_a_inv", an.A_OID_INV) & an.MASK64
        x.flags = d.get("flags", 0)
        ki = d.get("key_index")
        x.key_index = ki.h if ki is not None else None
    for i, (col, op1, th1, op2, th2, comb) in enumerate(preds):
        a.preds[i] = SqPred(col.desc(), op1, op2, comb, 0, th1 & an.MASK64, th2 & an.MASK64)
    for i, m in enumerate(meas):
        a.meas[i] = m.desc()
    a.a_result = a_result
    a.a_oid_err = a_oid_err
    return a


class NativeServer:
    """The request loop of libahead_hip.so (ahead_serve_*): slots are C++ threads with their own
    streams, workspaces and arenas; submit() returns a ticket, wait() the query's outputs."""

    def __init__(self, slots: int = 1, max_batch: int = 16, hbm_budget: int = 0, device: int = 0):
        h = _vp()
        check(load().ahead_serve_create(device, slots, max_batch, int(hbm_budget), ctypes.byref(h)))
        self.h = h

    def clock(self) -> float:
        return float(load().ahead_serve_clock(self.h))

    def submit(self, q: "ServeQuery") -> int:
        t = _u64()
        check(load().ahead_serve_submit(self.h, ctypes.byref(q), ctypes.byref(t)))
        return int(t.value)

    def wait(self, ticket: int, dims: int) -> dict:
        """The query's outputs as numpy (copied), its times on the server's clock; the ticket is
        released."""
        lib = load()
        r = ServeResult()
        st = lib.ahead_serve_wait(self.h, ticket, ctypes.byref(r))
        if st:
            msg = lib.ahead_hip_last_error().decode()
            lib.ahead_serve_release(self.h, ticket)
            raise AheadError(st, msg)

        def arr(ptr, n):
            n = int(n)
            if n == 0 or not ptr:
                return np.zeros(0, dtype=np.uint64)
            return np.ctypeslib.as_array((ctypes.c_uint64 * n).from_address(ptr)).copy()

        ng = int(r.groups)
        out = {"status": int(r.status), "groups": ng, "kept": int(r.kept), "events": int(r.events),
               "sums": arr(r.sums, ng), "first_row": arr(r.first_row, ng),
               "keys": {d: arr(r.keys[d], ng) for d in range(dims) if r.keys[d]},
               "err": [arr(r.err[k], r.n_err[k]) for k in range(SQ_KINDS)],
               "sel_err": [arr(r.sel_err[d], r.n_sel_err[d]) for d in range(dims)],
               "t_arrival": r.t_arrival, "t_start": r.t_start, "t_finish": r.t_finish}
        check(lib.ahead_serve_release(self.h, ticket))
        return out

    def close(self) -> None:
        if self.h is not None:
            check(load().ahead_serve_destroy(self.h))
            self.h = None


class StarQuery:
    """Argument and output blocks of one fused star query (ahead_hip_star_query), reusable.

    dims: list of dicts in the chain's order with keys fk, key (Columns), attr (Column or None),
    sel (DeviceBuffer of encoded OIDs or None), sel_count (DeviceBuffer holding the count at offset
    0), sel_a_inv, flags.  preds: list of (Column, op1, th1, op2, th2, comb).  meas: list of Columns.
    """

    def __init__(self, n_fact: int, dims: list, preds: list, meas: list, kind: int, pre_mask: int = 0,
                 a_result: int = an.A_RESULT, a_oid_err: int = an.A_OID_ERR, groups_cap: int = 4096,
                 err_cap: int = 1024, ev_scan: tuple | None = None, out_arena: tuple | None = None):
        self.keep = (dims, preds, meas)  # the columns and selections stay alive with the query
        self.args = star_args(n_fact, dims, preds, meas, kind, pre_mask, a_result, a_oid_err)
        if ev_scan is not None:
            self.args.ev_scan[0], self.args.ev_scan[1] = ev_scan
        self.attr_dims = [i for i, d in enumerate(dims) if d.get("attr") is not None]
        self.groups_cap = groups_cap
        self.err_cap = err_cap
        self._alloc(out_arena)

    @staticmethod
    def out_bytes(n_attr: int, groups_cap: int, err_cap: int) -> int:
        """Bytes of one query's outputs carved from an arena (counts, sums, first rows, attribute
        codes, error vectors), 256-aligned."""
        b = 8 * (4 + SQ_KINDS) + 8 * groups_cap * (2 + n_attr) + 8 * err_cap * SQ_KINDS
        return (b + 255) // 256 * 256

    def _alloc(self, arena: tuple | None = None):
        """The output buffers: own allocations, or consecutive views of arena = (DeviceBuffer,
        offset) in the order out_bytes counts them."""
        g, e = self.groups_cap, self.err_cap
        if arena is None:
            new = DeviceBuffer
        else:
            buf, off = arena
            pos = [off]

            def new(nbytes):
                v = buf.view(pos[0], nbytes)
                pos[0] += nbytes
                return v
        self.counts = new(8 * (4 + SQ_KINDS))
        self.sums = new(8 * g)
        self.first = new(8 * g)
        self.keys = {d: new(8 * g) for d in self.attr_dims}
        self.errs = [new(8 * e) for _ in range(SQ_KINDS)]
        o = SqOut()
        o.sums = self.sums.ptr
        o.first_row = self.first.ptr
        for d, b in self.keys.items():
            o.keys[d] = b.ptr
        o.groups_cap = g
        for k in range(SQ_KINDS):
            o.err[k] = self.errs[k].ptr
            o.err_cap[k] = e
        o.d_counts = self.counts.ptr
        self.out = o

    def ws_bytes(self) -> int:
        """Workspace bytes this query uses (ahead_hip_star_query_ws_bytes)."""
        return int(load().ahead_hip_star_query_ws_bytes(ctypes.byref(self.args)))

    def launch(self, stream: int | None = None) -> None:
        s = current_stream() if stream is None else stream
        st = load().ahead_hip_star_query(ctypes.byref(self.args), ctypes.byref(self.out), workspace().h, s)
        if st:
            check(st)

    def results_from(self, h: np.ndarray, off: int) -> dict | None:
        """The results from a host copy h (uint8) of the arena this query's outputs were carved
        from at off; None if an output was truncated (results() then re-runs the query alone)."""
        g, e = self.groups_cap, self.err_cap
        c = [int(x) for x in h[off: off + 8 * (4 + SQ_KINDS)].view(np.uint64)]
        if c[0] > g or max(c[4:]) > e:
            return None

        def arr(pos, count):
            return h[pos: pos + 8 * count].view(np.uint64).copy()

        pos = off + 8 * (4 + SQ_KINDS)
        ng = c[0]
        sums, first = arr(pos, ng), arr(pos + 8 * g, ng)
        pos += 16 * g
        keys = {}
        for d in self.attr_dims:
            keys[d] = arr(pos, ng)
            pos += 8 * g
        err = []
        for k in range(SQ_KINDS):
            err.append(arr(pos, c[4 + k]))
            pos += 8 * e
        return {"status": c[2], "groups": ng, "kept": c[1], "events": c[3], "sums": sums, "first_row": first,
                "keys": keys, "err": err}

    def results(self) -> dict:
        """Synchronises, re-runs with larger buffers if an output was truncated, returns numpy."""
        while True:
            synchronize()
            c = [int(x) for x in self.counts.to_numpy(np.uint64, 4 + SQ_KINDS)]
            need_e = max(c[4:])
            if c[0] <= self.groups_cap and need_e <= self.err_cap:
                break
            self.groups_cap = max(self.groups_cap, c[0])
            self.err_cap = max(self.err_cap, need_e)
            self._alloc()
            self.launch()
        ng = c[0]
        return {"status": c[2], "groups": ng, "kept": c[1], "events": c[3],
                "sums": self.sums.to_numpy(np.uint64, ng), "first_row": self.first.to_numpy(np.uint64, ng),
                "keys": {d: b.to_numpy(np.uint64, ng) for d, b in self.keys.items()},
                "err": [self.errs[k].to_numpy(np.uint64, c[4 + k]) for k in range(SQ_KINDS)]}


SQ_MAX_BATCH = 8
SQ_MAX_SHAPE_PASS = 32


def _batch_ptrs(queries: list):
    n = len(queries)
    args = (ctypes.POINTER(SqArgs) * n)(*[ctypes.pointer(q.args) for q in queries])
    outs = (ctypes.POINTER(SqOut) * n)(*[ctypes.pointer(q.out) for q in queries])
    return n, args, outs


def star_query_batch(queries: list, stream: int | None = None) -> None:
    """Launches StarQuery objects over one fact table with shared fact scans
    (ahead_hip_star_query_batch); each query's results() reads its own outputs."""
    if not queries:
        return
    n, args, outs = _batch_ptrs(queries)
    s = current_stream() if stream is None else stream
    st = load().ahead_hip_star_query_batch(args, outs, n, workspace().h, s)
    if st:
        check(st)


def star_query_batch_ws_bytes(queries: list) -> int:
    n, args, _ = _batch_ptrs(queries)
    return int(load().ahead_hip_star_query_batch_ws_bytes(args, n))


# ------------------------------------------------------------------------------------------------
# per-operator wrappers (synchronous; outputs re-allocated and the call repeated on truncation)
# ------------------------------------------------------------------------------------------------

def _run(call, outs, ncounts, count_of):
    """outs: list of [elem_bytes, cap]; call(bufs, counts) launches; count_of[i] = counts index of
    output i.  Returns (numpy uint8 buffers list, counts)."""
    while True:
        bufs = [DeviceBuffer(max(1, eb * cap)) for eb, cap in outs]
        cnt = DeviceBuffer(8 * ncounts, zero=True)
        call(bufs, cnt)
        synchronize()
        c = _counts(cnt, ncounts)
        grow = False
        for i, (eb, cap) in enumerate(outs):
            if count_of[i] is not None and c[count_of[i]] > cap:
                outs[i][1] = c[count_of[i]]
                grow = True
        if not grow:
            return bufs, c


def _b(buf: DeviceBuffer, cap: int) -> Buf:
    return Buf(buf.ptr, cap)


def fetchjoin_an(oids: Column, col: Column, reenc: int = 0, a_reenc: int = 0, a_oid: int = an.A_OID_ERR,
                 err_cap: int = 1024):
    lib, ws = load(), workspace()
    ow = col.meta.width if col.meta.type.encoded else (8 if col.meta.width >= 4 else 2 * col.meta.width)
    outs = [[ow, oids.n], [8, err_cap], [8, err_cap]]
    do, dc = oids.desc(), col.desc()

    def call(b, cnt):
        check(lib.ahead_hip_fetchjoin_an(ctypes.byref(do), ctypes.byref(dc), reenc, a_reenc & an.MASK64, a_oid,
                                         _b(b[0], outs[0][1]), _b(b[1], outs[1][1]), _b(b[2], outs[2][1]), cnt.ptr,
                                         ws.h, current_stream()))

    b, c = _run(call, outs, 3, [0, 1, 2])
    return (b[0].to_numpy(np.dtype(f"u{ow}"), c[0]), b[1].to_numpy(np.uint64, c[1]), b[2].to_numpy(np.uint64, c[2]))


def matchjoin_pos_an(h1: Column, t1: Column, t2: Column, reenc_h=0, f_h1=0, wout_h=8, reenc_t=0, f_t2=0, wout_t=8,
                     a_oid: int = an.A_OID_ERR, err_cap: int = 1024):
    lib, ws = load(), workspace()
    outs = [[wout_h, t1.n], [wout_t, t1.n], [8, err_cap], [8, err_cap], [8, err_cap]]
    d1, d2, d3 = h1.desc(), t1.desc(), t2.desc()

    def call(b, cnt):
        check(lib.ahead_hip_matchjoin_pos_an(ctypes.byref(d1), ctypes.byref(d2), ctypes.byref(d3), reenc_h, f_h1,
                                             wout_h, reenc_t, f_t2, wout_t, a_oid, *[_b(x, o[1]) for x, o in
                                                                                   zip(b, outs)], cnt.ptr, ws.h,
                                             current_stream()))

    b, c = _run(call, outs, 4, [0, 0, 1, 2, 3])
    return (b[0].to_numpy(np.dtype(f"u{wout_h}"), c[0]), b[1].to_numpy(np.dtype(f"u{wout_t}"), c[0]),
            b[2].to_numpy(np.uint64, c[1]), b[3].to_numpy(np.uint64, c[2]), b[4].to_numpy(np.uint64, c[3]))


def matchjoin_merge_an(h1: Column, t1: Column, h2: Column, t2: Column, reenc_h=0, f_h1=0, wout_h=8, reenc_t=0,
                       f_t2=0, wout_t=8, a_oid: int = an.A_OID_ERR, err_cap: int = 1024):
    lib, ws = load(), workspace()
    cap = max(min(t1.n, h2.n), 1)
    outs = [[wout_h, cap], [wout_t, cap], [8, err_cap], [8, err_cap], [8, err_cap], [8, err_cap]]
    d = [x.desc() for x in (h1, t1, h2, t2)]

    def call(b, cnt):
        check(lib.ahead_hip_matchjoin_merge_an(*[ctypes.byref(x) for x in d], reenc_h, f_h1, wout_h, reenc_t, f_t2,
                                               wout_t, a_oid, *[_b(x, o[1]) for x, o in zip(b, outs)], cnt.ptr,
                                               ws.h, current_stream()))

    b, c = _run(call, outs, 5, [0, 0, 1, 2, 3, 4])
    return (b[0].to_numpy(np.dtype(f"u{wout_h}"), c[0]), b[1].to_numpy(np.dtype(f"u{wout_t}"), c[0]),
            *[b[2 + k].to_numpy(np.uint64, c[1 + k]) for k in range(4)])


def matchjoin_seq_an(h1: Column, t1: Column, h2: Column, t2: Column, reenc_h=0, f_h1=0, wout_h=8, reenc_t=0,
                     f_t2=0, wout_t=8, a_oid: int = an.A_OID_ERR, err_cap: int = 1024):
    lib, ws = load(), workspace()
    cap = max(min(t1.n, h2.n), 1)
    outs = [[wout_h, cap], [wout_t, cap], [8, err_cap], [8, err_cap]]
    d = [x.desc() for x in (h1, t1, h2, t2)]

    def call(b, cnt):
        check(lib.ahead_hip_matchjoin_seq_an(*[ctypes.byref(x) for x in d], reenc_h, f_h1, wout_h, reenc_t, f_t2,
                                             wout_t, a_oid, *[_b(x, o[1]) for x, o in zip(b, outs)], cnt.ptr, ws.h,
                                             current_stream()))

    b, c = _run(call, outs, 3, [0, 0, 1, 2])
    return (b[0].to_numpy(np.dtype(f"u{wout_h}"), c[0]), b[1].to_numpy(np.dtype(f"u{wout_t}"), c[0]),
            b[2].to_numpy(np.uint64, c[1]), b[3].to_numpy(np.uint64, c[2]))


def check_and_decode2_an(head: Column, tail: Column, a_oid: int = an.A_OID_ERR, err_cap: int = 1024):
    """checkAndDecodeAN<Head, Tail>: (decoded heads, decoded tails, head errors, tail errors)."""
    lib, ws = load(), workspace()
    wh = head.meta.type.unenc_bits // 8 if head.meta.type.encoded else head.meta.width
    wt = tail.meta.type.unenc_bits // 8 if tail.meta.type.encoded else tail.meta.width
    n = tail.n
    outs = [[wh, max(n, 1)], [wt, max(n, 1)], [8, err_cap], [8, err_cap]]
    dh, dt = head.desc(), tail.desc()

    def call(b, cnt):
        check(lib.ahead_hip_check_and_decode2_an(ctypes.byref(dh), ctypes.byref(dt), a_oid, wh, wt, b[0].ptr, b[1].ptr,
                                                 outs[0][1], _b(b[2], outs[2][1]), _b(b[3], outs[3][1]), cnt.ptr,
                                                 ws.h, current_stream()))

    b, c = _run(call, outs, 3, [0, 0, 1, 2])
    return (b[0].to_numpy(np.dtype(f"u{wh}"), c[0]), b[1].to_numpy(np.dtype(f"u{wt}"), c[0]),
            b[2].to_numpy(np.uint64, c[1]), b[3].to_numpy(np.uint64, c[2]))


def hashjoin_an(h1: Column, t1: Column, h2: Column, t2: Column, build_right=1, reenc_h=0, f_h1=0, wout_h=8,
                reenc_t=0, f_t2=0, wout_t=8, a_oid: int = an.A_OID_ERR, out_cap: int | None = None,
                err_cap: int = 1024):
    lib, ws = load(), workspace()
    cap = max(t1.n, t2.n) if out_cap is None else out_cap
    outs = [[wout_h, cap], [wout_t, cap], [8, err_cap], [8, err_cap], [8, err_cap], [8, err_cap]]
    ds = [x.desc() for x in (h1, t1, h2, t2)]

    def call(b, cnt):
        check(lib.ahead_hip_hashjoin_an(*[ctypes.byref(d) for d in ds], build_right, reenc_h, f_h1, wout_h, reenc_t,
                                        f_t2, wout_t, a_oid, *[_b(x, o[1]) for x, o in zip(b, outs)], cnt.ptr, ws.h,
                                        current_stream()))

    b, c = _run(call, outs, 6, [0, 0, 1, 2, 3, 4])
    if c[5] != 0:
        raise AheadError(6, f"hashjoin status {c[5]}")
    return (b[0].to_numpy(np.dtype(f"u{wout_h}"), c[0]), b[1].to_numpy(np.dtype(f"u{wout_t}"), c[0]),
            *(b[i].to_numpy(np.uint64, c[i - 1]) for i in range(2, 6)))


def groupby_an(head: Column, tail: Column, a_oid: int = an.A_OID, err_cap: int = 1024, table_bits: int = 0):
    lib, ws = load(), workspace()
    dh, dt = head.desc(), tail.desc()
    while True:
        outs = [[8, tail.n], [8, 4096], [8, err_cap], [8, err_cap]]

        def call(b, cnt):
            check(lib.ahead_hip_groupby_an(ctypes.byref(dh), ctypes.byref(dt), a_oid, table_bits, b[0].ptr,
                                           _b(b[1], outs[1][1]), _b(b[2], outs[2][1]), _b(b[3], outs[3][1]), cnt.ptr,
                                           ws.h, current_stream()))

        b, c = _run(call, outs, 6, [None, 1, 2, 3])
        if c[5] == 0:
            break
        table_bits = max(table_bits, 22) + 2
    return (b[0].to_numpy(np.uint64, c[0]), b[1].to_numpy(np.uint64, c[1]), b[2].to_numpy(np.uint64, c[2]),
            b[3].to_numpy(np.uint64, c[3]))


def groupby2_an(head: Column, tail: Column, grouping: Column, num_groups: int, larger_bits: int,
                a_oid: int = an.A_OID, err_cap: int = 1024, table_bits: int = 0):
    lib, ws = load(), workspace()
    dh, dt, dg = head.desc(), tail.desc(), grouping.desc()
    while True:
        outs = [[8, tail.n], [8, 4096], [8, err_cap], [8, err_cap], [8, err_cap]]

        def call(b, cnt):
            check(lib.ahead_hip_groupby2_an(ctypes.byref(dh), ctypes.byref(dt), ctypes.byref(dg), num_groups,
                                            larger_bits, a_oid, table_bits, b[0].ptr, *[_b(x, o[1]) for x, o in
                                                                                       zip(b[1:], outs[1:])],
                                            cnt.ptr, ws.h, current_stream()))

        b, c = _run(call, outs, 6, [None, 1, 2, 3, 4])
        if c[5] == 0:
            break
        table_bits = max(table_bits, 22) + 2
    return (b[0].to_numpy(np.uint64, c[0]), *(b[i].to_numpy(np.uint64, c[i]) for i in range(1, 5)))


def aggregate_sum_grouped_an(head: Column, tail: Column, grouping: Column, num_groups: int,
                             a_result: int = an.A_RESULT, a_oid: int = an.A_OID, err_cap: int = 1024):
    lib, ws = load(), workspace()
    outs = [[8, num_groups], [8, err_cap], [8, err_cap], [8, err_cap]]
    ds = [x.desc() for x in (head, tail, grouping)]

    def call(b, cnt):
        check(lib.ahead_hip_aggregate_sum_grouped_an(*[ctypes.byref(d) for d in ds], num_groups, a_result, a_oid,
                                                     b[0].ptr, *[_b(x, o[1]) for x, o in zip(b[1:], outs[1:])],
                                                     cnt.ptr, ws.h, current_stream()))

    b, c = _run(call, outs, 4, [None, 1, 2, 3])
    return (b[0].to_numpy(np.uint64, num_groups), *(b[i].to_numpy(np.uint64, c[i]) for i in range(1, 4)))


def aggregate_mul_sum_an(t1: Column, t2: Column, init: int = 0, result_encoded: int = 1, ra: int = an.A_RESULT,
                         a_oid: int = an.A_OID_ERR, err_cap: int = 1024):
    lib, ws = load(), workspace()
    outs = [[8, 1], [8, err_cap], [8, err_cap]]
    d1, d2 = t1.desc(), t2.desc()

    def call(b, cnt):
        check(lib.ahead_hip_aggregate_mul_sum_an(ctypes.byref(d1), ctypes.byref(d2), init, result_encoded, ra, a_oid,
                                                 b[0].ptr, _b(b[1], outs[1][1]), _b(b[2], outs[2][1]), cnt.ptr, ws.h,
                                                 current_stream()))

    b, c = _run(call, outs, 3, [None, 1, 2])
    return (int(b[0].to_numpy(np.uint64, 1)[0]), b[1].to_numpy(np.uint64, c[1]), b[2].to_numpy(np.uint64, c[2]))


def arithmetic_an(h1: Column, t1: Column, h2: Column, t2: Column, op: int, a_result: int, result_width: int,
                  a_oid: int = an.A_OID_ERR, err_cap: int = 1024):
    lib, ws = load(), workspace()
    outs = [[result_width, t1.n], [8, err_cap], [8, err_cap], [8, err_cap], [8, err_cap]]
    ds = [x.desc() for x in (h1, t1, h2, t2)]

    def call(b, cnt):
        check(lib.ahead_hip_arithmetic_an(*[ctypes.byref(d) for d in ds], op, a_result, result_width, a_oid, b[0].ptr,
                                          *[_b(x, o[1]) for x, o in zip(b[1:], outs[1:])], cnt.ptr, ws.h,
                                          current_stream()))

    b, c = _run(call, outs, 5, [None, 1, 2, 3, 4])
    return (b[0].to_numpy(np.dtype(f"u{result_width}"), c[0]), *(b[i].to_numpy(np.uint64, c[i]) for i in range(1, 5)))


class StarKeyIndex:
    """ahead_hip_star_key_index_create over a resident key column (kept alive with the index)."""

    def __init__(self, key: Column, stream: int | None = None):
        self.key = key
        h = ctypes.c_void_p()
        d = key.desc()
        check(load().ahead_hip_star_key_index_create(ctypes.byref(d), ctypes.byref(h),
                                                     current_stream() if stream is None else stream))
        self.h = h

    @property
    def status(self) -> int:
        return int(load().ahead_hip_star_key_index_status(self.h))

    def close(self) -> None:
        if self.h:
            check(load().ahead_hip_star_key_index_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HipEvents:
    """Raw hipEvent_t pairs (libamdhip64) for timing one kernel inside an entry point."""

    def __init__(self, count: int):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), _vp, _vp]
        self.hip.hipEventSynchronize.argtypes = [_vp]
        self.hip.hipEventDestroy.argtypes = [_vp]
        self.pairs = []
        for _ in range(count):
            e = [_vp(), _vp()]
            for x in e:
                if self.hip.hipEventCreate(ctypes.byref(x)) != 0:
                    raise AheadError(3, "hipEventCreate failed")
            self.pairs.append((e[0].value, e[1].value))

    def ms(self, i: int) -> float:
        a, b = self.pairs[i]
        self.hip.hipEventSynchronize(b)
        out = ctypes.c_float()
        if self.hip.hipEventElapsedTime(ctypes.byref(out), a, b) != 0:
            raise AheadError(3, "hipEventElapsedTime failed")
        return float(out.value)

    def close(self) -> None:
        for a, b in self.pairs:
            self.hip.hipEventDestroy(a)
            self.hip.hipEventDestroy(b)
        self.pairs = []
