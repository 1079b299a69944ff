"""ctypes binding of libahead_hip.so (include/ahead_hip.h) for tests, smoke and bench.

PyTorch only supplies device memory and the stream: every operator runs in the HIP library.
If the library is missing, loading raises instead of falling back to anything.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import an

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libahead_hip.so"

AHEAD_OK = 0
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
    "ahead_hip_encode_an": ([_colp, _u64, _u32, _vp, _vp], _int),
    "ahead_hip_check_an": ([_colp, _u64, _vp, _u64, _vp, _vp, _vp], _int),
    "ahead_hip_decode_an": ([_colp, _u64, _vp, _vp, _u64, _vp, _vp, _vp], _int),
    "ahead_hip_check_and_decode_an": ([_colp, _u64, _vp, _u64, _vp, _u64, _vp, _vp, _vp], _int),
    "ahead_hip_star1_pipeline": ([ctypes.POINTER(Star1Args), ctypes.POINTER(Star1Out), _vp, _vp], _int),
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
    _torch().cuda.synchronize()


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


_ws: dict[int, Workspace] = {}


def workspace(device: int = 0) -> Workspace:
    if device not in _ws:
        _ws[device] = Workspace(device)
    return _ws[device]


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
        s = current_stream() if stream is None else stream
        st = load().ahead_hip_star1_pipeline(ctypes.byref(self.args), ctypes.byref(self.out), workspace().h, s)
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
