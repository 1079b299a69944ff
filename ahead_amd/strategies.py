"""The reference's other execution strategies for the SSB queries, on the same operators.

  normal      the plain operators over unhardened values (ssbm-q*_normal.cpp): here the values are
              kept at their code width with A = 1 and no check (an.unchecked_type), which is the
              special case A = 1 of every AN operator;
  early       every column the query reads is checked and decoded first (checkAndDecodeAN, corrupt
              rows dropped), then the plain chain runs on the values (ssbm-q11_early.cpp:72-94);
  late        the plain operators run on the AN codes themselves (thresholds encoded, codes
              compared, joined and grouped as they are) and only the fetched group attributes and
              measures are checked and decoded, right before grouping and aggregation
              (ssbm-q11_late.cpp:70-120, ssbm-q21_late.cpp:85-160);
  continuous_reenc  the hardened chain re-encoding every join's and fetch's outputs with a
              different A (ssbm-q*_continuous_reenc.cpp);
  dmr / tmr   the plain chain on 2 / 3 replicas of the unhardened values, then a vote
              (ModularRedundant.hpp:111-180, ssbm-q11_dmr_seq.cpp:98-147); `concurrent` runs the
              replicas at the same time, one stream each (the OpenMP replicas of
              ssbm-q11_dmr_mt.cpp:85).

Every function takes the operator module (ahead_amd.bat on the GPU; the oracle's host BATs in the
tests) and returns the chain's result form.
"""
from __future__ import annotations

import threading
from concurrent.futures import ThreadPoolExecutor
from dataclasses import replace

import numpy as np

from . import an, ssb
from . import bat as B


class ModularityError(RuntimeError):
    """The replicas' results disagree (modularity_exception, ModularRedundant.hpp:32-35)."""


def same(a: dict, b: dict) -> bool:
    """Equal query results: rows, the Q1.x result or the groups, their sums and keys."""
    if a["rows"] != b["rows"] or a.get("result") != b.get("result") or a.get("groups") != b.get("groups"):
        return False
    if "sums" in a:
        return np.array_equal(a["sums"], b["sums"]) and all(np.array_equal(x, y) for x, y in zip(a["keys"], b["keys"]))
    return True


def vote(results: list) -> dict:
    """vote_majority_value (ModularRedundant.hpp:111-180): with two replicas their results must be
    equal; with three, the first if another equals it, else the second if the third equals it."""
    if len(results) == 2:
        if same(results[0], results[1]):
            return results[0]
        raise ModularityError("DMR: the replicas' results differ")
    if len(results) == 3:
        r0, r1, r2 = results
        if same(r0, r1) or same(r0, r2):
            return r0
        if same(r1, r2):
            return r1
        raise ModularityError("TMR: no two replicas agree")
    raise ValueError("modular redundancy takes 2 or 3 replicas")


def plain_tables(t: dict, ops=B) -> dict:
    """The schema's values, unhardened: every column decoded (decodeAN) and stored at its code
    width with A = 1 and no check."""
    return {name: ops.decodeAN(b)[0] for name, b in t.items()}


def replicas(plain: dict, k: int, ops=B) -> list:
    """k replicas of the unhardened tables: the tables themselves and k - 1 copies."""
    return [plain] + [{name: ops.copy(b) for name, b in plain.items()} for _ in range(k - 1)]


def normal(plain: dict, q: ssb.Query, ops=B) -> dict:
    return ssb.chain(plain, q, ops)


_POOLS: dict = {}
_local = threading.local()


def _pool(k: int) -> ThreadPoolExecutor:
    """k long-lived replica threads: each keeps its stream and its workspace across queries."""
    if k not in _POOLS:
        _POOLS[k] = ThreadPoolExecutor(max_workers=k)
    return _POOLS[k]


def _on_stream(fn):
    import torch

    s = getattr(_local, "stream", None)
    if s is None:
        s = _local.stream = torch.cuda.Stream()
    with torch.cuda.stream(s):
        r = fn()
    s.synchronize()
    return r


def modular(reps: list, q: ssb.Query, ops=B, concurrent: bool = False) -> dict:
    """DMR (2 replicas) or TMR (3): the plain chain once per replica, then the vote."""
    if concurrent and ops is B:
        results = list(_pool(len(reps)).map(lambda r: _on_stream(lambda: ssb.chain(r, q, ops)), reps))
    else:
        results = [ssb.chain(r, q, ops) for r in reps]
    return vote(results)


def columns_read(q: ssb.Query) -> list:
    """The columns a query's chain reads: fact predicates and measures, and per dimension its
    foreign key, key, selected attributes and group attribute."""
    cols = [c for c, _ in q.preds] + list(q.meas)
    for d in q.dims:
        fk, key = ssb.DIM_KEYS[d.name]
        cols += [fk, key] + [c for c, _ in d.sels] + ([d.attr] if d.attr else [])
    return list(dict.fromkeys(cols))


def early_decode(t: dict, q: ssb.Query, ops=B) -> tuple:
    """The early strategy's first step: every column the query reads checked and decoded
    (checkAndDecodeAN; a column with corrupt codes comes back shorter, as the reference's scalar
    operator drops them, encdecAN_scalar.tcc:200-207).  Returns (tables, errors)."""
    E = ssb.Errors()
    out = {}
    for name in columns_read(q):
        b, e = ops.checkAndDecodeAN(t[name])
        E.add(f"decode_{name}", e)
        out[name] = b
    return out, E


def early(t: dict, q: ssb.Query, ops=B) -> dict:
    """The plain chain on the early-decoded columns; with corrupt codes the columns' lengths differ
    and a size check of the chain can raise, as the reference's does."""
    dec, E = early_decode(t, q, ops)
    r = ssb.chain(dec, q, ops)
    r["errors"] = E
    return r


def late_tables(t: dict) -> dict:
    """The AN codes as the plain operators see them: unchecked, raw codes (A^-1 = 1), thresholds
    still encoded with the column's A."""
    out = {}
    for name, b in t.items():
        m = b.tail.meta
        out[name] = type(b)(b.head, replace(b.tail, meta=an.unchecked_meta(m.type, m.a)))
    return out


class _LateOps:
    """The operator module with the late strategy's check-and-decode of fetched attributes and
    measures (the chains call ops.late_decode right before grouping and aggregation)."""

    def __init__(self, ops):
        self._ops = ops

    def __getattr__(self, name):
        return getattr(self._ops, name)

    def late_decode(self, bat, E: ssb.Errors, name: str):
        t = bat.tail
        real = an.an_meta(an.checked_of(t.meta.type), t.meta.a)
        d, e = self._ops.checkAndDecodeAN(type(bat)(bat.head, replace(t, meta=real)))
        E.add(f"decode_{name}", e)
        return d


def late(t: dict, q: ssb.Query, ops=B) -> dict:
    return ssb.chain(late_tables(t), q, _LateOps(ops))


class _ReencOps:
    """The operator module with the re-encoding of ssbm-q*_continuous_reenc.cpp: every join and
    fetch re-encodes its outputs (matchjoinAN.tcc:394-432, hashjoinAN.tcc:561-585) with the next A
    of the code width's parameter table, counting down from the one below the default (the drivers
    take std::get<14>, <13>, ... of v2_resoid_t::As / v2_resint_t::As, <6>, <5>, ... of
    v2_restiny_t::As), so consecutive operators' outputs carry different As.  selectAN's
    re-encoding arguments are ignored by the reference (selectAN.tcc:51-70) and are not passed."""

    TABLES = {2: an.RESTINY_AS, 4: an.RESSHORT_AS, 8: an.RESINT_AS}

    def __init__(self, ops):
        self._ops = ops
        self._next: dict = {}

    def __getattr__(self, name):
        return getattr(self._ops, name)

    def _a(self, side) -> tuple:
        w = side.meta.width if side.meta.type.encoded else 8  # void / plain OIDs become 8-byte OID codes
        table = self.TABLES[w]
        i = self._next.get(w, len(table) - 2)
        self._next[w] = i - 1 if i > 1 else len(table) - 2
        return table[i], an.inverse(table[i], 8 * w)

    def matchjoinAN(self, b1, b2, reenc_h=None, reenc_t=None, **kw):
        return self._ops.matchjoinAN(b1, b2, self._a(b1.head), self._a(b2.tail), **kw)

    def hashjoinAN(self, b1, b2, reenc_h=None, reenc_t=None, **kw):
        return self._ops.hashjoinAN(b1, b2, self._a(b1.head), self._a(b2.tail), **kw)

    def fetchjoinAN(self, oids, col, reenc=None, **kw):
        return self._ops.fetchjoinAN(oids, col, self._a(col.tail), **kw)


def continuous_reenc(t: dict, q: ssb.Query, ops=B) -> dict:
    """The hardened chain with re-encoding at every join and fetch (its results and error vectors
    are the continuous chain's: decoding and error positions do not depend on the As)."""
    return ssb.chain(t, q, _ReencOps(ops))


STRATEGIES = ("continuous", "continuous_reenc", "normal", "early", "late", "dmr", "dmr_concurrent", "tmr")
