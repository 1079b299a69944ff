"""Served star-schema queries: a device-resident schema, Poisson arrivals, continuous batching.

The tables of ahead_amd.ssb stay in HBM for the server's lifetime.  Queries (the four pipeline
shapes with seeded random predicate constants) arrive on a Poisson clock.  `workers` GPU slots
each run one query's operator chain on their own HIP stream, so the operators of concurrent
queries overlap on the device.  A query is admitted as soon as a slot is free and the HBM its
intermediates need (an upper bound per shape and fact row) fits the budget beside those running.
Queries that arrive meanwhile wait in arrival order.  Latency runs from arrival to the query's
last result on the host.  With workers=1 the same loop is the unbatched path: one query on the
GPU at a time.
"""
from __future__ import annotations

import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import ssb

# upper bound of the device bytes a query's intermediates hold at once, per fact row: the peaks
# measured at 600M fact rows (P1 14.0, P2 17.6, P3 17.6, P4 19.8) plus about 15%
HBM_PER_ROW = {"P1": 16, "P2": 20, "P3": 20, "P4": 23}


@dataclass
class Query:
    qid: int
    shape: str
    params: dict
    arrival: float = 0.0
    start: float = 0.0
    finish: float = 0.0
    result: dict = field(default_factory=dict)

    @property
    def latency(self) -> float:
        return self.finish - self.arrival


def random_queries(n: int, rate: float, seed: int) -> list[Query]:
    """n queries with exponential inter-arrival gaps of mean 1/rate, shapes uniform over P1-P4,
    predicate constants uniform over the schema's value ranges (seeded)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n)) if rate > 0 else np.zeros(n)
    out = []
    for i in range(n):
        shape = ["P1", "P2", "P3", "P4"][int(rng.integers(0, 4))]
        if shape == "P1":
            lo = int(rng.integers(0, 9))
            p = {"year": int(rng.integers(1992, 1999)), "disc_lo": lo, "disc_hi": lo + int(rng.integers(0, 3)),
                 "qty_lt": int(rng.integers(10, 51))}
        elif shape == "P2":
            p = {"category": int(rng.integers(0, 25)), "s_region": int(rng.integers(0, 5))}
        elif shape == "P3":
            ylo = int(rng.integers(1992, 1996))
            p = {"c_region": int(rng.integers(0, 5)), "s_region": int(rng.integers(0, 5)), "year_lo": ylo,
                 "year_hi": ylo + int(rng.integers(0, 4))}
        else:
            m = rng.choice(5, 2, replace=False)
            p = {"c_region": int(rng.integers(0, 5)), "s_region": int(rng.integers(0, 5)),
                 "mfgr": (int(m[0]), int(m[1]))}
        out.append(Query(i, shape, p, float(t[i])))
    return out


class Server:
    """A long-lived query server: its worker threads, their streams and their workspaces persist
    across `run` calls (close() ends them), so device memory cached for one stream's queries is
    reused by that stream's later queries."""

    def __init__(self, tables: dict, n_fact: int, workers: int = 4, hbm_budget: float = 64e9):
        self.tables = tables
        self.n_fact = n_fact
        self.workers = workers
        self.budget = hbm_budget
        self._in_use = 0.0
        self._cv = threading.Condition()
        self._local = threading.local()
        self._ex = ThreadPoolExecutor(max_workers=workers)

    def close(self) -> None:
        self._ex.shutdown(wait=True)

    def __enter__(self) -> "Server":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def _stream(self):
        import torch

        s = getattr(self._local, "stream", None)
        if s is None:
            s = self._local.stream = torch.cuda.Stream()
        return s

    def _need(self, q: Query) -> float:
        return min(self.budget, HBM_PER_ROW[q.shape] * self.n_fact)

    def _run_one(self, q: Query, t0: float) -> Query:
        import torch

        need = self._need(q)
        with self._cv:
            while self._in_use + need > self.budget:
                self._cv.wait()
            self._in_use += need
        try:
            q.start = time.perf_counter() - t0
            s = self._stream()
            with torch.cuda.stream(s):
                q.result = ssb.PIPELINES[q.shape](self.tables, **q.params)
            s.synchronize()
            q.finish = time.perf_counter() - t0
        finally:
            with self._cv:
                self._in_use -= need
                self._cv.notify_all()
        return q

    def run(self, queries: list[Query], realtime: bool = True) -> list[Query]:
        """Replays the arrivals (in real time, or all at once) and returns the queries, completed."""
        t0 = time.perf_counter()
        futs = []
        for q in queries:
            if realtime:
                delay = q.arrival - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
            else:
                q.arrival = time.perf_counter() - t0
            futs.append(self._ex.submit(self._run_one, q, t0))
        for f in futs:
            f.result()
        return queries


def summarize(queries: list[Query]) -> dict:
    lat = np.array([q.latency for q in queries]) * 1e3
    span = max(q.finish for q in queries) - min(q.arrival for q in queries)
    return {"queries": len(queries), "qps": len(queries) / span if span > 0 else 0.0,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean())}


def same_result(a: dict, b: dict) -> bool:
    """Bit-exact equality of two pipeline results, error vectors included."""
    if a["rows"] != b["rows"] or a.get("result") != b.get("result") or a.get("groups") != b.get("groups"):
        return False
    if "sums" in a and not (np.array_equal(a["sums"], b["sums"]) and
                            all(np.array_equal(x, y) for x, y in zip(a["keys"], b["keys"]))):
        return False
    ea, eb = a["errors"].items, b["errors"].items
    return [k for k, _ in ea] == [k for k, _ in eb] and all(np.array_equal(x, y) for (_, x), (_, y) in zip(ea, eb))
