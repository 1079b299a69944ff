"""Served star-schema queries: a device-resident schema, Poisson arrivals, continuous batching.

The tables of ahead_amd.ssb stay in HBM for the server's lifetime.  Queries (the four query shapes
Q1.1 / Q2.1 / Q3.1 / Q4.1 with seeded random predicate constants) arrive on a Poisson clock.
`workers` GPU slots each run one query at a time on their own HIP stream and workspace, so
concurrent queries overlap on the device: one query's dimension selections, planning and result
copies run beside another's fact scan.  Admission is first come, first served: the query at the
head of the arrival queue starts as soon as a slot is free and the device memory it needs fits the
budget beside the queries running; later arrivals wait behind it even if they would fit.
Latency runs from arrival to the query's results on the host.

Three execution forms:
  fused    each query is a fused star query (ssb.fused): the dimension selections, then one pass
           over the fact rows; it needs its workspace (dimension maps, group tables, as
           ahead_hip_star_query_ws_bytes reports) and its output buffers;
  batched  a free slot takes the query at the head of the queue together with the queued queries
           of the same shape behind it (at most MAX_BATCH), and runs them as one batch
           (ssb.fused_batch): queries of one shape share one pass over the fact rows; each result
           is the one the query gets alone; a batch needs its queries' workspaces and one more for
           the shared scan's combined maps;
  native   the batched form's loop in C++ (ahead_serve_*, capi.NativeServer): C++ slot threads
           with their own streams, workspaces and arenas form the same-shape batches, admit them
           first come, first served by the bytes the library computes for them, run the dimension
           selections and the batch, and hand back each query's outputs after one copy per batch;
           Python only submits arrivals and reads results;
  chain    each query is its operator chain (ssb.chain), one launch per operator, every operator
           waiting for its counts: the plain path; it needs its materialised intermediates (a bound
           per fact row measured at 600 M rows).
"""
from __future__ import annotations

import threading
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import ssb

# upper bound of the device bytes an operator chain's intermediates hold at once, per fact row: the
# peaks measured at 600M fact rows (P1 14.0, P2 17.6, P3 17.6, P4 19.8) plus about 15%
HBM_PER_ROW = {"P1": 16, "P2": 20, "P3": 20, "P4": 23}
# the fused form's output buffers (4096 groups x 6 arrays, 16 error vectors of 1024 positions)
FUSED_OUT_BYTES = 4096 * 8 * 6 + 16 * 1024 * 8 + 4096
# queries per batch in the batched form (ahead_hip_star_query_batch's same-shape pass,
# AHEAD_SQ_MAX_SHAPE_PASS)
MAX_BATCH = 32


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

    def spec(self) -> ssb.Query:
        return ssb.SHAPES[self.shape](**self.params)


# a query of each shape with its default constants (the workspace depends on the shape only)
SHAPE_DEFAULTS = {k: (lambda k=k: ssb.SHAPES[k]()) for k in ("P1", "P2", "P3", "P4")}


def random_queries(n: int, rate: float, seed: int) -> list[Query]:
    """n queries with exponential inter-arrival gaps of mean 1/rate (0: all at once), shapes uniform
    over P1-P4, predicate constants uniform over the schema's value ranges (seeded)."""
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


class Admission:
    """First-come-first-served admission of device memory: callers take tickets in arrival order;
    the holder of the oldest unserved ticket is admitted once its bytes fit the budget beside the
    bytes in use, and no later ticket is admitted before it (no overtaking, so a large query is not
    starved by a stream of small ones)."""

    def __init__(self, budget: float):
        self.budget = budget
        self.in_use = 0.0
        self._cv = threading.Condition()
        self._next = 0     # next ticket handed out
        self._serving = 0  # oldest ticket not yet admitted

    def ticket(self) -> int:
        with self._cv:
            t = self._next
            self._next += 1
            return t

    def acquire(self, ticket: int, need: float) -> None:
        need = min(need, self.budget)
        with self._cv:
            while ticket != self._serving or self.in_use + need > self.budget:
                self._cv.wait()
            self._serving += 1
            self.in_use += need
            self._cv.notify_all()

    def release(self, need: float) -> None:
        need = min(need, self.budget)
        with self._cv:
            self.in_use -= need
            self._cv.notify_all()


def form_batch(queue: deque, max_batch: int) -> list:
    """Takes the query at the head of the queue and the queued queries of its shape behind it, at
    most max_batch in all, in arrival order; the other queries stay queued in their order."""
    head = queue.popleft()
    batch, rest = [head], deque()
    while queue and len(batch) < max_batch:
        q = queue.popleft()
        (batch if q.shape == head.shape else rest).append(q)
    rest.extend(queue)
    queue.clear()
    queue.extend(rest)
    return batch


class Server:
    """A long-lived query server: its worker threads, their streams and their workspaces persist
    across `run` calls (close() ends them), so device memory cached for one stream's queries is
    reused by that stream's later queries."""

    def __init__(self, tables: dict, n_fact: int, workers: int = 4, hbm_budget: float = 64e9, mode: str = "fused"):
        assert mode in ("fused", "batched", "native", "chain")
        self.tables = tables
        self.n_fact = n_fact
        self.workers = workers
        self.mode = mode
        self.max_batch = MAX_BATCH if mode == "batched" else 1
        self.admission = Admission(hbm_budget)
        self._need_cache: dict = {}
        self._local = threading.local()
        self._ex = ThreadPoolExecutor(max_workers=workers)
        self._queue: deque = deque()
        self._qcv = threading.Condition()
        self._feeding = False
        self._native = None
        if mode == "native":
            from . import capi

            self._native = capi.NativeServer(slots=workers, max_batch=MAX_BATCH, hbm_budget=int(hbm_budget))

    def close(self) -> None:
        self._ex.shutdown(wait=True)
        if self._native is not None:
            self._native.close()
            self._native = None

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

    def need(self, shape: str, n: int = 1) -> float:
        """Device bytes a query of this shape (a batch of n of them) holds while it runs: for the
        fused forms the workspace (as the library computes it from the dimension sizes; a batch
        one more for the shared scan's combined maps) and outputs, for the chain the bound of its
        intermediates."""
        if self.mode == "chain":
            return HBM_PER_ROW[shape] * self.n_fact
        if shape not in self._need_cache:
            spec = SHAPE_DEFAULTS[shape]()
            self._need_cache[shape] = ssb.fused_ws_bytes(self.tables, spec) + FUSED_OUT_BYTES
        return self._need_cache[shape] * (n + 1 if n > 1 else 1)

    def _execute(self, batch: list) -> list:
        specs = [q.spec() for q in batch]
        if self.mode == "chain":
            return [ssb.chain(self.tables, specs[0])]
        if len(batch) == 1:
            return [ssb.fused(self.tables, specs[0])]
        return ssb.fused_batch(self.tables, specs)

    def _run_batch(self, batch: list) -> list:
        """Runs the batch on this slot's stream; returns once its results are on the host."""
        import torch

        s = self._stream()
        with torch.cuda.stream(s):
            results = self._execute(batch)
        s.synchronize()
        return results

    def _worker(self, t0: float) -> None:
        """A slot: takes batches from the queue in arrival order until the feed ends and the queue
        is empty; each batch is admitted first come, first served, and runs on the slot's stream."""
        while True:
            with self._qcv:
                while not self._queue and self._feeding:
                    self._qcv.wait()
                if not self._queue:
                    return
                batch = form_batch(self._queue, self.max_batch)
                # nothing that can fail lies between taking a ticket and serving it (a ticket
                # never served would hold every later one)
                need = self.need(batch[0].shape, len(batch))
                ticket = self.admission.ticket()
            self.admission.acquire(ticket, need)
            try:
                start = time.perf_counter() - t0
                results = self._run_batch(batch)
                finish = time.perf_counter() - t0
                for q, r in zip(batch, results):
                    q.start, q.finish, q.result = start, finish, r
            finally:
                self.admission.release(need)

    def _run_native(self, queries: list[Query], realtime: bool) -> list[Query]:
        specs = [q.spec() for q in queries]
        reqs = [ssb.serve_query(self.tables, sp) for sp in specs]  # built before the clock starts
        srv = self._native
        c0, t0 = srv.clock(), time.perf_counter()
        tickets = []
        for q, rq in zip(queries, reqs):
            if realtime:
                delay = q.arrival - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
            tickets.append(srv.submit(rq))
        for q, sp, tk in zip(queries, specs, tickets):
            r = srv.wait(tk, len(sp.dims))
            q.arrival, q.start, q.finish = r["t_arrival"] - c0, r["t_start"] - c0, r["t_finish"] - c0
            q.result = ssb.serve_result(self.tables, sp, r)
            if r["status"] != 0:  # the chain ran instead, here
                q.finish = time.perf_counter() - t0
        return queries

    def run(self, queries: list[Query], realtime: bool = True) -> list[Query]:
        """Replays the arrivals (in real time, or all at once) into the queue the slots serve, and
        returns the queries, completed."""
        if self._native is not None:
            return self._run_native(queries, realtime)
        for shape in {q.shape for q in queries}:
            self.need(shape)  # computed once per shape, before the clock starts
        t0 = time.perf_counter()
        self._feeding = True
        futs = [self._ex.submit(self._worker, t0) for _ in range(self.workers)]
        try:
            for q in queries:
                if realtime:
                    delay = q.arrival - (time.perf_counter() - t0)
                    if delay > 0:
                        time.sleep(delay)
                else:
                    q.arrival = time.perf_counter() - t0
                with self._qcv:
                    self._queue.append(q)
                    self._qcv.notify()
        finally:
            with self._qcv:
                self._feeding = False
                self._qcv.notify_all()
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
    """Bit-exact equality of two query results, error vectors included."""
    if a["rows"] != b["rows"] or a.get("result") != b.get("result") or a.get("groups") != b.get("groups"):
        return False
    if "sums" in a and not (np.array_equal(a["sums"], b["sums"]) and
                            all(np.array_equal(x, y) for x, y in zip(a["keys"], b["keys"]))):
        return False
    ea, eb = dict(a["errors"].items), dict(b["errors"].items)
    return sorted(ea) == sorted(eb) and all(np.array_equal(ea[k], eb[k]) for k in eb)
