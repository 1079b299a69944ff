"""Host-side pieces of the served form: the seeded query stream and the latency summary."""
import numpy as np

from ahead_amd import serve


def test_random_queries_are_seeded_and_poisson():
    a = serve.random_queries(2000, 500.0, 1)
    b = serve.random_queries(2000, 500.0, 1)
    assert [(q.shape, q.params, q.arrival) for q in a] == [(q.shape, q.params, q.arrival) for q in b]
    gaps = np.diff([q.arrival for q in a])
    assert abs(gaps.mean() - 1 / 500.0) < 0.1 / 500.0   # exponential gaps, mean 1/rate
    assert {q.shape for q in a} == {"P1", "P2", "P3", "P4"}
    for q in a:
        if q.shape == "P1":
            assert 0 <= q.params["disc_lo"] <= q.params["disc_hi"] <= 10
        if q.shape == "P4":
            assert q.params["mfgr"][0] != q.params["mfgr"][1]


def test_summary_percentiles():
    qs = serve.random_queries(100, 0, 2)
    for i, q in enumerate(qs):
        q.arrival, q.finish = 0.0, (i + 1) / 1000.0
    s = serve.summarize(qs)
    assert s["queries"] == 100 and abs(s["p50_ms"] - 50.5) < 1e-6 and s["p99_ms"] > 99


def test_admission_is_first_come_first_served():
    """A query that does not fit yet holds back the later ones even when they would fit: with a
    budget of 10 and 6 in use, ticket 1 (needs 8) waits, and ticket 2 (needs 1) waits behind it."""
    import threading
    import time

    adm = serve.Admission(10.0)
    order = []
    t0, t1, t2 = adm.ticket(), adm.ticket(), adm.ticket()
    adm.acquire(t0, 6.0)

    def take(t, need):
        adm.acquire(t, need)
        order.append(t)

    th = [threading.Thread(target=take, args=(t2, 1.0)), threading.Thread(target=take, args=(t1, 8.0))]
    for x in th:
        x.start()
    time.sleep(0.2)
    assert order == [] and adm.in_use == 6.0
    adm.release(6.0)
    for x in th:
        x.join(5)
    assert order == [t1, t2] and adm.in_use == 9.0
    adm.release(8.0)
    adm.release(1.0)
    assert adm.in_use == 0.0


def test_admission_caps_a_need_at_the_budget():
    adm = serve.Admission(4.0)
    t = adm.ticket()
    adm.acquire(t, 100.0)   # larger than the budget: runs alone
    assert adm.in_use == 4.0
    adm.release(100.0)
    assert adm.in_use == 0.0


def test_form_batch_takes_the_heads_shape_in_arrival_order():
    from collections import deque

    qs = [serve.Query(i, s, {}) for i, s in enumerate(["P2", "P1", "P2", "P3", "P2", "P1", "P2"])]
    q = deque(qs)
    b = serve.form_batch(q, 3)
    assert [x.qid for x in b] == [0, 2, 4]            # the head and the next two of its shape
    assert [x.qid for x in q] == [1, 3, 5, 6]         # the rest, in arrival order
    b = serve.form_batch(q, 8)
    assert [x.qid for x in b] == [1, 5] and [x.qid for x in q] == [3, 6]
    assert [x.qid for x in serve.form_batch(q, 1)] == [3] and [x.qid for x in q] == [6]
    assert [x.qid for x in serve.form_batch(q, 8)] == [6] and not q


def test_batched_server_serves_every_query_once_in_batches_of_one_shape():
    """The slots' loop on the host (a stand-in for the device work): every query is served exactly
    once, and each batch holds one shape, at most MAX_BATCH queries, in arrival order."""
    import threading

    srv = serve.Server({}, 1000, workers=3, mode="batched")
    srv.need = lambda shape, n=1: 1.0
    batches, lock = [], threading.Lock()

    def execute(batch):
        with lock:
            batches.append([q.qid for q in batch])
        return [{"rows": q.qid} for q in batch]

    srv._run_batch = execute
    try:
        qs = srv.run(serve.random_queries(200, 0, 8), realtime=False)
    finally:
        srv.close()
    seen = sorted(i for b in batches for i in b)
    assert seen == list(range(200))
    shape = {q.qid: q.shape for q in qs}
    for b in batches:
        assert len(b) <= serve.MAX_BATCH and len({shape[i] for i in b}) == 1 and b == sorted(b)
    assert all(q.result["rows"] == q.qid and q.finish >= q.start >= q.arrival for q in qs)

