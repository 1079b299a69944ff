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
