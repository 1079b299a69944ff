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
