"""Served queries (ahead_amd.serve), in the fused form, in same-shape batches (Python slots and the C++
request loop) and as operator chains:
every query's result, error vectors included, is the same when it runs alone (one query or batch on
the GPU at a time) and in the mix (four slots overlapping on four streams, all arriving at once),
and both equal the oracle's operator chain for the query; the admission budget serialises queries
(batches) that do not fit together."""
import pytest

import oracle_bat as OB
from ahead_amd import serve, ssb

pytestmark = pytest.mark.gpu

SMALL = dict(n_fact=250_000, n_date=2556, n_cust=3_000, n_supp=200, n_part=1_400)


@pytest.mark.parametrize("mode", ["fused", "batched", "native", "chain"])
def test_results_alone_equal_results_in_the_mix(mode):
    spec = ssb.SsbSpec(rate=2e-3, seed=21, **SMALL)
    tables = ssb.device_tables(spec)
    print(f"serve {mode}: 24 queries, 1 then 4 slots", flush=True)
    with serve.Server(tables, spec.n_fact, workers=1, mode=mode) as srv:
        alone = srv.run(serve.random_queries(24, 0, 5), realtime=False)
    with serve.Server(tables, spec.n_fact, workers=4, mode=mode) as srv:
        mixed = srv.run(serve.random_queries(24, 0, 5), realtime=False)
        again = srv.run(serve.random_queries(24, 0, 5), realtime=False)  # the same threads, warm
    host = OB.tables(spec)
    shapes = set()
    for a, m in zip(alone, mixed):
        assert (a.qid, a.shape, a.params) == (m.qid, m.shape, m.params)
        ref = ssb.PIPELINES[a.shape](host, ops=OB, **a.params)
        print(a.qid, a.shape, a.params, "rows", a.result["rows"], m.result["rows"], ref["rows"], flush=True)
        assert serve.same_result(a.result, ref), ("alone", a.qid, a.shape, a.params)
        assert serve.same_result(m.result, ref), ("mixed", a.qid, a.shape, a.params)
        assert serve.same_result(again[a.qid].result, ref), ("mixed again", a.qid, a.shape, a.params)
        if mode != "chain":
            assert a.result["status"] == 0 and m.result["status"] == 0
        shapes.add(a.shape)
    assert shapes == {"P1", "P2", "P3", "P4"}
    s = serve.summarize(mixed)
    assert s["queries"] == 24 and s["p99_ms"] >= s["p50_ms"] > 0


@pytest.mark.parametrize("mode", ["fused", "batched", "chain"])
def test_admission_budget_serializes_when_memory_is_short(mode):
    spec = ssb.SsbSpec(rate=0.0, seed=22, **SMALL)
    tables = ssb.device_tables(spec)
    print(f"serve {mode}: admission budget", flush=True)
    with serve.Server(tables, spec.n_fact, workers=4, mode=mode) as srv:
        needs = {s: srv.need(s) for s in ("P1", "P2", "P3", "P4")}
        assert all(n > 0 for n in needs.values())
        # a budget below any two needs together (a need above it is capped at it): no overlap
        srv.admission.budget = 2 * min(needs.values()) - 1
        qs = srv.run(serve.random_queries(8, 0, 6), realtime=False)
    heads = {}  # start of each batch (one query each, but in the batched form) -> its first query
    for q in qs:
        heads.setdefault(q.start, q.qid)
    assert list(heads) == sorted(heads)   # batches admitted in the arrival order of their first queries
    spans = sorted({(q.start, q.finish) for q in qs})
    for (s0, f0), (s1, f1) in zip(spans, spans[1:]):
        assert s1 >= f0 - 1e-4


def test_native_loop_admission_serializes_and_keeps_arrival_order():
    """The C++ request loop with a budget below any batch's bytes: batches run one at a time, in the
    arrival order of their first queries, and every result still equals the oracle's chain."""
    spec = ssb.SsbSpec(rate=1e-3, seed=23, **SMALL)
    tables = ssb.device_tables(spec)
    print("serve native: admission budget of one byte", flush=True)
    with serve.Server(tables, spec.n_fact, workers=4, mode="native", hbm_budget=1) as srv:
        qs = srv.run(serve.random_queries(12, 0, 7), realtime=False)
    heads = {}
    for q in qs:
        heads.setdefault(q.start, q.qid)
    assert list(heads) == sorted(heads)
    spans = sorted({(q.start, q.finish) for q in qs})
    for (s0, f0), (s1, f1) in zip(spans, spans[1:]):
        assert s1 >= f0 - 1e-4
    host = OB.tables(spec)
    for q in qs:
        assert serve.same_result(q.result, ssb.PIPELINES[q.shape](host, ops=OB, **q.params)), (q.qid, q.shape)
