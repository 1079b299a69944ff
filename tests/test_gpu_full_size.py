"""Configs 3-5 at their full single-GPU sizes.

* Config 3 (600 M fact rows): the fused form of Q1.1 and Q4.1 on fault-free data against the host
  textbook GROUP BY / SUM over all rows (closed form I2), of Q2.1 and Q3.1 against their operator
  chains; with the schema's flips, every selection error vector against the generator's flip
  positions (closed form I1: every single-bit flip is detected).
* Config 4: the bench's 512 lognormal instances (about 10^9 rows) in one persistent launch, after
  three replays of a captured hipGraph, each instance against the oracle's select + decode + sum.
* Config 5: queries over the resident 600 M-row schema alone (one slot) and in the mix (four
  slots) give the same results, which equal the operator chains'."""
import numpy as np
import pytest

from ahead_amd import an, capi, serve, ssb

pytestmark = pytest.mark.gpu

AOE = np.uint64(an.A_OID_ERR)


@pytest.fixture(scope="module")
def clean600():
    spec = ssb.SsbSpec(rate=0.0)
    print(f"generating the fault-free schema, {spec.n_fact} fact rows", flush=True)
    return spec, ssb.device_tables(spec)


def same_result(got, ref):
    assert got["rows"] == ref["rows"]
    if "result" in ref:
        assert got["result"] == ref["result"]
    else:
        assert got["groups"] == ref["groups"]
        assert np.array_equal(got["sums"], ref["sums"])
        for a, b in zip(got["keys"], ref["keys"]):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("shape", ["P1", "P4"])
def test_config3_fused_against_textbook(clean600, shape):
    spec, t = clean600
    q = ssb.SHAPES[shape]()
    print(f"config 3 {shape}: fused vs textbook over {spec.n_fact} rows", flush=True)
    got = ssb.fused(t, q)
    assert got["status"] == 0 and got["errors"].total() == 0
    same_result(got, ssb.textbook(spec, q, workers=16))


@pytest.mark.parametrize("shape", ["P2", "P3"])
def test_config3_fused_against_chain(clean600, shape):
    spec, t = clean600
    q = ssb.SHAPES[shape]()
    print(f"config 3 {shape}: fused vs chain over {spec.n_fact} rows", flush=True)
    got = ssb.fused(t, q)
    assert got["status"] == 0
    same_result(got, ssb.chain(t, q))


def test_config3_selection_errors_are_the_flip_positions():
    spec = ssb.SsbSpec(rate=1e-6)
    print(f"config 3 with flips at {spec.rate}: selection error vectors", flush=True)
    t = ssb.device_tables(spec)
    seen = set()
    for shape in ("P1", "P2", "P3", "P4"):
        r = ssb.fused(t, ssb.SHAPES[shape]())
        assert r["status"] == 0
        for k, v in r["errors"].items:
            assert k.startswith("select_"), k
            col = k[len("select_"):-len(".e1")]
            expect = ssb.flip_positions(spec, col, workers=16) * AOE
            assert expect.size > 0 and np.array_equal(v, expect), col
            seen.add(col)
    assert "lo_quantity" in seen and "c_region" in seen  # 200 k suppliers at 1e-6 hold no flip


def test_config4_batch_after_graph_replays_against_oracle():
    import torch

    import oracle_lib as O
    from ahead_amd import workloads as W

    spec = W.BatchSpec()
    print(f"config 4: {spec.count} instances", flush=True)
    jobs = W.batch_device(spec)
    batch = capi.SelectBatch(jobs)
    batch.launch()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(g, stream=side):
            batch.launch()
    for _ in range(3):
        g.replay()
        torch.cuda.synchronize()
    res = batch.results()
    plan = W.batch_plan(spec)
    rows = 0
    for i, (p, (oids, err, s)) in enumerate(zip(plan, res)):
        codes, meta = W.batch_host_column(spec, i)
        ref = O.select_an_simd(O.Col(codes, meta), capi.GE, p["lo"] * p["a"], capi.LE, p["hi"] * p["a"],
                               capi.COMB_AND, threads=16)
        assert np.array_equal(oids, ref[0]) and np.array_equal(err, ref[1]) and s == ref[2], i
        rows += p["n"]
    assert rows > 5 * 10**8
    batch.close()


def test_config5_alone_equals_mixed_at_600M(clean600):
    spec, t = clean600
    qs = serve.random_queries(8, 0, 71)
    print("config 5: 8 queries alone and in the mix", flush=True)
    with serve.Server(t, spec.n_fact, workers=1) as s1:
        alone = s1.run(serve.random_queries(8, 0, 71), realtime=False)
    with serve.Server(t, spec.n_fact, workers=4) as s4:
        mixed = s4.run(serve.random_queries(8, 0, 71), realtime=False)
    for q, a, m in zip(qs, alone, mixed):
        assert a.result["status"] == 0
        assert serve.same_result(a.result, m.result), (q.shape, q.params)
        assert serve.same_result(a.result, ssb.chain(t, q.spec())), (q.shape, q.params)


def test_config5_batched_native_loop_equals_alone_at_600M(clean600):
    """The C++ request loop over the resident 600M-row schema: 48 queries arriving at once are
    served in same-shape batches of up to 24; every result equals the query run alone."""
    spec, t = clean600
    print("config 5: 48 queries through the C++ request loop vs alone", flush=True)
    with serve.Server(t, spec.n_fact, workers=2, mode="native") as srv:
        got = srv.run(serve.random_queries(48, 0, 72), realtime=False)
    for q in got:
        ref = ssb.fused(t, q.spec())
        assert q.result["status"] == 0 and ref["status"] == 0, (q.shape, q.params)
        assert serve.same_result(q.result, ref), (q.shape, q.params)
