"""GPU parity of the batched selectAN (many instances, mixed code widths, one persistent launch)
against the oracle, per job: OIDs, error positions and decoded sums.  Also: repeated launches of
one batch, several tilings, and replays of a hipGraph capture of the launch."""
import numpy as np
import pytest

import oracle_lib as O
from ahead_amd import an, capi, datagen

pytestmark = pytest.mark.gpu

TYPES = [(an.RESTINY, 233, 0xFF), (an.RESSHORT, 233, 0xFFFF), (an.RESSHORT, 63877, 0xFFFF),
         (an.RESINT, 65117, 0xFFFFFFFF)]


def make_jobs(count, seed, max_rows):
    rng = np.random.default_rng(seed)
    jobs, host = [], []
    sizes = [0, 1, 2047, 2049, 8193, 65_537] + [int(x) for x in rng.integers(1, max_rows, count)]
    for i, n in enumerate(sizes[:count]):
        t, a, vmax = TYPES[i % len(TYPES)]
        hi = min(vmax, 1000)
        vals = datagen.uniform(n, seed + i, 1, 0, hi)
        codes = datagen.encode(vals, a, t.width)
        datagen.inject_flips(codes, seed + i, 2, datagen.flip_threshold(1e-3))
        meta = an.an_meta(t, a)
        lo, up = sorted(int(x) for x in rng.integers(0, hi + 1, 2))
        kind = i % 3
        if kind == 0:
            q = dict(op1=capi.GE, th1=lo * a, op2=capi.LE, th2=up * a, comb=capi.COMB_AND)
        elif kind == 1:
            q = dict(op1=capi.LT, th1=lo * a, op2=capi.GT, th2=up * a, comb=capi.COMB_OR)
        else:
            q = dict(op1=capi.LT, th1=up * a, comb=capi.COMB_NONE)
        q["col"] = capi.Column.from_numpy(codes, meta)
        q["err_cap"] = 4
        jobs.append(q)
        host.append((codes, meta, q))
    return jobs, host


def check(batch, host):
    res = batch.results()
    for (codes, meta, q), (oids, err, total) in zip(host, res):
        ro, re, rs = O.select_an(O.Col(codes, meta), q["op1"], q["th1"], q.get("op2", capi.LT), q.get("th2", 0),
                                 q.get("comb", capi.COMB_NONE))
        assert np.array_equal(oids, ro)
        assert np.array_equal(err, re)
        assert total == rs


@pytest.mark.parametrize("max_subtiles,blocks", [(16, 0), (1, 64), (4, 300)])
def test_select_batch_parity(max_subtiles, blocks):
    jobs, host = make_jobs(40, 7 + max_subtiles, 400_000)
    b = capi.SelectBatch(jobs, max_subtiles=max_subtiles, blocks=blocks)
    b.launch()
    check(b, host)
    # a second launch of the same batch (fresh epoch, reset tickets) gives the same results
    b.counts.t.zero_()
    b.sums.t.zero_()
    b.launch()
    check(b, host)
    b.close()


def test_select_batch_graph_replays():
    import torch

    jobs, host = make_jobs(24, 99, 200_000)
    b = capi.SelectBatch(jobs, max_subtiles=8)
    b.launch()
    b.results()          # grows error buffers if needed, before capture
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            b.launch()
    torch.cuda.synchronize()
    for rep in range(3):
        b.counts.t.zero_()
        b.sums.t.zero_()
        for o in b.oids:
            o.t.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        check(b, host)
    b.close()
