#!/usr/bin/env python3
"""Benchmark of the AN-hardened operator path on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is one pass of the hardened hot path over one batch of synthetic input already resident in
HBM.  Workloads:
  star1       : hardened hash join + group-by: 60,000,000 fact rows with a foreign key
                uniform over a dense dimension of 200,000 keys and a measure uniform in
                [90000, 10500000]; the dimension has two 8-bit attributes with 25 and 5 distinct
                values coded to uint16 (A=29); keys are 32-bit values coded to uint64 (A=63877),
                the measure to uint64 (A=65117); 1e-7 seeded single-bit flips in every column;
                join, group by both attributes (125 groups), int64 sum -- one fused launch chain.
  select_sum  : range select + check-and-decode + sum over 6,000,000 16-bit values in [1,50]
                AN-coded to uint32 with A=233, 1e-6 seeded single-bit flips, predicate 1<=v<=24
                on encoded bounds; outputs OIDs, decoded sum and sorted error positions.
  batch       : 512 independent select + check-and-decode + sum instances, lengths lognormal
                (median 707k rows, sigma 1.42, clipped to [1e4, 5e7]), code widths 16/32/64 bits
                (8-bit values A=233; 16-bit values A=233 or 63877; 32-bit values A=65117), range
                predicates of selectivity 1-50%, 1e-6 flips; one persistent launch (autotuned
                tiling and workgroup count, replayed from a hipGraph) against one launch per
                instance.  A step is the whole batch.
  ssb         : (default) the star-schema queries Q1.1, Q2.1, Q3.1 and Q4.1 of the reference's SSB
                drivers over a 600,000,000-row fact table with 4 foreign keys into date (2,556
                rows), customer (3M), supplier (200k) and part (1.4M), integer-coded attributes,
                each run as a fused star query (the dimension selections, then one fact scan).  A
                step runs all four.  Before timing, each query's result is compared with its
                operator chain (one operator launch after another, as the reference's drivers run
                them), and after timing Q1.1 with a host textbook evaluation and every selection
                error vector with the generator's flip positions.
  serve       : the same schema resident in HBM, queries of the four shapes with seeded random
                predicate constants arriving on Poisson clocks; fused star queries on 4 GPU slots
                (own streams and workspaces, first-come-first-served admission by the bytes each
                query needs) against the plain operator chains (one launch per operator) on 1 and
                4 slots; reports queries/s with p50/p99 latency per arrival rate.
  strategies  : the four query shapes under the reference's other strategies (early / late
                check-and-decode, unhardened, DMR and TMR replicas) beside the hardened chain and
                the fused star query.
  hbm_copy    : a plain device-to-device copy of 4 GiB (read 4 GiB, write 4 GiB): the calibration
                kernel for the FETCH_SIZE / WRITE_SIZE counters behind `roofline.traffic`.

The line printed carries the whole-job throughput (`value`), the dominant kernel's roofline
(algorithmic bytes / its average duration from HIP events on the launch stream) and a CPU
baseline: the oracle (a scalar restatement of the reference's operator) timed on a bounded sample
on this host.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="ssb", choices=["star1", "select_sum", "ops", "batch", "ssb", "serve",
                                                                 "hbm_copy", "strategies", "shape_batches"])
    ap.add_argument("--rows", type=int, default=None, help="override the row count (scaling studies)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-check", action="store_true", help="skip the host-side checks of the ssb outputs")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    return ap.parse_args()


def cpu_baseline_select(codes_np, meta, lo, hi, seconds, expect):
    """The all-core vectorised CPU restatement of selectAN + check-and-decode + sum
    (oracle/an_simd.cpp) on this host's usable cores; its outputs are asserted equal to the GPU's."""
    sys.path.insert(0, str(ROOT / "tests"))
    import numpy as np

    import oracle_lib as O

    col = O.Col(codes_np, meta)
    threads = len(os.sched_getaffinity(0))
    reps, t0 = 0, time.perf_counter()
    while True:
        oids, err, s = O.select_an_simd(col, 3, lo, 1, hi, 1, threads=threads)
        reps += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    ok = (oids.size == expect[0] and np.array_equal(err, expect[1]) and s == expect[2])
    if not ok:
        raise RuntimeError("the CPU baseline's select differs from the GPU's")
    cpu = host_cpu()
    isa = {4: "AVX-512 (x86-64-v4)", 3: "AVX2 (x86-64-v3)", 0: "x86-64 baseline"}[O.simd_level()]
    return {
        "value": codes_np.size * reps / el,
        "unit": "tuples/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{codes_np.size}-row column x {reps} repetitions ({el:.1f} s); oracle/an_simd.cpp "
                  f"or_select_an_simd, {isa} build, OpenMP on {threads} threads of {cpu['model']} "
                  f"({cpu['nproc']} logical CPUs); bit-exact vs GPU: {ok}",
    }


def profiled_traffic(name: str):
    """HBM bytes per launch of a workload's dominant kernel from its committed rocprofv3 counter
    summary (profiles/<name>.md, line `traffic_bytes_per_launch: N`), or None."""
    path = ROOT / "profiles" / f"{name}.md"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":", 1)[1])
    return None


def run_select_sum(args):
    import numpy as np
    import torch

    from ahead_amd import an, capi, datagen

    lib = capi.load()
    n = args.rows or 6_000_000
    seed = 42
    meta = an.an_meta(an.RESSHORT, 233)
    codes = capi.DeviceBuffer(4 * n)
    capi.gen_uniform(codes, 4, n, seed, 1, 1, 50, 233)
    capi.inject_flips(codes, 4, n, seed, 2, datagen.flip_threshold(1e-6))
    col = capi.Column(codes, n, meta)
    d = col.desc()
    err_cap = 1 << 16
    oids = capi.DeviceBuffer(8 * n)
    err = capi.DeviceBuffer(8 * err_cap)
    cnt = capi.DeviceBuffer(16)
    dsum = capi.DeviceBuffer(8)
    ws = capi.workspace()
    lo, hi = 1 * 233, 24 * 233
    stream = capi.current_stream()

    def step():
        st = lib.ahead_hip_select_an(ctypes.byref(d), capi.GE, lo, capi.LE, hi, capi.COMB_AND, an.A_OID,
                                     an.A_OID_ERR, oids.ptr, n, err.ptr, err_cap, cnt.ptr, dsum.ptr, ws.h, stream)
        if st:
            capi.check(st)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for e0, e1 in ev:
        e0.record()
        step()
        e1.record()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    kern_ms = sum(e0.elapsed_time(e1) for e0, e1 in ev) / len(ev)
    n_oids, n_err = (int(x) for x in cnt.to_numpy(np.uint64, 2))
    total = int(dsum.to_numpy(np.uint64, 1)[0])
    err_np = err.to_numpy(np.uint64, min(n_err, err_cap))
    algo_bytes = 4 * n + 8 * n_oids + 8 * n_err
    achieved = algo_bytes / (kern_ms * 1e-3) / 1e9
    res = {
        "metric": "AN-hardened tuples/s (select + check-and-decode + sum)",
        "value": n * args.steps / el,
        "unit": "tuples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded, generated on device)",
        "config": {"workload": "select_sum: 1 column x %d rows, u16 values in [1,50] AN-coded u32 A=233, "
                               "1e-6 flips, 1<=v<=24" % n, "rows": n, "matches": n_oids, "errors": n_err,
                   "decoded_sum": total},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": profiled_traffic("select_sum_traffic_1e9") if n == 1_000_000_000 else None,
                     "kernel": "k_select<uint32_t,AND,SUM>", "kernel_ms": kern_ms,
                     "algorithmic_bytes": algo_bytes},
    }
    if not args.no_cpu:
        host_codes = codes.to_numpy(np.uint32, n)
        res["cpu_baseline"] = cpu_baseline_select(host_codes, meta, lo, hi, args.cpu_seconds,
                                                  (n_oids, err_np, total))
    return res


def cpu_baseline_star1(spec_full, seconds):
    """Oracle (scalar restatement of the operator chain) on a bounded sample of the workload."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as O
    from ahead_amd import workloads as W

    frac = 10  # sample: 1/10 of the fact rows, the full dimension
    spec = W.Star1Spec(n_fact=spec_full.n_fact // frac, n_dim=spec_full.n_dim, rate=spec_full.rate,
                       seed=spec_full.seed)
    h = W.star1_host(spec)
    cols = (O.Col(h["fk"], spec.meta_key), O.Col(h["m"], spec.meta_m), O.Col(h["key"], spec.meta_key),
            O.Col(h["a1"], spec.meta_attr), O.Col(h["a2"], spec.meta_attr))
    reps, t0 = 0, time.perf_counter()
    while True:
        O.star1_pipeline(*cols)
        reps += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    return {
        "value": spec.n_fact * reps / el,
        "unit": "tuples/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{spec.n_fact}-row fact sample (1/{frac} of the workload, same generator) x {reps} runs "
                  f"({el:.1f} s), single thread, oracle/an_oracle.c (-O3 -fno-tree-vectorize)",
    }


def run_star1(args):
    import numpy as np
    import torch

    from ahead_amd import capi, workloads as W

    spec = W.Star1Spec()
    if args.rows:
        spec.n_fact = args.rows
    cols = W.star1_device(spec)
    st = capi.Star1(*cols)
    stream = capi.current_stream()
    for _ in range(args.warmup):
        st.launch(stream)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for e0, e1 in ev:
        e0.record()
        st.launch(stream)
        e1.record()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    step_ms = sum(e0.elapsed_time(e1) for e0, e1 in ev) / len(ev)
    res_out = st.results()
    # the probe side (dominant kernel k_star_probe2) timed with HIP events on the launch stream
    probe_ms = time_star_probe(st, args.steps)
    nf = spec.n_fact
    algo_bytes = 16 * nf  # FK + measure codes read once; dimension (2.4 MB) and outputs excluded
    achieved = algo_bytes / (probe_ms * 1e-3) / 1e9
    res = {
        "metric": "AN-hardened tuples/s (hash join + group-by + sum, fused)",
        "value": nf * args.steps / el,
        "unit": "tuples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded, generated on device)",
        "config": {"workload": f"star1: fact {nf} rows (FK u64 A=63877, measure u64 A=65117), dimension "
                               f"{spec.n_dim} keys with 8-bit attributes of 25 and 5 values (u16 A=29), "
                               f"flips {spec.rate} per code, join + group by 2 attributes + sum",
                   "fact_rows": nf, "dim_rows": spec.n_dim, "groups": int(res_out["sums"].size),
                   "kept_rows": res_out["kept"], "status": res_out["status"],
                   "errors": {k: int(res_out[k].size) for k in ("e_fk", "e_dk", "e_a1", "e_a2", "e_m")},
                   "step_kernels_ms": step_ms},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": profiled_traffic("star1_traffic_60M") if (nf, spec.n_dim) == (60_000_000, 200_000)
                     else None,
                     "kernel": "probe side: k_star_probe2 (+ k_star_scan2, k_star_probe, k_star_reduce, k_star_final)",
                     "kernel_ms": probe_ms, "algorithmic_bytes": algo_bytes},
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_star1(spec, args.cpu_seconds)
    return res


def time_star_probe(st, steps):
    """Average duration of the probe side (k_star_probe2 and the small kernels that finish it) from
    HIP events on the launch stream, with the dimension table built once beforehand."""
    import torch

    from ahead_amd import capi

    stream = capi.current_stream()
    st.build(stream)
    for _ in range(3):
        st.probe(stream)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for e0, e1 in ev:
        e0.record()
        st.probe(stream)
        e1.record()
    torch.cuda.synchronize()
    return sum(e0.elapsed_time(e1) for e0, e1 in ev) / len(ev)


def run_batch(args):
    import numpy as np
    import torch

    from ahead_amd import autotune, capi, workloads as W

    spec = W.BatchSpec()
    if args.rows:
        spec.count = max(1, min(1024, args.rows))
    jobs = W.batch_device(spec)
    lib = capi.load()
    stream = capi.current_stream()
    # the batched form: autotune, then time replays of a hipGraph capture of the tuned launch
    (sub, blocks), tune_ms, table = autotune.tune_select_batch(jobs)
    batch = capi.SelectBatch(jobs, max_subtiles=sub, blocks=blocks)
    batch.launch()
    res_b = batch.results()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(g, stream=side):
            batch.launch()
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        g.replay()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        g.replay()
    e1.record()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    graph_ms = e0.elapsed_time(e1) / args.steps
    # the same launch, not captured: the kernel's own duration on its stream
    ek0, ek1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ek0.record()
    for _ in range(args.steps):
        batch.launch()
    ek1.record()
    torch.cuda.synchronize()
    kern_ms = ek0.elapsed_time(ek1) / args.steps
    res_g = batch.results()

    # the plain form: one ahead_hip_select_an launch per instance (same outputs, own buffers)
    ws = capi.workspace()
    descs = [j["col"].desc() for j in jobs]
    outs = [(capi.DeviceBuffer(8 * max(1, j["col"].n)), capi.DeviceBuffer(8 * 4096), capi.DeviceBuffer(16),
             capi.DeviceBuffer(8)) for j in jobs]

    def plain_step():
        for j, d, (o, e, c, sm) in zip(jobs, descs, outs):
            capi.check(lib.ahead_hip_select_an(ctypes.byref(d), j["op1"], j["th1"], j["op2"], j["th2"], j["comb"],
                                               capi.an.A_OID, capi.an.A_OID_ERR, o.ptr, j["col"].n, e.ptr, 4096,
                                               c.ptr, sm.ptr, ws.h, stream))

    for _ in range(max(1, args.warmup // 2)):
        plain_step()
    torch.cuda.synchronize()
    p0, p1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    plain_steps = max(1, args.steps // 2)
    p0.record()
    for _ in range(plain_steps):
        plain_step()
    p1.record()
    torch.cuda.synchronize()
    plain_ms = p0.elapsed_time(p1) / plain_steps
    # parity between the forms (and between the eager and graph launches)
    same = True
    for (ob, eb, sb), (og, eg, sg), (o, e, c, sm) in zip(res_b, res_g, outs):
        cnt = c.to_numpy(np.uint64, 2)
        same &= bool(np.array_equal(ob, og) and np.array_equal(eb, eg) and sb == sg)
        same &= int(cnt[0]) == ob.size and int(cnt[1]) == eb.size and int(sm.to_numpy(np.uint64, 1)[0]) == sb
        same &= bool(np.array_equal(o.to_numpy(np.uint64, int(cnt[0])), ob))
        same &= bool(np.array_equal(e.to_numpy(np.uint64, min(int(cnt[1]), 4096)), eb[:4096]))
    rows = sum(j["col"].n for j in jobs)
    matched = sum(r[0].size for r in res_b)
    errors = sum(r[1].size for r in res_b)
    algo_bytes = sum(j["col"].n * j["col"].meta.width for j in jobs) + 8 * matched + 8 * errors
    achieved = algo_bytes / (kern_ms * 1e-3) / 1e9
    res = {
        "metric": "batched hardened select+sum: instances/s (one persistent launch vs one launch per instance)",
        "value": spec.count * args.steps / el,
        "unit": "instances/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16/u32/u64",
        "data": "synthetic (seeded, generated on device)",
        "config": {"workload": f"batch: {spec.count} select+check-and-decode+sum instances, lognormal lengths "
                               f"(median 707k, sigma 1.42, [1e4, 5e7]), 16/32/64-bit codes, selectivity "
                               f"{spec.sel_min:.0%}-{spec.sel_max:.0%}, flips {spec.rate}",
                   "instances": spec.count, "rows": rows, "matched": matched, "errors": errors,
                   "tuned": {"max_subtiles": sub, "blocks": blocks, "sweep": table},
                   "graph_ms": graph_ms, "plain_ms": plain_ms,
                   "speedup_vs_plain": plain_ms / graph_ms, "tuples_per_s": rows / (graph_ms * 1e-3),
                   "parity_batched_vs_plain": same},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_select_batch (+sums)",
                     "kernel_ms": kern_ms, "algorithmic_bytes": algo_bytes},
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_batch(spec, jobs, res_b, args.cpu_seconds)
    batch.close()
    return res


def cpu_baseline_batch(spec, jobs, res_b, seconds):
    """The all-core vectorised CPU select (oracle/an_simd.cpp) on instances of the batch (host-built
    columns), in order, until the time budget is spent; each result is asserted equal to the batched
    GPU result of the same instance."""
    sys.path.insert(0, str(ROOT / "tests"))
    import numpy as np

    import oracle_lib as O
    from ahead_amd import workloads as W

    threads = len(os.sched_getaffinity(0))
    rows, done, el = 0, 0, 0.0
    for i, j in enumerate(jobs):
        codes, meta = W.batch_host_column(spec, i)
        t0 = time.perf_counter()
        oids, err, s = O.select_an_simd(O.Col(codes, meta), j["op1"], j["th1"], j["op2"], j["th2"], j["comb"],
                                        threads=threads)
        el += time.perf_counter() - t0
        if not (np.array_equal(oids, res_b[i][0]) and np.array_equal(err, res_b[i][1]) and s == res_b[i][2]):
            raise RuntimeError(f"the CPU baseline's instance {i} differs from the GPU's")
        rows += codes.size
        done += 1
        if el >= seconds:
            break
    cpu = host_cpu()
    return {"value": done / el, "unit": "instances/s", "cores": threads, "kind": "port",
            "sample": f"first {done} of {len(jobs)} instances ({rows} rows, {el:.1f} s of select time); "
                      f"oracle/an_simd.cpp or_select_an_simd on {threads} threads of {cpu['model']}; "
                      f"{rows / el:.3e} tuples/s; bit-exact vs GPU: True"}


def run_ops(args):
    """Every hardened operator alone at the join + group-by workload's sizes (60 M fact rows, 200 k
    dimension keys): HIP-event time of one entry-point call on preallocated outputs, algorithmic
    bytes (inputs read once, outputs written once), and the fraction of HBM peak."""
    import numpy as np
    import torch

    from ahead_amd import an, capi, workloads as W

    lib = capi.load()
    spec = W.Star1Spec()
    if args.rows:
        spec.n_fact = args.rows
    n, nd = spec.n_fact, spec.n_dim
    fk, m, key, a1, a2 = W.star1_device(spec)
    A_OID, A_ERR = an.A_OID, an.A_OID_ERR
    resoid = an.an_meta(an.RESOID, A_OID)
    # derived fact-side columns: OIDs into the dimension, 8-bit fact attribute, group ids, a second measure
    oids = capi.DeviceBuffer(8 * n)
    capi.gen_uniform(oids, 8, n, spec.seed, 31, 0, nd - 1, A_OID)
    c_oids = capi.Column(oids, n, resoid)
    fattr = capi.DeviceBuffer(2 * n)
    capi.gen_uniform(fattr, 2, n, spec.seed, 32, 0, 24, W.A_ATTR)
    c_fattr = capi.Column(fattr, n, spec.meta_attr)
    gids = capi.DeviceBuffer(8 * n)
    capi.gen_uniform(gids, 8, n, spec.seed, 33, 0, 124, A_OID)
    c_gids = capi.Column(gids, n, resoid)
    m2 = capi.DeviceBuffer(8 * n)
    capi.gen_uniform(m2, 8, n, spec.seed, 34, 0, 90_000, W.A_MEASURE)
    c_m2 = capi.Column(m2, n, spec.meta_m)
    void_n = capi.Column.void(n)
    void_d = capi.Column.void(nd)
    ws = capi.workspace()
    stream = capi.current_stream()
    ecap = 1 << 16
    errs = [capi.DeviceBuffer(8 * ecap) for _ in range(4)]
    eb = [capi.Buf(e.ptr, ecap) for e in errs]
    cnt = capi.DeviceBuffer(8 * 8)
    out_a = capi.DeviceBuffer(8 * n)
    out_b = capi.DeviceBuffer(8 * n)
    total = capi.DeviceBuffer(8)
    sums = capi.DeviceBuffer(8 * 1024)
    torch.cuda.synchronize()

    def desc(c):
        return ctypes.byref(c.desc())

    def timed(call, reps):
        d = []
        for i in range(reps + 1):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            capi.check(call())
            e1.record()
            if i:
                d.append((e0, e1))
        torch.cuda.synchronize()
        return sum(a.elapsed_time(b) for a, b in d) / len(d)

    def counts(k):
        return [int(x) for x in cnt.to_numpy(np.uint64, k)]

    reps = max(3, min(args.steps, 20))
    ops = []

    def record(name, call, bytes_of, note=""):
        ms = timed(call, reps)
        c = counts(8)
        b = bytes_of(c)
        gbs = b / (ms * 1e-3) / 1e9
        ops.append({"op": name, "ms": ms, "algorithmic_bytes": b, "achieved_gbs": gbs, "frac": gbs / HBM_PEAK_GBS,
                    "counts": c[:6], "note": note})

    # selectAN<>=, <=> on the measure codes (about half the rows match)
    mid = (90_000 + 10_500_000) // 2
    record("selectAN (range, 64-bit codes)",
           lambda: lib.ahead_hip_select_an(desc(m), capi.GE, 90_000 * W.A_MEASURE, capi.LE, mid * W.A_MEASURE,
                                           capi.COMB_AND, A_OID, A_ERR, out_a.ptr, n, errs[0].ptr, ecap, cnt.ptr,
                                           None, ws.h, stream),
           lambda c: 8 * n + 8 * c[0] + 8 * c[1])
    record("checkAN (64-bit codes)",
           lambda: lib.ahead_hip_check_an(desc(m), A_ERR, errs[0].ptr, ecap, cnt.ptr, ws.h, stream),
           lambda c: 8 * n + 8 * c[0])
    record("checkAndDecodeAN (64-bit codes -> 32-bit values)",
           lambda: lib.ahead_hip_check_and_decode_an(desc(m), A_ERR, out_a.ptr, n, errs[0].ptr, ecap, cnt.ptr, ws.h,
                                                     stream),
           lambda c: 8 * n + 4 * c[0] + 8 * c[1])
    record("fetchjoinAN (60M OIDs into a 200k-row 16-bit column)",
           lambda: lib.ahead_hip_fetchjoin_an(desc(c_oids), desc(a1), 0, 0, A_ERR, capi.Buf(out_a.ptr, n), eb[0], eb[1],
                                              cnt.ptr, ws.h, stream),
           lambda c: 8 * n + 2 * c[0] + 8 * (c[1] + c[2]), "dimension column (400 KB) cache-resident, not counted")
    # hashjoinAN(fact <void, FK>, reverse(dim <void, key>)): pairs (fact OID, dim OID)
    record("hashjoinAN (60M probe rows, 200k build keys)",
           lambda: lib.ahead_hip_hashjoin_an(desc(void_n), desc(fk), desc(key), desc(void_d), 1, 0, 0, 4, 0, 0, 4,
                                             A_ERR, capi.Buf(out_a.ptr, n), capi.Buf(out_b.ptr, n), *eb, cnt.ptr, ws.h,
                                             stream),
           lambda c: 8 * n + 8 * nd + 8 * c[0] + 8 * (c[1] + c[2] + c[3] + c[4]),
           "output pairs as 4-byte OIDs (void heads); build side (1.6 MB) counted once")
    record("groupbyAN unary (60M 16-bit codes, 25 groups)",
           lambda: lib.ahead_hip_groupby_an(desc(void_n), desc(c_fattr), A_OID, 0, out_a.ptr,
                                            capi.Buf(out_b.ptr, n), eb[0], eb[1], cnt.ptr, ws.h, stream),
           lambda c: 2 * n + 8 * n + 8 * c[1] + 8 * (c[2] + c[3]))
    record("aggregate_sum_groupedAN (60M measures, 125 groups)",
           lambda: lib.ahead_hip_aggregate_sum_grouped_an(desc(void_n), desc(m), desc(c_gids), 125, an.A_RESULT, A_OID,
                                                          sums.ptr, eb[0], eb[1], eb[2], cnt.ptr, ws.h, stream),
           lambda c: 8 * n + 8 * n + 8 * 125 + 8 * (c[1] + c[2] + c[3]))
    record("aggregate_mul_sumAN (60M 64-bit x 16-bit codes)",
           lambda: lib.ahead_hip_aggregate_mul_sum_an(desc(m), desc(c_fattr), 0, 1, an.A_RESULT, A_ERR, total.ptr, eb[0],
                                                      eb[1], cnt.ptr, ws.h, stream),
           lambda c: 8 * n + 2 * n + 8 * (c[1] + c[2]))
    record("arithmeticAN SUB (60M + 60M 64-bit codes -> 64-bit)",
           lambda: lib.ahead_hip_arithmetic_an(desc(void_n), desc(m), desc(void_n), desc(c_m2), 1, an.A_RESULT, 8, A_ERR,
                                               out_a.ptr, *eb, cnt.ptr, ws.h, stream),
           lambda c: 8 * n + 8 * n + 8 * n + 8 * (c[1] + c[2] + c[3] + c[4]))
    # matchjoinAN merge form on two ascending OID lists (the Q1.1 selection intersection)
    capi.check(lib.ahead_hip_select_an(desc(m), capi.LE, mid * W.A_MEASURE, capi.LT, 0, capi.COMB_NONE, A_OID, A_ERR,
                                       out_a.ptr, n, errs[0].ptr, ecap, cnt.ptr, None, ws.h, stream))
    na = counts(1)[0]
    capi.check(lib.ahead_hip_select_an(desc(c_fattr), capi.LE, 12 * W.A_ATTR, capi.LT, 0, capi.COMB_NONE, A_OID, A_ERR,
                                       out_b.ptr, n, errs[0].ptr, ecap, cnt.ptr, None, ws.h, stream))
    nb = counts(1)[0]
    la, lb = capi.Column(out_a, na, resoid), capi.Column(out_b, nb, resoid)
    mo_h, mo_t = capi.DeviceBuffer(8 * min(na, nb)), capi.DeviceBuffer(8 * min(na, nb))
    record("matchjoinAN merge (two ascending OID lists)",
           lambda: lib.ahead_hip_matchjoin_merge_an(desc(la), desc(la), desc(lb), desc(lb), 0, 0, 8, 0, 0, 8, A_ERR,
                                                    capi.Buf(mo_h.ptr, min(na, nb)), capi.Buf(mo_t.ptr, min(na, nb)),
                                                    *eb, cnt.ptr, ws.h, stream),
           lambda c: 8 * na + 8 * nb + 16 * c[0] + 8 * (c[1] + c[2] + c[3] + c[4]),
           f"lists of {na} and {nb} OIDs; matched rows' second side re-read counted once")
    worst = min(ops, key=lambda o: o["frac"])
    return {
        "metric": "AN-hardened operators alone at the 60M-row join workload's sizes: % of HBM peak (lowest operator)",
        "value": worst["frac"],
        "unit": "fraction of HBM peak",
        "n_gpus": 1,
        "steps": reps,
        "warmup": 1,
        "ms_per_step": sum(o["ms"] for o in ops),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16/u32/u64",
        "data": "synthetic (seeded, generated on device)",
        "config": {"workload": f"ops: {n} fact rows, {nd} dimension keys; one entry-point call per operator",
                   "ops": ops},
        "roofline": {"bound": "hbm", "achieved": worst["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": worst["frac"], "traffic": None, "kernel": worst["op"], "kernel_ms": worst["ms"],
                     "algorithmic_bytes": worst["algorithmic_bytes"]},
    }


def run_hbm_copy(args):
    import torch

    nbytes = args.rows or (4 << 30)
    src = torch.empty(nbytes, dtype=torch.uint8, device="cuda").fill_(7)
    dst = torch.empty_like(src)
    for _ in range(max(1, args.warmup)):
        dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    gbs = 2 * nbytes / (ms * 1e-3) / 1e9
    return {"metric": "device copy GB/s (read + write)", "value": gbs, "unit": "GB/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
            "config": {"workload": f"hbm_copy: {nbytes} bytes read and {nbytes} bytes written per step"},
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS, "traffic": 2 * nbytes, "kernel": "torch copy",
                         "kernel_ms": ms, "algorithmic_bytes": 2 * nbytes}}


def run_serve(args):
    import torch

    from ahead_amd import serve, ssb

    spec = ssb.SsbSpec()
    if args.rows:
        spec.n_fact = args.rows
    tables = ssb.device_tables(spec)
    torch.cuda.synchronize()
    n_q = max(16, args.steps * 16)
    # long-lived servers: same-shape batches of fused queries and single fused queries on 1 and 4
    # slots, the plain operator chains on 1 and 4
    forms = {"native4": ("native", 4), "native1": ("native", 1), "batched4": ("batched", 4),
             "batched1": ("batched", 1), "fused4": ("fused", 4), "fused1": ("fused", 1), "chain1": ("chain", 1),
             "chain4": ("chain", 4)}
    servers = {k: serve.Server(tables, spec.n_fact, workers=w, mode=m) for k, (m, w) in forms.items()}
    for k, srv in servers.items():   # warm-up on every shape
        srv.run(serve.random_queries(4 * srv.workers, 0, 99), realtime=False)
    torch.cuda.reset_peak_memory_stats()
    retries0 = torch.cuda.memory_stats().get("num_alloc_retries", 0)
    # capacity of each form: everything arrives at once
    sat = {}
    for k, srv in servers.items():
        sat[k] = serve.summarize(srv.run(serve.random_queries(n_q, 0, 7), realtime=False))
        print(f"serve: {k}, all at once: {sat[k]['qps']:.1f} queries/s, p99 {sat[k]['p99_ms']:.1f} ms",
              file=sys.stderr, flush=True)
    # Poisson arrivals at multiples of the plain path's capacity; every form sees the same arrivals
    sweep = []
    base = sat["chain1"]["qps"]
    for frac in (0.5, 0.9, 2.0, 4.0, 8.0, 16.0, 24.0, 32.0):
        rate = base * frac
        row = {"rate_qps": rate}
        for k, srv in servers.items():
            if rate > 1.2 * sat[k]["qps"]:
                continue   # beyond the form's capacity: its queue only grows
            row[k] = serve.summarize(srv.run(serve.random_queries(n_q, rate, 11), realtime=True))
            print(f"serve: {k}, Poisson {rate:.1f} queries/s: p50 {row[k]['p50_ms']:.1f} ms, "
                  f"p99 {row[k]['p99_ms']:.1f} ms", file=sys.stderr, flush=True)
        sweep.append(row)
    t0 = time.perf_counter()
    final = servers["batched4"].run(serve.random_queries(n_q, 0, 13), realtime=False)
    el = time.perf_counter() - t0
    s4 = serve.summarize(final)
    fin_f = serve.summarize(servers["fused4"].run(serve.random_queries(n_q, 0, 13), realtime=False))
    needs = {k: {sh: servers[k].need(sh) for sh in SSB_SHAPES} for k in ("fused4", "chain1")}
    for srv in servers.values():
        srv.close()
    mem = torch.cuda.memory_stats()
    memory = {"alloc_retries": mem.get("num_alloc_retries", 0) - retries0,
              "peak_allocated_gb": mem.get("allocated_bytes.all.peak", 0) / 1e9,
              "peak_reserved_gb": mem.get("reserved_bytes.all.peak", 0) / 1e9,
              "admission_bytes_fused": needs["fused4"], "admission_bytes_chain": needs["chain1"]}
    return {
        "metric": "served star-schema queries/s (fused star queries in same-shape batches, 4 slots) with p50/p99 latency",
        "value": s4["qps"],
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 1,
        "ms_per_step": el / max(1, args.steps) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16/u32/u64",
        "data": "synthetic (seeded; fact columns generated on device)",
        "config": {"workload": f"serve: ssb schema with {spec.n_fact} fact rows resident in HBM; {n_q} queries "
                               f"per run, shapes P1-P4 with seeded random predicates",
                   "saturated": sat,
                   "speedup_saturated_vs_plain": sat["batched4"]["qps"] / base if base else None,
                   "poisson_sweep": sweep, "final_batched_4_slots": s4, "final_fused_4_slots": fin_f,
                   "device_memory": memory},
        "roofline": None,
    }


def run_shape_batches(args):
    """Sixteen queries of each shape (seeded random constants) as one ahead_hip_star_query_batch
    call (same-shape shared scans) against the same sixteen one at a time; results asserted equal."""
    import torch

    from ahead_amd import capi, serve, ssb

    spec = ssb.SsbSpec()
    if args.rows:
        spec.n_fact = args.rows
    t = ssb.device_tables(spec)
    torch.cuda.synchronize()
    per_shape, tot_single, tot_batch = {}, 0.0, 0.0
    nb = capi.SQ_MAX_SHAPE_PASS
    ev = capi.HipEvents(nb + 1)
    for shape in SSB_SHAPES:
        qs = [q.spec() for q in serve.random_queries(400, 0, 5) if q.shape == shape][:nb]
        singles = [ssb.fused(t, q) for q in qs]
        if not all(serve.same_result(a, b) for a, b in zip(singles, ssb.fused_batch(t, qs))):
            raise RuntimeError(f"{shape}: a batched query differs from the query alone")
        best_s = best_b = float("inf")
        scan_s = scan_b = float("inf")
        for _ in range(max(1, args.steps)):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for i, q in enumerate(qs):
                ssb.fused(t, q, ev_scan=ev.pairs[i])
            torch.cuda.synchronize()
            best_s = min(best_s, time.perf_counter() - t0)
            scan_s = min(scan_s, sum(ev.ms(i) for i in range(len(qs))))
            t0 = time.perf_counter()
            ssb.fused_batch(t, qs, ev_scan=ev.pairs[nb])
            torch.cuda.synchronize()
            best_b = min(best_b, time.perf_counter() - t0)
            scan_b = min(scan_b, ev.ms(nb))
        per_shape[shape] = {"queries": len(qs), "single_wall_ms": best_s * 1e3, "batch_wall_ms": best_b * 1e3,
                            "single_scans_ms": scan_s, "batch_maps_and_scan_ms": scan_b}
        tot_single += best_s
        tot_batch += best_b
        print(f"shape_batches: {shape}: {nb} single {best_s * 1e3:.1f} ms, one batch {best_b * 1e3:.1f} ms",
              file=sys.stderr, flush=True)
    ev.close()
    nq = nb * len(SSB_SHAPES)
    return {
        "metric": f"star-schema queries/s, {nb} queries of one shape per ahead_hip_star_query_batch call",
        "value": nq / tot_batch,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 1,
        "ms_per_step": tot_batch * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16/u32/u64",
        "data": "synthetic (seeded; fact columns generated on device)",
        "config": {"workload": f"shape_batches: ssb schema with {spec.n_fact} fact rows, {nb} queries per shape P1-P4",
                   "single_queries_per_s": nq / tot_single, "per_shape": per_shape},
        "roofline": None,
    }


def run_strategies(args):
    """The four SSB query shapes under each of the reference's strategies, one after another: the
    hardened chain (continuous), early and late check-and-decode, the unhardened chain (normal),
    DMR (sequential and concurrent replicas) and TMR, plus the fused star query."""
    import torch

    from ahead_amd import ssb
    from ahead_amd import strategies as S

    # fault-free (early drops corrupt rows, which misaligns the columns); 300 M fact rows: the schema,
    # its unhardened values and two more replicas of them stay resident together
    spec = ssb.SsbSpec(rate=0.0, n_fact=args.rows or 300_000_000)
    t = ssb.device_tables(spec)
    plain = S.plain_tables(t)
    reps3 = S.replicas(plain, 3)
    torch.cuda.synchronize()
    qs = [ssb.SHAPES[k]() for k in SSB_SHAPES]
    forms = {
        "continuous": lambda q: ssb.chain(t, q),
        "continuous_reenc": lambda q: S.continuous_reenc(t, q),
        "early": lambda q: S.early(t, q),
        "late": lambda q: S.late(t, q),
        "normal": lambda q: S.normal(plain, q),
        "dmr": lambda q: S.modular(reps3[:2], q),
        "dmr_concurrent": lambda q: S.modular(reps3[:2], q, concurrent=True),
        "tmr": lambda q: S.modular(reps3, q),
        "fused": lambda q: ssb.fused(t, q),
    }
    ref = [ssb.chain(t, q) for q in qs]
    steps = max(1, args.steps)
    out = {}
    for name, fn in forms.items():
        got = [fn(q) for q in qs]  # warm-up, and the results the step must reproduce
        for q, g, r in zip(SSB_SHAPES, got, ref):
            if not S.same(g, r):
                raise RuntimeError(f"{name} differs from the hardened chain on {q}")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            for q in qs:
                fn(q)
        torch.cuda.synchronize()
        el = (time.perf_counter() - t0) / steps
        out[name] = {"ms_per_step": el * 1e3, "queries_per_s": len(qs) / el}
        del got
        torch.cuda.empty_cache()  # the next strategy's intermediates (and stream workspaces) start afresh
        print(f"strategies: {name} {el * 1e3:.1f} ms per step of {len(qs)} queries", file=sys.stderr, flush=True)
    base = out["normal"]["ms_per_step"]
    for v in out.values():
        v["overhead_vs_normal"] = v["ms_per_step"] / base
    return {
        "metric": "SSB queries/s per execution strategy (Q1.1, Q2.1, Q3.1, Q4.1 per step)",
        "value": out["continuous"]["queries_per_s"],
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 1,
        "ms_per_step": out["continuous"]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16/u32/u64",
        "data": "synthetic (seeded; fact columns generated on device)",
        "config": {"workload": f"strategies: ssb schema with {spec.n_fact} fact rows; every strategy's results "
                               f"checked equal to the hardened chain's", "strategies": out},
        "roofline": None,
    }


SSB_SHAPES = ("P1", "P2", "P3", "P4")  # Q1.1, Q2.1, Q3.1, Q4.1 of the reference's SSB drivers
SSB_BYTES_PER_ROW = {"P1": 20, "P2": 32, "P3": 32, "P4": 48}  # fact-column bytes each query reads once
SSB_SCAN_KERNEL = {"P1": "k_sq_scan<1, 2>", "P2": "k_sq_scan<3, 0>", "P3": "k_sq_scan<3, 0>", "P4": "k_sq_scan<4, 0>"}


def host_cpu() -> dict:
    """This host's CPU: model name, logical CPUs and the CPUs this process may run on."""
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def _same_query(a: dict, b: dict) -> bool:
    import numpy as np

    if a["rows"] != b["rows"]:
        return False
    if "result" in b:
        if a["result"] != b["result"]:
            return False
    elif a["groups"] != b["groups"] or not np.array_equal(a["sums"], b["sums"]) or \
            not all(np.array_equal(x, y) for x, y in zip(a["keys"], b["keys"])):
        return False
    if "errors" in a and "errors" in b:
        ea, eb = dict(a["errors"].items), dict(b["errors"].items)
        return sorted(ea) == sorted(eb) and all(np.array_equal(ea[k], eb[k]) for k in eb)
    return True


def run_ssb(args):
    import numpy as np
    import torch

    from ahead_amd import an, capi, ssb

    spec = ssb.SsbSpec()
    if args.rows:
        spec.n_fact = args.rows
    n = spec.n_fact
    print(f"ssb: generating the schema ({n} fact rows)", file=sys.stderr, flush=True)
    t = ssb.device_tables(spec)
    torch.cuda.synchronize()
    queries = {k: ssb.SHAPES[k]() for k in SSB_SHAPES}
    # the operator chains (the per-operator drop-in path) once, then the fused form: bit-exact
    print("ssb: operator chains vs fused star queries", file=sys.stderr, flush=True)
    chain_ms, chain_res, fused_res = {}, {}, {}
    for q in queries.values():  # once untimed: workspace growth and first-call costs stay out of the times
        ssb.chain(t, q)
    for k, q in queries.items():
        torch.cuda.synchronize()
        q0 = time.perf_counter()
        chain_res[k] = ssb.chain(t, q)
        torch.cuda.synchronize()
        chain_ms[k] = (time.perf_counter() - q0) * 1e3
        fused_res[k] = ssb.fused(t, q, fallback=False)
        if fused_res[k]["status"] != 0 or not _same_query(fused_res[k], chain_res[k]):
            raise RuntimeError(f"{k}: fused star query differs from the operator chain")
    for _ in range(args.warmup):
        for q in queries.values():
            ssb.fused(t, q)
    torch.cuda.synchronize()
    ev = capi.HipEvents(len(queries) * args.steps)
    wall = {k: 0.0 for k in queries}
    scan_ms = {k: [] for k in queries}
    print(f"ssb: timing {args.steps} steps of {len(queries)} fused queries", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    i = 0
    for _ in range(args.steps):
        for k, q in queries.items():
            q0 = time.perf_counter()
            r = ssb.fused(t, q, ev_scan=ev.pairs[i])  # returns after the query's results are on the host
            wall[k] += time.perf_counter() - q0
            if r["rows"] != fused_res[k]["rows"]:
                raise RuntimeError(f"{k}: result changed between runs")
            i += 1
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    i = 0
    for _ in range(args.steps):
        for k in queries:
            scan_ms[k].append(ev.ms(i))
            i += 1
    ev.close()
    # host checks of the step's outputs: Q1.1 against a textbook evaluation over the whole fact
    # table, and every selection error vector against the generator's flip positions (closed form:
    # every single-bit flip is detected)
    print("ssb: host textbook and flip-position checks", file=sys.stderr, flush=True)
    workers = min(16, len(os.sched_getaffinity(0)))
    host_ok = flips_ok = None
    if not args.no_check:
        tb = ssb.textbook(spec, queries["P1"], chunk=20_000_000, workers=workers)
        host_ok = tb["rows"] == fused_res["P1"]["rows"] and tb["result"] == fused_res["P1"]["result"]
        flips_ok = True
    aoe = np.uint64(an.A_OID_ERR)
    for k, q in queries.items():
        if args.no_check:
            break
        e = dict(fused_res[k]["errors"].items)
        cols = [c for c, _ in q.preds] + [c for d in q.dims for c, _ in d.sels]
        for c in cols:
            exp = ssb.flip_positions(spec, c, chunk=20_000_000, workers=workers) * aoe
            got = e.get(f"select_{c}.e1", np.zeros(0, np.uint64))
            flips_ok = flips_ok and np.array_equal(got, exp)
    if not args.no_check and not (host_ok and flips_ok):
        raise RuntimeError(f"host checks failed: textbook {host_ok}, flip positions {flips_ok}")
    per = {}
    for k in queries:
        sm = float(np.mean(scan_ms[k]))
        ab = SSB_BYTES_PER_ROW[k] * n
        per[k] = {"wall_ms": wall[k] / args.steps * 1e3, "scan_ms": sm, "scan_kernel": SSB_SCAN_KERNEL[k],
                  "scan_algorithmic_bytes": ab, "scan_gbs": ab / (sm * 1e-3) / 1e9,
                  "scan_frac": ab / (sm * 1e-3) / 1e9 / HBM_PEAK_GBS, "chain_ms": chain_ms[k],
                  "rows": fused_res[k]["rows"], "groups": fused_res[k].get("groups", 1),
                  "errors": fused_res[k]["errors"].total()}
    dom = max(per, key=lambda k: per[k]["scan_ms"])
    nq = len(queries) * args.steps
    res = {
        "metric": "AN-hardened tuples/s and % HBM peak per op; batched queries/s + p99 vs plain",
        "value": n * nq / el,
        "unit": "tuples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16/u64",
        "data": "synthetic (seeded; fact columns generated on device)",
        "config": {"workload": f"ssb: star schema, fact {n} rows (4 FKs, quantity/discount u8, price/revenue/"
                               f"supplycost int32, AN-coded per width), date {spec.n_date}, customer {spec.n_cust}, "
                               f"supplier {spec.n_supp}, part {spec.n_part}; flips {spec.rate} in the selected "
                               f"columns; a step runs Q1.1, Q2.1, Q3.1, Q4.1 as fused star queries",
                   "queries_per_s": nq / el, "queries": per,
                   "chain_queries_per_s": len(queries) / (sum(chain_ms.values()) * 1e-3),
                   "checks": {"fused_equals_operator_chain": True, "q11_equals_host_textbook": host_ok,
                              "selection_errors_equal_flip_positions": flips_ok}},
        "roofline": {"bound": "hbm", "achieved": per[dom]["scan_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": per[dom]["scan_frac"],
                     "traffic": profiled_traffic("ssb_scan_traffic_600M") if n == 600_000_000 else None,
                     "kernel": f"{SSB_SCAN_KERNEL[dom]} ({dom})", "kernel_ms": per[dom]["scan_ms"],
                     "algorithmic_bytes": per[dom]["scan_algorithmic_bytes"]},
    }
    tr = res["roofline"]["traffic"]
    if tr:  # the bytes the kernel really moves (it skips measure lines of rows that do not join)
        res["roofline"]["traffic_gbs"] = tr / (per[dom]["scan_ms"] * 1e-3) / 1e9
        res["roofline"]["traffic_frac"] = res["roofline"]["traffic_gbs"] / HBM_PEAK_GBS
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_ssb(spec, queries, args.cpu_seconds)
    return res


def cpu_baseline_ssb(spec_full, queries, seconds):
    """The same four queries on the all-core vectorised CPU restatement of the fused star query
    (oracle/an_simd.cpp: OpenMP over every usable core, AVX-512 / AVX2 build picked by the CPU), on a
    fact sample; the sample's results are asserted equal to the GPU's on the same sample."""
    sys.path.insert(0, str(ROOT / "tests"))
    import dataclasses

    import oracle_bat as OB
    import oracle_lib as O
    from ahead_amd import ssb

    n = min(spec_full.n_fact, 30_000_000)
    spec = dataclasses.replace(spec_full, n_fact=n)
    print(f"ssb cpu baseline: {n}-row sample", file=sys.stderr, flush=True)
    host = OB.tables(spec)
    dev = ssb.device_tables(spec)
    threads = len(os.sched_getaffinity(0))
    OB.SIMD_THREADS = threads
    try:
        exact = all(_same_query(ssb.fused(dev, q), ssb.fused(host, q, OB)) for q in queries.values())
        if not exact:
            raise RuntimeError("the CPU baseline's results differ from the GPU's on the sample")
        done, el = 0, 0.0
        while el < seconds:
            for q in queries.values():
                t0 = time.perf_counter()
                ssb.fused(host, q, OB)
                el += time.perf_counter() - t0
                done += 1
    finally:
        OB.SIMD_THREADS = None
    cpu = host_cpu()
    isa = {4: "AVX-512 (x86-64-v4)", 3: "AVX2 (x86-64-v3)", 0: "x86-64 baseline"}[O.simd_level()]
    return {"value": n * done / el, "unit": "tuples/s", "cores": threads, "kind": "port",
            "sample": f"{done} fused star queries (Q1.1/Q2.1/Q3.1/Q4.1) on a {n}-row fact sample of the same "
                      f"generator ({el:.1f} s); oracle/an_simd.cpp or_star_query_simd, {isa} build, OpenMP on "
                      f"{threads} threads of {cpu['model']} ({cpu['nproc']} logical CPUs); dimension selections "
                      f"by the scalar oracle; results bit-exact vs the GPU on the sample: {exact}"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: the AN operator path has no multi-GPU code "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)
    import torch

    assert torch.cuda.is_available(), "bench.py needs an MI355X (cuda:0)"
    torch.cuda.set_device(0)
    run = {"star1": run_star1, "select_sum": run_select_sum, "batch": run_batch, "ssb": run_ssb,
           "serve": run_serve, "hbm_copy": run_hbm_copy, "ops": run_ops, "strategies": run_strategies,
           "shape_batches": run_shape_batches}[args.workload]
    res = run(args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
