#!/usr/bin/env python3
"""Benchmark of the AN-hardened operator path on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is one pass of the hardened hot path over one batch of synthetic input already resident in
HBM.  Workloads:
  star1       : (default) hardened hash join + group-by: 60,000,000 fact rows with a foreign key
                uniform over a dense dimension of 200,000 keys and a measure uniform in
                [90000, 10500000]; the dimension has two 8-bit attributes with 25 and 5 distinct
                values coded to uint16 (A=29); keys are 32-bit values coded to uint64 (A=63877),
                the measure to uint64 (A=65117); 1e-7 seeded single-bit flips in every column;
                join, group by both attributes (125 groups), int64 sum -- one fused launch chain.
  select_sum  : range select + check-and-decode + sum over 6,000,000 16-bit values in [1,50]
                AN-coded to uint32 with A=233, 1e-6 seeded single-bit flips, predicate 1<=v<=24
                on encoded bounds; outputs OIDs, decoded sum and sorted error positions.

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
    ap.add_argument("--workload", default="star1", choices=["star1", "select_sum"])
    ap.add_argument("--rows", type=int, default=None, help="override the row count (scaling studies)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    return ap.parse_args()


def cpu_baseline_select(codes_np, meta, lo, hi, seconds, expect):
    """Oracle (scalar restatement of selectAN + check-and-decode + sum) on this host's cores."""
    sys.path.insert(0, str(ROOT / "tests"))
    import numpy as np

    import oracle_lib as O

    col = O.Col(codes_np, meta)
    reps, t0 = 0, time.perf_counter()
    while True:
        oids, err, s = O.select_an(col, 3, lo, 1, hi, 1)
        reps += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    ok = (oids.size == expect[0] and np.array_equal(err, expect[1]) and s == expect[2])
    return {
        "value": codes_np.size * reps / el,
        "unit": "tuples/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{codes_np.size}-row column x {reps} repetitions ({el:.1f} s), single thread, "
                  f"oracle/an_oracle.c built -O3 -fno-tree-vectorize; bit-exact vs GPU: {ok}",
    }


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
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
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
    # the dominant kernel (k_star_probe) timed alone with HIP events on the launch stream
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
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_star_probe (+k_star_final)",
                     "kernel_ms": probe_ms, "algorithmic_bytes": algo_bytes},
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_star1(spec, args.cpu_seconds)
    return res


def time_star_probe(st, steps):
    """Average duration of the fact-side launch (k_star_probe + the one-workgroup k_star_final)
    from HIP events on the launch stream, with the dimension table built once beforehand."""
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


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: the AN operator path has no multi-GPU code "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)
    import torch

    assert torch.cuda.is_available(), "bench.py needs an MI355X (cuda:0)"
    torch.cuda.set_device(0)
    res = run_star1(args) if args.workload == "star1" else run_select_sum(args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
