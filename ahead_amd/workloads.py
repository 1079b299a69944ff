"""Synthetic workloads of the benchmark configurations, generated from seeded counter-based streams.

star1 (join + 2-key group-by + sum): a fact table of n_fact rows with a foreign key uniform over a
dense dimension of n_dim keys and a measure uniform in [90000, 10500000]; the dimension holds its
key (= row id) and two 8-bit attributes with 25 and 5 distinct values.  Keys are 32-bit values
AN-coded to uint64 with A=63877, the measure is coded to uint64 with the RESINT load default
A=65117, the attributes to uint16 with A=29.  Every column receives seeded single-bit flips at the
given rate.  `*_host` builds the columns with numpy (ahead_amd.datagen), `*_device` builds the
identical columns on the GPU (csrc/datagen.hip).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import an, datagen

A_KEY = 63877
A_MEASURE = 65117
A_ATTR = 29
M_LO, M_HI = 90_000, 10_500_000

# stream ids: values use s, flips use s (hit) and s+1 (bit)
S_FK, S_M, S_A1, S_A2 = 1, 2, 3, 4
F_FK, F_M, F_KEY, F_A1, F_A2 = 11, 13, 15, 17, 19


@dataclass
class Star1Spec:
    n_fact: int = 60_000_000
    n_dim: int = 200_000
    card1: int = 25
    card2: int = 5
    rate: float = 1e-7
    seed: int = 2024

    @property
    def meta_key(self):
        return an.an_meta(an.RESINT, A_KEY)

    @property
    def meta_m(self):
        return an.an_meta(an.RESINT, A_MEASURE)

    @property
    def meta_attr(self):
        return an.an_meta(an.RESTINY, A_ATTR)


def star1_host(spec: Star1Spec) -> dict:
    th = datagen.flip_threshold(spec.rate)
    fk = datagen.encode(datagen.uniform(spec.n_fact, spec.seed, S_FK, 0, spec.n_dim - 1), A_KEY, 8)
    m = datagen.encode(datagen.uniform(spec.n_fact, spec.seed, S_M, M_LO, M_HI), A_MEASURE, 8)
    key = datagen.encode(np.arange(spec.n_dim, dtype=np.uint64), A_KEY, 8)
    a1 = datagen.encode(datagen.uniform(spec.n_dim, spec.seed, S_A1, 0, spec.card1 - 1), A_ATTR, 2)
    a2 = datagen.encode(datagen.uniform(spec.n_dim, spec.seed, S_A2, 0, spec.card2 - 1), A_ATTR, 2)
    flips = {
        "fk": datagen.inject_flips(fk, spec.seed, F_FK, th),
        "m": datagen.inject_flips(m, spec.seed, F_M, th),
        "key": datagen.inject_flips(key, spec.seed, F_KEY, th),
        "a1": datagen.inject_flips(a1, spec.seed, F_A1, th),
        "a2": datagen.inject_flips(a2, spec.seed, F_A2, th),
    }
    return {"fk": fk, "m": m, "key": key, "a1": a1, "a2": a2, "flips": flips}


def star1_device(spec: Star1Spec):
    """Returns the five device columns (capi.Column) of the star1 workload."""
    from . import capi

    th = datagen.flip_threshold(spec.rate)
    nf, nd = spec.n_fact, spec.n_dim
    fk = capi.DeviceBuffer(8 * nf)
    m = capi.DeviceBuffer(8 * nf)
    key = capi.DeviceBuffer(8 * nd)
    a1 = capi.DeviceBuffer(2 * nd)
    a2 = capi.DeviceBuffer(2 * nd)
    capi.gen_uniform(fk, 8, nf, spec.seed, S_FK, 0, nd - 1, A_KEY)
    capi.gen_uniform(m, 8, nf, spec.seed, S_M, M_LO, M_HI, A_MEASURE)
    capi.encode_void_into(key, nd, A_KEY, 8)
    capi.gen_uniform(a1, 2, nd, spec.seed, S_A1, 0, spec.card1 - 1, A_ATTR)
    capi.gen_uniform(a2, 2, nd, spec.seed, S_A2, 0, spec.card2 - 1, A_ATTR)
    capi.inject_flips(fk, 8, nf, spec.seed, F_FK, th)
    capi.inject_flips(m, 8, nf, spec.seed, F_M, th)
    capi.inject_flips(key, 8, nd, spec.seed, F_KEY, th)
    capi.inject_flips(a1, 2, nd, spec.seed, F_A1, th)
    capi.inject_flips(a2, 2, nd, spec.seed, F_A2, th)
    return (capi.Column(fk, nf, spec.meta_key), capi.Column(m, nf, spec.meta_m), capi.Column(key, nd, spec.meta_key),
            capi.Column(a1, nd, spec.meta_attr), capi.Column(a2, nd, spec.meta_attr))


# ------------------------------------------------------------------------------------------------
# batched select + aggregate: many independent instances of uneven length and mixed code width
# ------------------------------------------------------------------------------------------------

# (column type, A): 8-bit values in 16-bit codes, 16-bit values in 32-bit codes (two As), 32-bit
# values in 64-bit codes.  The three widths are equally likely; the 32-bit codes split between the
# two As.
BATCH_KINDS = [(an.RESTINY, 233), (an.RESSHORT, 233), (an.RESSHORT, 63877), (an.RESINT, 65117)]
BATCH_VMAX = {2: 0xFF, 4: 0xFFFF, 8: 0xFFFFFFFF}


@dataclass
class BatchSpec:
    count: int = 512
    mu: float = float(np.log(7.07e5))   # lognormal length: median 707k rows
    sigma: float = 1.42
    n_min: int = 10_000
    n_max: int = 50_000_000
    sel_min: float = 0.01
    sel_max: float = 0.50
    rate: float = 1e-6
    seed: int = 4242


def batch_plan(spec: BatchSpec) -> list[dict]:
    """Per instance: rows, column type, A, value range [lo, hi] of its predicate (seeded)."""
    rng = np.random.default_rng(spec.seed)
    n = np.clip(np.round(np.exp(rng.normal(spec.mu, spec.sigma, spec.count))), spec.n_min, spec.n_max)
    width_class = rng.integers(0, 3, spec.count)
    two_a = rng.integers(0, 2, spec.count)
    sel = rng.uniform(spec.sel_min, spec.sel_max, spec.count)
    start = rng.uniform(0.0, 1.0, spec.count)
    plan = []
    for i in range(spec.count):
        kind = [0, 1 + int(two_a[i]), 3][int(width_class[i])]
        t, a = BATCH_KINDS[kind]
        vmax = BATCH_VMAX[t.width]
        span = max(1, int(sel[i] * vmax))
        lo = int(start[i] * (vmax - span))
        plan.append({"n": int(n[i]), "type": t, "a": a, "lo": lo, "hi": lo + span, "vmax": vmax,
                     "stream": 100 + 4 * i})  # values: stream; flips: stream+1 (hit), +2 (bit)
    return plan


def batch_device(spec: BatchSpec) -> list[dict]:
    """The instances as select jobs (capi.SelectBatch / select_an arguments) over device columns
    generated from the seeded streams: values uniform in [0, vmax], AN-coded, flipped at `rate`."""
    from . import capi

    th = datagen.flip_threshold(spec.rate)
    jobs = []
    for p in batch_plan(spec):
        t, a = p["type"], p["a"]
        buf = capi.DeviceBuffer(t.width * p["n"])
        capi.gen_uniform(buf, t.width, p["n"], spec.seed, p["stream"], 0, p["vmax"], a)
        capi.inject_flips(buf, t.width, p["n"], spec.seed, p["stream"] + 1, th)
        col = capi.Column(buf, p["n"], an.an_meta(t, a))
        jobs.append({"col": col, "op1": capi.GE, "th1": p["lo"] * a, "op2": capi.LE, "th2": p["hi"] * a,
                     "comb": capi.COMB_AND, "err_cap": 256})
    return jobs


def batch_host_column(spec: BatchSpec, i: int) -> tuple[np.ndarray, an.ColMeta]:
    """Instance i's codes built with numpy (the same streams as batch_device)."""
    p = batch_plan(spec)[i]
    t, a = p["type"], p["a"]
    codes = datagen.encode(datagen.uniform(p["n"], spec.seed, p["stream"], 0, p["vmax"]), a, t.width)
    datagen.inject_flips(codes, spec.seed, p["stream"] + 1, datagen.flip_threshold(spec.rate))
    return codes, an.an_meta(t, a)
