"""Autotuner of the batched select launch: sweeps the look-back tile size (16 KiB sub-tiles per
tile) and the number of resident workgroups (per CU), times each candidate on the actual batch
with HIP events, and returns the fastest.  Results are kept per (device name, batch signature)
for the life of the process."""
from __future__ import annotations

import itertools

from . import capi

_cache: dict = {}

SUBTILES = (2, 4, 8, 16)
BLOCKS_PER_CU = (2, 3, 4)


def _time_launches(batch: "capi.SelectBatch", reps: int) -> float:
    import torch

    batch.launch()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        batch.launch()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def tune_select_batch(jobs: list[dict], reps: int = 3, subtiles=SUBTILES, blocks_per_cu=BLOCKS_PER_CU):
    """Returns ((max_subtiles, blocks), ms, table) for the fastest candidate; table lists all."""
    import torch

    props = torch.cuda.get_device_properties(0)
    cus = props.multi_processor_count
    key = (props.name, len(jobs), sum(j["col"].n for j in jobs), tuple(subtiles), tuple(blocks_per_cu))
    if key in _cache:
        return _cache[key]
    table = []
    for sub, bpc in itertools.product(subtiles, blocks_per_cu):
        b = capi.SelectBatch(jobs, max_subtiles=sub, blocks=cus * bpc)
        try:
            ms = _time_launches(b, reps)
        finally:
            b.close()
        table.append({"max_subtiles": sub, "blocks": cus * bpc, "ms": ms})
    best = min(table, key=lambda r: r["ms"])
    _cache[key] = ((best["max_subtiles"], best["blocks"]), best["ms"], table)
    return _cache[key]
