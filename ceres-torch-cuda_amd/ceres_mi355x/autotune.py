"""Launch-shape autotuner for the persistent batched solves.

A batched launch is shaped by the workgroup size (threads per problem), the
number of persistent workgroups pulling problems from the work queue, and the
order in which problems are queued (index order, or largest first, which
balances uneven sizes).  ``curve_schedule`` times every candidate on the
caller's batch (on a scratch copy of the start values) and keeps the fastest;
results are cached per batch signature (problem count, sample count and size
spread, precision, iteration limit) for the life of the process.
"""
from __future__ import annotations

import math
from typing import Optional

import numpy as np
import torch

THREADS = (64, 128, 256, 512)
GROUPS_PER_CU = (2, 4, 8)
ORDERS = ("index", "largest")

_CACHE: dict = {}


def _signature(batch, options) -> tuple:
    sizes = batch.sizes
    bucket = lambda v: int(round(math.log2(max(float(v), 1.0)) * 4))  # noqa: E731
    return (bucket(batch.num_problems), bucket(sizes.sum()), bucket(sizes.max() if sizes.size else 1),
            bucket(np.median(sizes) if sizes.size else 1), int(options.precision), int(options.max_num_iterations),
            batch.loss_native[0])


def compute_units(device) -> int:
    return int(torch.cuda.get_device_properties(device).multi_processor_count)


def candidates(device) -> list[tuple[tuple[int, int], str]]:
    cus = compute_units(device)
    return [((t, g * cus), o) for t in THREADS for g in GROUPS_PER_CU for o in ORDERS]


def time_schedule(batch, options, schedule, order, repeats: int = 2) -> float:
    """Mean device time (ms) of the batched solve with this shape, from HIP events on the current stream."""
    scratch = torch.zeros(batch.num_problems, 2, dtype=torch.float64, device=batch.device)
    batch.launch(options, scratch.clone(), schedule=schedule, order=order)  # warm-up
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(repeats):
        batch.launch(options, scratch.clone(), schedule=schedule, order=order)
    end.record()
    end.synchronize()
    return start.elapsed_time(end) / repeats


def curve_schedule(batch, options, *, retune: bool = False) -> tuple[tuple[int, int], str]:
    """Fastest ((threads, groups), order) for this batch, timed once per signature."""
    key = _signature(batch, options)
    if not retune and key in _CACHE:
        return _CACHE[key][0]
    best: Optional[tuple] = None
    timings = []
    for sched, order in candidates(batch.device):
        ms = time_schedule(batch, options, sched, order)
        timings.append((sched, order, ms))
        if best is None or ms < best[1]:
            best = ((sched, order), ms)
    _CACHE[key] = (best[0], timings)
    return best[0]


def tuning_table() -> dict:
    """The cached decisions: signature -> (choice, [(schedule, order, ms), ...])."""
    return dict(_CACHE)
