"""Launch-shape autotuner for the persistent batched solves.

A batched launch is shaped by the workgroup size (threads per problem), the
number of persistent workgroups pulling problems from the work queues, the
order in which problems are queued (index order, or largest first, which
balances uneven sizes) and a size split: with the largest-first order, the
problems of at most a threshold number of samples are solved by one wave each
and only the larger ones by a whole workgroup.  ``curve_schedule`` times every candidate on the
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
GROUPS_PER_CU = (2, 4, 8, 16)
ORDERS = ("index", "largest")
SPLIT_SAMPLES = (128, 256, 512, 1024)  # problems of at most this many samples go one per wave
MAX_THREADS_PER_CU = 2048              # 32 waves resident per compute unit

_CACHE: dict = {}


def _signature(batch, options) -> tuple:
    sizes = batch.sizes
    bucket = lambda v: int(round(math.log2(max(float(v), 1.0)) * 4))  # noqa: E731
    return (bucket(batch.num_problems), bucket(sizes.sum()), bucket(sizes.max() if sizes.size else 1),
            bucket(np.median(sizes) if sizes.size else 1), int(options.precision), int(options.max_num_iterations),
            batch.loss_native[0])


def compute_units(device) -> int:
    return int(torch.cuda.get_device_properties(device).multi_processor_count)


def candidates(batch) -> list[tuple[tuple[int, int, int], str]]:
    """((threads, groups, wave_limit), order) for every shape worth timing on this batch.

    wave_limit > 0 (largest-first order only): problems of at most that many
    samples are solved one per wave; 0: every problem by a whole workgroup.
    """
    cus = compute_units(batch.device)
    out = []
    for t in THREADS:
        for g in GROUPS_PER_CU:
            if t * g > MAX_THREADS_PER_CU:
                continue
            out += [((t, g * cus, 0), o) for o in ORDERS]
            if t == 64:
                continue  # a 64-thread workgroup is one wave already
            for limit in SPLIT_SAMPLES:
                if 0 < wave_problems(batch, limit) < batch.num_problems:
                    out.append(((t, g * cus, limit), "largest"))
    return out


def wave_problems(batch, limit: int) -> int:
    """Problems of at most ``limit`` samples: the trailing slots of the largest-first order."""
    return int(np.sum(batch.sizes <= limit)) if limit > 0 else 0


def launch_shape(batch, choice) -> tuple[tuple[int, int, int], str]:
    """A tuned ((threads, groups, wave_limit), order) as the launch schedule of this batch."""
    (threads, groups, limit), order = choice
    return (threads, groups, wave_problems(batch, limit)), order


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


def curve_schedule(batch, options, *, retune: bool = False) -> tuple[tuple[int, int, int], str]:
    """Fastest ((threads, groups, wave_problems), order) for this batch.

    Candidates are timed once per batch signature; the cache keeps the size
    limit of the wave split, which is turned into this batch's problem count.
    """
    key = _signature(batch, options)
    if retune or key not in _CACHE:
        best: Optional[tuple] = None
        timings = []
        for choice in candidates(batch):
            sched, order = launch_shape(batch, choice)
            ms = time_schedule(batch, options, sched, order)
            timings.append((choice, ms))
            if best is None or ms < best[1]:
                best = (choice, ms)
        _CACHE[key] = (best[0], timings)
    return launch_shape(batch, _CACHE[key][0])


def tuning_table() -> dict:
    """The cached decisions: signature -> (choice, [(choice, ms), ...]), choice = ((threads, groups,
    wave_limit), order)."""
    return dict(_CACHE)
