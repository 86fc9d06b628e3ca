"""Batched curve fits ``y = exp(m x + c)`` on the device.

The reference solves one curve-fit ``Problem`` at a time: one autograd
residual block per sample, a robust loss on each, the trust-region loop of
``solver.py:35-357`` and a dense QR step (``examples/robust_curve_fitting.py:80-97``,
``linear.py:359-371``).  Here many such problems of uneven size are solved in
one persistent device launch (``ct_curve_solve``): a team (a whole workgroup
for the leading problems of the processing order, one wave for the trailing
ones) runs the whole LM loop of one problem at a time, taking problems from
work queues in a chosen order.  The launch shape (workgroup size, number of
persistent workgroups, how many problems go one per wave, order) comes from
``ceres_mi355x.autotune``.  The mixed form reads FP32 copies of the samples.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np
import torch

from . import _native as nat
from .losses import LOSS_IDS, native_loss
from .problem import Problem
from .types import SolverOptions, SolverSummary


@dataclass
class CurveResult:
    params: torch.Tensor              # [B, 2] (m, c), best values
    reports: np.ndarray               # [B] structured (REPORT_DTYPE)
    records: Optional[np.ndarray]     # [B, max_iter + 2, CT_REC_FIELDS] or None
    seconds: float                    # host wall time of the launch, synchronised


class CurveBatch:
    """Device-resident samples of many curve-fit problems.

    ``offsets`` [B+1] int64 (problem p owns samples offsets[p]..offsets[p+1]),
    ``x`` / ``y`` [S] float64, kept with FP32 copies for the mixed form (which
    solves the FP32-rounded samples).  The batch is validated once on the host.
    """

    def __init__(self, offsets, x, y, loss=None, device="cuda"):
        nat.load()
        self.device = torch.device(device)
        self.offsets = torch.as_tensor(np.asarray(offsets, dtype=np.int64)).to(self.device)
        self.x = torch.as_tensor(np.asarray(x, dtype=np.float64)).to(self.device)
        self.y = torch.as_tensor(np.asarray(y, dtype=np.float64)).to(self.device)
        self.x32 = self.x.to(torch.float32)
        self.y32 = self.y.to(torch.float32)
        self.num_problems = int(self.offsets.numel()) - 1
        self.loss_native = native_loss(loss)
        self.sizes = np.diff(np.asarray(offsets, dtype=np.int64))
        self._order_cache: dict = {}
        self._workspace = torch.zeros(256, dtype=torch.uint8, device=self.device)
        bad = ctypes.c_int32(-1)
        # ct_curve_validate reads the offsets with a synchronous copy, which does not wait for a
        # non-blocking stream: the uploads above (on the current stream) finish first
        torch.cuda.current_stream(self.device).synchronize()
        nat.check(nat.load().ct_curve_validate(ctypes.byref(self._struct(None)), ctypes.byref(bad)),
                  "ct_curve_validate")

    def order(self, kind: str) -> Optional[torch.Tensor]:
        """Processing order: 'index' (None) or 'largest' (decreasing sample count)."""
        if kind == "index":
            return None
        if kind not in self._order_cache:
            if kind != "largest":
                raise ValueError("order must be 'index' or 'largest'")
            perm = np.argsort(-self.sizes, kind="stable").astype(np.int32)
            self._order_cache[kind] = torch.from_numpy(perm).to(self.device)
        return self._order_cache[kind]

    def _struct(self, order: Optional[torch.Tensor]) -> nat.ct_curve_batch:
        b = nat.ct_curve_batch()
        b.num_problems = self.num_problems
        b.num_samples = int(self.x.numel())
        b.offsets = nat.ptr(self.offsets)
        b.x = nat.ptr(self.x)
        b.y = nat.ptr(self.y)
        b.x32 = nat.ptr(self.x32)
        b.y32 = nat.ptr(self.y32)
        b.order = nat.ptr(order)
        kind, a, bb, scale = self.loss_native
        b.loss = nat.ct_loss(LOSS_IDS[kind], float(a), float(bb), float(scale))
        return b

    def launch(self, options: nat.ct_lm_options, params: torch.Tensor, *, schedule=None, order: str = "index",
               results: Optional[torch.Tensor] = None, records: Optional[torch.Tensor] = None) -> None:
        """Enqueue one batched solve on the current stream (no synchronisation; graph-capturable).

        ``schedule`` is (threads, groups) or (threads, groups, wave_problems).
        """
        shape = tuple(schedule) if schedule is not None else (256, 0)
        threads, groups, wave_problems = shape + (0,) * (3 - len(shape))
        sched = nat.ct_batch_schedule(int(threads), int(groups), int(wave_problems))
        nat.check(nat.load().ct_curve_solve(ctypes.byref(self._struct(self.order(order))), ctypes.byref(options),
                                            ctypes.byref(sched), nat.ptr(params), nat.ptr(results),
                                            nat.ptr(records), nat.ptr(self._workspace), self._workspace.numel(),
                                            nat.current_stream_handle()),
                  "ct_curve_solve")

    def solve(self, options: nat.ct_lm_options, params=None, *, schedule=None, order: Optional[str] = None,
              records: bool = False) -> CurveResult:
        """Solve every problem from ``params`` ([B,2], default m = c = 0); returns best values and reports."""
        B = self.num_problems
        if params is None:
            p = torch.zeros(B, 2, dtype=torch.float64, device=self.device)
        else:
            p = torch.as_tensor(params, dtype=torch.float64).to(self.device).reshape(B, 2).clone()
        if schedule is None or order is None:
            from .autotune import curve_schedule

            tuned = curve_schedule(self, options)
            schedule = schedule if schedule is not None else tuned[0]
            order = order if order is not None else tuned[1]
        res = torch.zeros(B * ctypes.sizeof(nat.ct_lm_report), dtype=torch.uint8, device=self.device)
        rec = None
        if records:
            rec = torch.zeros(B * (options.max_num_iterations + 2) * nat.CT_REC_FIELDS, dtype=torch.float64,
                              device=self.device)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        self.launch(options, p, schedule=schedule, order=order, results=res, records=rec)
        torch.cuda.synchronize()
        seconds = time.perf_counter() - t0
        reports = np.frombuffer(res.cpu().numpy().tobytes(), dtype=nat.REPORT_DTYPE).copy()
        recs = None
        if rec is not None:
            recs = rec.cpu().numpy().reshape(B, options.max_num_iterations + 2, nat.CT_REC_FIELDS)
        return CurveResult(params=p, reports=reports, records=recs, seconds=seconds)


# ---------------------------------------------------------------------------
# Problem-level API
# ---------------------------------------------------------------------------


class _CurveExtract:
    """Samples and the (m, c) parameter blocks of a Problem made of ExpCurveResidual blocks."""

    def __init__(self, problem: Problem):
        rbs = problem._residual_blocks
        if not rbs or any(rb.cost_function.device_kind != "curve" for rb in rbs):
            raise NotImplementedError("not a curve-fit problem")
        blocks = {tuple(rb.parameter_blocks) for rb in rbs}
        if len(blocks) != 1:
            raise NotImplementedError("the device curve fit needs every residual block on the same (m, c) blocks")
        self.m_block, self.c_block = next(iter(blocks))
        losses = {native_loss(rb.loss_function) for rb in rbs}
        if len(losses) != 1:
            raise NotImplementedError("the device path needs one loss function for all residual blocks")
        self.loss = losses.pop()
        for b in (self.m_block, self.c_block):
            if b.lower_bound is not None or b.upper_bound is not None:
                raise NotImplementedError("parameter bounds are not on the MI355X device path")
            if b.constant or not all(b.free_mask()):
                raise NotImplementedError("constant curve parameters are not on the MI355X device path")
        self.x = np.array([rb.cost_function.x for rb in rbs], dtype=np.float64)
        self.y = np.array([rb.cost_function.y for rb in rbs], dtype=np.float64)

    def start(self) -> list[float]:
        return [float(self.m_block.tensor.detach().reshape(-1)[0]), float(self.c_block.tensor.detach().reshape(-1)[0])]

    def write_back(self, m: float, c: float) -> None:
        with torch.no_grad():
            self.m_block.tensor.reshape(-1)[0] = m
            self.c_block.tensor.reshape(-1)[0] = c


class _Report:
    """Attribute view of one REPORT_DTYPE row (the shape summary_from_report reads)."""

    def __init__(self, row):
        for name in nat.REPORT_DTYPE.names:
            setattr(self, name, row[name])


def _curve_options(options: SolverOptions):
    from .ba import lm_options

    precision = "mixed" if options.use_mixed_precision_solves else "fp64"
    return lm_options(options, precision=precision, loop="device", linear="dense_schur"), precision


def solve_batch(options: SolverOptions, problems: Sequence[Problem]) -> list[SolverSummary]:
    """Solve many curve-fit problems with one device launch; one SolverSummary per problem.

    Problems that are not curve fits are solved one at a time with ``solve``.
    """
    from .solver import _check_supported, solve, summary_from_report

    options.validate()
    _check_supported(options)
    start = time.perf_counter()
    out: list[Optional[SolverSummary]] = [None] * len(problems)
    curve_idx, extracts = [], []
    for i, pr in enumerate(problems):
        kinds = {rb.cost_function.device_kind for rb in pr._residual_blocks}
        if kinds == {"curve"}:
            curve_idx.append(i)
            extracts.append(_CurveExtract(pr))
        else:
            out[i] = solve(options, pr)
    if curve_idx:
        losses = {e.loss for e in extracts}
        if len(losses) != 1:
            raise NotImplementedError("one batched launch needs the same loss in every problem")
        sizes = np.array([e.x.size for e in extracts], dtype=np.int64)
        offsets = np.zeros(len(extracts) + 1, dtype=np.int64)
        offsets[1:] = np.cumsum(sizes)
        batch = CurveBatch(offsets, np.concatenate([e.x for e in extracts]), np.concatenate([e.y for e in extracts]),
                           loss=None, device="cuda")
        batch.loss_native = losses.pop()
        opts, precision = _curve_options(options)
        res = batch.solve(opts, np.array([e.start() for e in extracts]), records=True)
        params = res.params.cpu().numpy()
        for k, (i, e) in enumerate(zip(curve_idx, extracts)):
            rep = _Report(res.reports[k])
            rows = res.records[k][: int(rep.num_records)]
            s = summary_from_report(options, problems[i], rep, rows, start, precision)
            e.write_back(float(params[k, 0]), float(params[k, 1]))
            out[i] = s
    return out  # type: ignore[return-value]


def solve_curve_problem(options: SolverOptions, problem: Problem, start: float) -> SolverSummary:
    """``solve`` for one curve-fit problem (a batch of one)."""
    s = solve_batch(options, [problem])[0]
    s.total_time_in_seconds = time.perf_counter() - start
    return s
