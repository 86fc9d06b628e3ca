"""Continuous-batching solve server on one MI355X.

Requests of three kinds arrive at any time:

* ``CurveRequest``: a curve fit y = exp(m x + c) (robust loss per sample);
* ``BARequest``: a bundle adjustment (cameras, points, observations);
* ``PoseRequest``: an SE2 pose graph.

Admission is by device-memory budget: a request is admitted when its
estimated working set fits beside the requests already on the GPU, in arrival
order.  Bundle adjustments and pose graphs run on ``num_slots`` worker
threads, each with its own HIP stream, so their kernels overlap on the
device (the library calls release the GIL).  Curve fits are batched
continuously: one worker launches a persistent batched solve of every curve
request that is waiting when the previous batch finishes, so requests that
arrive while a batch runs join the next one.  Each request's result is the
same as solving it alone (per-problem arithmetic does not depend on the other
problems of a batch with a fixed launch shape; plans are per request).

Slots run the host-driven LM loop by default (one synchronisation of the
slot's own stream per iteration): a graph capture in one slot thread does not
survive device-memory allocations made meanwhile by other slot threads, so
concurrent slots capture nothing.  Their kernels still overlap on the device.
"""
from __future__ import annotations

import threading
import time
from collections import deque
from concurrent.futures import Future
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _native as nat
from .ba import BADevice, lm_options
from .curve import CurveBatch
from .losses import native_loss
from .pose2d import PoseGraphDevice
from .types import SolverOptions


@dataclass
class CurveRequest:
    x: np.ndarray
    y: np.ndarray
    start: tuple = (0.0, 0.0)
    loss: object = None


@dataclass
class BARequest:
    cameras: np.ndarray       # [C, 9]
    points: np.ndarray        # [P, 3]
    obs_camera: np.ndarray    # [N]
    obs_point: np.ndarray     # [N]
    observations: np.ndarray  # [N, 2]
    loss: object = None


@dataclass
class PoseRequest:
    poses: np.ndarray         # [N, 3]
    edges: np.ndarray         # [E, 2]
    measurements: np.ndarray  # [E, 3]
    loss: object = None
    fixed_node: int = 0


@dataclass
class Result:
    kind: str
    params: np.ndarray        # curve [2]; BA [C*9 + P*3]; pose [N, 3]
    initial_cost: float
    final_cost: float
    termination: int
    iterations: int
    arrival: float = 0.0      # host clock (perf_counter)
    admitted: float = 0.0
    finished: float = 0.0

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


@dataclass
class _Ticket:
    request: object
    future: Future
    arrival: float
    nbytes: int
    admitted: float = 0.0


def estimate_bytes(req) -> int:
    """Device working set of a request, generously rounded: inputs, Jacobians, the dense normal matrix, its
    factor's inverse and scratch."""
    if isinstance(req, CurveRequest):
        return 64 * int(np.asarray(req.x).size) + 4096
    if isinstance(req, BARequest):
        C, P, N = len(req.cameras), len(req.points), len(req.obs_camera)
        n = 9 * C
        pairs = int(np.sum(np.bincount(np.asarray(req.obs_point), minlength=P) ** 2))
        return 400 * N + 200 * P + 32 * n * n + 16 * pairs + (1 << 20)
    if isinstance(req, PoseRequest):
        n = 3 * len(req.poses)
        return 200 * len(req.edges) + 100 * len(req.poses) + 32 * n * n + (1 << 20)
    raise TypeError(f"unknown request type {type(req).__name__}")


class SolveServer:
    """Solve requests concurrently on the GPU; ``submit`` returns a Future of ``Result``."""

    def __init__(self, options: Optional[SolverOptions] = None, *, hbm_budget_bytes: int = 16 << 30,
                 num_slots: int = 4, max_curve_batch: int = 4096, curve_schedule=(64, 0), loop: str = "host"):
        nat.load()
        self.options = options or SolverOptions()
        self.options.validate()
        self.budget = int(hbm_budget_bytes)
        self.in_use = 0
        self.max_curve_batch = int(max_curve_batch)
        self.curve_schedule = curve_schedule
        self.loop = loop
        self._lock = threading.Condition()
        self._pending: deque[_Ticket] = deque()       # arrival order, not yet admitted
        self._solo: deque[_Ticket] = deque()          # admitted BA / pose requests
        self._curves: deque[_Ticket] = deque()        # admitted curve requests
        self._closed = False
        self._threads = [threading.Thread(target=self._slot_worker, name=f"slot{i}", daemon=True)
                         for i in range(num_slots)]
        self._threads.append(threading.Thread(target=self._curve_worker, name="curves", daemon=True))
        for t in self._threads:
            t.start()

    # -- submission and admission ------------------------------------------
    def submit(self, request) -> Future:
        nbytes = estimate_bytes(request)
        if nbytes > self.budget:
            raise ValueError(f"request needs ~{nbytes} bytes, more than the whole budget ({self.budget})")
        fut: Future = Future()
        with self._lock:
            if self._closed:
                raise RuntimeError("server is closed")
            self._pending.append(_Ticket(request, fut, time.perf_counter(), nbytes))
            self._admit_locked()
        return fut

    def _admit_locked(self) -> None:
        # in arrival order while the head fits in the budget
        while self._pending and self.in_use + self._pending[0].nbytes <= self.budget:
            t = self._pending.popleft()
            t.admitted = time.perf_counter()
            self.in_use += t.nbytes
            (self._curves if isinstance(t.request, CurveRequest) else self._solo).append(t)
        self._lock.notify_all()

    def _release(self, nbytes: int) -> None:
        with self._lock:
            self.in_use -= nbytes
            self._admit_locked()

    def close(self) -> None:
        with self._lock:
            self._closed = True
            self._lock.notify_all()
        for t in self._threads:
            t.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- workers -------------------------------------------------------------
    def _next(self, queue: deque, batch: int = 1) -> list[_Ticket]:
        with self._lock:
            while not queue and not (self._closed and not self._pending):
                self._lock.wait()
            out = []
            while queue and len(out) < batch:
                out.append(queue.popleft())
            return out

    def _slot_worker(self) -> None:
        stream = torch.cuda.Stream()
        while True:
            tickets = self._next(self._solo)
            if not tickets:
                return
            t = tickets[0]
            try:
                with torch.cuda.stream(stream):
                    res = self._solve_one(t.request)
                    torch.cuda.current_stream().synchronize()
                res.arrival, res.admitted, res.finished = t.arrival, t.admitted, time.perf_counter()
                t.future.set_result(res)
            except Exception as exc:  # noqa: BLE001 - delivered to the caller
                t.future.set_exception(exc)
            finally:
                self._release(t.nbytes)

    def _curve_worker(self) -> None:
        stream = torch.cuda.Stream()
        while True:
            tickets = self._next(self._curves, self.max_curve_batch)
            if not tickets:
                return
            try:
                with torch.cuda.stream(stream):
                    results = self._solve_curves([t.request for t in tickets])
                    torch.cuda.current_stream().synchronize()
                done = time.perf_counter()
                for t, res in zip(tickets, results):
                    res.arrival, res.admitted, res.finished = t.arrival, t.admitted, done
                    t.future.set_result(res)
            except Exception as exc:  # noqa: BLE001
                for t in tickets:
                    t.future.set_exception(exc)
            finally:
                self._release(sum(t.nbytes for t in tickets))

    # -- solves ----------------------------------------------------------------
    def _lm(self, precision: str = "fp64"):
        return lm_options(self.options, precision=precision, loop=self.loop, linear="dense_schur")

    def _solve_one(self, req) -> Result:
        if isinstance(req, BARequest):
            dev = BADevice(torch.as_tensor(req.cameras, dtype=torch.float64),
                           torch.as_tensor(req.points, dtype=torch.float64), req.obs_camera, req.obs_point,
                           req.observations, loss_native=native_loss(req.loss))
            try:
                rep, _ = dev.run(self._lm())
                params = torch.cat([dev.cameras.reshape(-1), dev.points.reshape(-1)]).cpu().numpy()
            finally:
                dev.close()
            return Result("ba", params, rep.initial_cost, rep.final_cost, rep.termination, rep.num_records - 1)
        if isinstance(req, PoseRequest):
            dev = PoseGraphDevice(req.poses, req.edges, req.measurements, loss_native=native_loss(req.loss),
                                  fixed_node=req.fixed_node)
            try:
                rep, _ = dev.run(self._lm())
                params = dev.poses.cpu().numpy()
            finally:
                dev.close()
            return Result("pose2d", params, rep.initial_cost, rep.final_cost, rep.termination, rep.num_records - 1)
        raise TypeError(f"unknown request type {type(req).__name__}")

    def _solve_curves(self, reqs: list) -> list[Result]:
        losses = {native_loss(r.loss) for r in reqs}
        out: list[Optional[Result]] = [None] * len(reqs)
        for loss in losses:  # one batched launch per loss kind
            idx = [i for i, r in enumerate(reqs) if native_loss(r.loss) == loss]
            sizes = np.array([np.asarray(reqs[i].x).size for i in idx], dtype=np.int64)
            offsets = np.zeros(len(idx) + 1, dtype=np.int64)
            offsets[1:] = np.cumsum(sizes)
            batch = CurveBatch(offsets, np.concatenate([np.asarray(reqs[i].x, np.float64) for i in idx]),
                               np.concatenate([np.asarray(reqs[i].y, np.float64) for i in idx]), loss=None)
            batch.loss_native = loss
            params = torch.as_tensor(np.array([reqs[i].start for i in idx], dtype=np.float64)).to(batch.device)
            reports = torch.zeros(len(idx) * nat.REPORT_DTYPE.itemsize, dtype=torch.uint8, device=batch.device)
            # launch on this worker's stream and wait for that stream only
            batch.launch(self._lm(), params, schedule=self.curve_schedule, order="largest", results=reports)
            torch.cuda.current_stream().synchronize()
            host = params.cpu().numpy()
            reps = np.frombuffer(reports.cpu().numpy().tobytes(), dtype=nat.REPORT_DTYPE)
            for k, i in enumerate(idx):
                rep = reps[k]
                out[i] = Result("curve", host[k], float(rep["initial_cost"]), float(rep["final_cost"]),
                                int(rep["termination"]), int(rep["num_records"]) - 1)
        return out  # type: ignore[return-value]

    def solve_alone(self, req) -> Result:
        """The same request solved by itself (no batching, no concurrency), for comparison."""
        t0 = time.perf_counter()
        res = self._solve_curves([req])[0] if isinstance(req, CurveRequest) else self._solve_one(req)
        torch.cuda.synchronize()
        res.arrival = res.admitted = t0
        res.finished = time.perf_counter()
        return res
