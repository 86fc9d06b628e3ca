"""Continuous-batching solve server on one MI355X.

Requests of three kinds arrive at any time:

* ``CurveRequest``: a curve fit y = exp(m x + c) (robust loss per sample);
* ``BARequest``: a bundle adjustment (cameras, points, observations);
* ``PoseRequest``: an SE2 pose graph.

Admission is by device-memory budget: a request is admitted when the device
bytes it needs (its inputs, tables and the library's exact workspace for it,
from ``ct_ba_batch_problem_bytes`` / ``ct_pose2d_workspace_bytes``) fit beside
the requests already on the GPU, in arrival order.

Curve fits and bundle adjustments of at most 21 cameras are batched
continuously: a worker launches a persistent batched solve of the admitted
requests of its kind that are waiting when its previous batch finishes (up to
a batch limit), so requests that arrive while a batch runs join the next one.
Bundle adjustments have two such workers, so one batch's host tables are
built while another runs on the device, and all batches overlap on the
device.  Pose graphs (and larger bundle
adjustments) run on ``num_slots`` worker threads, each with its own HIP stream
and per-request plan, so their kernels overlap too (the library calls release
the GIL).  Each request's result is the same as solving it alone through the
same path: a batched problem's arithmetic does not depend on the other
problems of its batch, and plans are per request.

Slots run the host-driven LM loop by default (one synchronisation of the
slot's own stream per iteration): a graph capture in one slot thread does not
survive device-memory allocations made meanwhile by other threads, so
concurrent slots capture nothing.  Their kernels still overlap on the device.
"""
from __future__ import annotations

import ctypes
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
from .ba_batch import MAX_CAMERAS, BABatch, BAProblemInput, device_bytes as ba_device_bytes
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


def _ba_terms(num_cameras: int, num_points: int, obs_camera, obs_point) -> int:
    """Observation pairs of a common point with cam(a) <= cam(b) (the camera-pair assembly's terms), counted by
    the library's native table builder; requests with more cameras than a batch holds are bounded instead."""
    oc = np.ascontiguousarray(obs_camera, dtype=np.int32).reshape(-1)
    op = np.ascontiguousarray(obs_point, dtype=np.int32).reshape(-1)
    if num_cameras > MAX_CAMERAS:
        k = np.bincount(op, minlength=num_points).astype(np.int64)
        return int(np.sum(k * k))
    lib = nat.load()
    sizes = np.array([num_cameras, num_points, oc.size], dtype=np.int32)
    T = np.zeros(1, dtype=np.int64)
    nat.check(lib.ct_ba_batch_tables(1, sizes[0:].ctypes.data, sizes[1:].ctypes.data, sizes[2:].ctypes.data,
                                     oc.ctypes.data if oc.size else None, op.ctypes.data if op.size else None,
                                     T.ctypes.data, *([None] * 8)), "ct_ba_batch_tables")
    return int(T[0])


def _pose_pairs(edges, num_nodes: int) -> int:
    e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    if e.size == 0:
        return 0
    return int(np.unique(np.minimum(e[:, 0], e[:, 1]) * num_nodes + np.maximum(e[:, 0], e[:, 1])).size)


def request_bytes(req, options: Optional[SolverOptions] = None) -> int:
    """Device bytes a request holds while it is on the GPU: inputs, tables, parameters, results and the
    library's exact workspace for its solve path."""
    if isinstance(req, CurveRequest):
        return 24 * int(np.asarray(req.x).size) + 128
    if isinstance(req, BARequest):
        C, P, N = len(req.cameras), len(req.points), len(req.obs_camera)
        T = _ba_terms(C, P, req.obs_camera, req.obs_point)
        if C <= MAX_CAMERAS:
            return ba_device_bytes(C, P, N, T)
        # per-request plan: the plan's own workspace query needs the device tables; bound it from its layout
        n = 9 * C
        return 400 * N + 200 * P + 32 * n * n + 16 * T + (1 << 20)
    if isinstance(req, PoseRequest):
        lib = nat.load()
        prob = nat.ct_pose2d_problem()
        prob.num_nodes, prob.num_edges = len(req.poses), len(req.edges)
        prob.num_pairs, prob.fixed_node = _pose_pairs(req.edges, len(req.poses)), int(req.fixed_node)
        opts = lm_options(options or SolverOptions())
        out = ctypes.c_size_t(0)
        nat.check(lib.ct_pose2d_workspace_bytes(ctypes.byref(prob), ctypes.byref(opts), ctypes.byref(out)),
                  "ct_pose2d_workspace_bytes")
        E, N = len(req.edges), len(req.poses)
        tables = 4 * (2 * E + (N + 1) + 2 * E + 3 * prob.num_pairs + 1 + E)
        return int(out.value) + tables + 24 * E + 24 * N
    raise TypeError(f"unknown request type {type(req).__name__}")


estimate_bytes = request_bytes  # the name earlier callers use


class SolveServer:
    """Solve requests concurrently on the GPU; ``submit`` returns a Future of ``Result``."""

    def __init__(self, options: Optional[SolverOptions] = None, *, hbm_budget_bytes: int = 16 << 30,
                 num_slots: int = 4, max_curve_batch: int = 4096, max_ba_batch: int = 64, ba_workers: int = 2,
                 curve_schedule=(64, 0), loop: str = "host"):
        nat.load()
        self.options = options or SolverOptions()
        self.options.validate()
        self.budget = int(hbm_budget_bytes)
        self.in_use = 0
        self.max_curve_batch = int(max_curve_batch)
        self.max_ba_batch = int(max_ba_batch)
        self.slot_workgroups = 64  # persistent workgroups of a slot's reduced-system factorisation
        self.curve_schedule = curve_schedule
        self.loop = loop
        self._lock = threading.Condition()
        self._pending: deque[_Ticket] = deque()       # arrival order, not yet admitted
        self._solo: deque[_Ticket] = deque()          # admitted BA / pose requests
        self._curves: deque[_Ticket] = deque()        # admitted curve requests
        self._bas: deque[_Ticket] = deque()           # admitted bundle adjustments of at most MAX_CAMERAS cameras
        self._closed = False
        self._threads = [threading.Thread(target=self._slot_worker, name=f"slot{i}", daemon=True)
                         for i in range(num_slots)]
        self._threads.append(threading.Thread(target=self._curve_worker, name="curves", daemon=True))
        # two batch workers per kind that batches large numbers: while one batch runs on the device the next
        # one's host tables are built
        self._threads += [threading.Thread(target=self._ba_worker, name=f"bundle-adjustments{i}", daemon=True)
                          for i in range(max(1, int(ba_workers)))]
        for t in self._threads:
            t.start()

    # -- submission and admission ------------------------------------------
    def submit(self, request) -> Future:
        nbytes = request_bytes(request, self.options)
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
            self._queue_for(t.request).append(t)
        self._lock.notify_all()

    def _queue_for(self, req) -> deque:
        if isinstance(req, CurveRequest):
            return self._curves
        if isinstance(req, BARequest) and len(req.cameras) <= MAX_CAMERAS:
            return self._bas
        return self._solo

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
                    res = self._solve_one(t.request, self.slot_workgroups)
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

    def _ba_worker(self) -> None:
        stream = torch.cuda.Stream()
        while True:
            tickets = self._next(self._bas, self.max_ba_batch)
            if not tickets:
                return
            try:
                with torch.cuda.stream(stream):
                    results = self._solve_bas([t.request for t in tickets])
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
    def _lm(self, precision: str = "fp64", solver_workgroups: int = 0):
        return lm_options(self.options, precision=precision, loop=self.loop, linear="dense_schur",
                          solver_workgroups=solver_workgroups)

    def _solve_one(self, req, solver_workgroups: int = 0) -> Result:
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
                # the separator factorisation's column chain keeps few workgroups busy: a share of the chip
                # per solve lets the slots' solves run side by side
                rep, _ = dev.run(self._lm(solver_workgroups=solver_workgroups))
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

    def _solve_bas(self, reqs: list) -> list[Result]:
        """One persistent launch for every bundle adjustment of the batch (ct_ba_batch_solve)."""
        batch = BABatch([BAProblemInput(r.cameras, r.points, r.obs_camera, r.obs_point, r.observations,
                                        native_loss(r.loss)) for r in reqs])
        reps, _, cams, pts = batch.solve(self._lm())
        out = []
        for k in range(len(reqs)):
            c, p = batch.split(cams, pts, k)
            rep = reps[k]
            out.append(Result("ba", np.concatenate([c.reshape(-1), p.reshape(-1)]), float(rep["initial_cost"]),
                              float(rep["final_cost"]), int(rep["termination"]), int(rep["num_records"]) - 1))
        return out

    def _solve_path(self, req) -> Result:
        if isinstance(req, CurveRequest):
            return self._solve_curves([req])[0]
        if isinstance(req, BARequest) and len(req.cameras) <= MAX_CAMERAS:
            return self._solve_bas([req])[0]
        return self._solve_one(req, self.slot_workgroups)

    def solve_plain(self, req) -> Result:
        """The request solved by itself through the per-problem path (one plan per bundle adjustment or pose
        graph, one launch per curve fit), as a caller without the server would: the plain per-problem GPU
        baseline."""
        t0 = time.perf_counter()
        res = self._solve_curves([req])[0] if isinstance(req, CurveRequest) else self._solve_one(req)
        torch.cuda.synchronize()
        res.arrival = res.admitted = t0
        res.finished = time.perf_counter()
        return res

    def solve_alone(self, req) -> Result:
        """The same request solved by itself through the path the server takes for it (no other request in its
        batch, no concurrency), for comparison."""
        t0 = time.perf_counter()
        res = self._solve_path(req)
        torch.cuda.synchronize()
        res.arrival = res.admitted = t0
        res.finished = time.perf_counter()
        return res
