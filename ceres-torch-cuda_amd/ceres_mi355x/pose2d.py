"""SE2 pose graphs on the device (examples/pose_graph_2d.py:6-46).

The reference builds one autograd residual block per edge, holds the first
pose constant and solves with DENSE_QR.  Here the whole LM solve runs in the
HIP library (``ct_pose2d_plan_*``): per-edge residuals and Jacobians, the
damped normal equations assembled per node and per node pair in a fixed
order, and the FP64 Cholesky; the step is the one DENSE_QR takes.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np
import torch

from . import _native as nat
from .losses import LOSS_IDS, native_loss
from .problem import Problem
from .types import LinearSolverType, SolverOptions, SolverSummary


def build_pose_tables(edges, num_nodes: int):
    """Per-node incident edges and per-node-pair edges, tagged ``2*edge + (node is the second end)``."""
    edges = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    E = edges.shape[0]
    if E and (edges.min() < 0 or edges.max() >= num_nodes or np.any(edges[:, 0] == edges[:, 1])):
        raise ValueError("edges must join two distinct nodes in [0, num_nodes)")
    ends = np.concatenate([edges[:, 0], edges[:, 1]])
    tags = np.concatenate([2 * np.arange(E), 2 * np.arange(E) + 1])
    order = np.lexsort((tags, ends))
    node_offsets = np.zeros(num_nodes + 1, dtype=np.int64)
    node_offsets[1:] = np.cumsum(np.bincount(ends, minlength=num_nodes))
    u = np.minimum(edges[:, 0], edges[:, 1])
    v = np.maximum(edges[:, 0], edges[:, 1])
    key = u * num_nodes + v
    porder = np.lexsort((np.arange(E), key))
    uniq, first = np.unique(key[porder], return_index=True)
    pair_offsets = np.append(first, E).astype(np.int64)
    pair_nodes = np.stack([uniq // num_nodes, uniq % num_nodes], axis=1) if E else np.zeros((0, 2), np.int64)
    u_is_b = (edges[porder, 1] == u[porder]).astype(np.int64)
    pair_edges = 2 * porder + u_is_b
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)  # noqa: E731
    return dict(node_offsets=i32(node_offsets), node_edges=i32(tags[order]), pair_nodes=i32(pair_nodes),
                pair_offsets=i32(pair_offsets), pair_edges=i32(pair_edges))


class PoseGraphDevice:
    """Device-resident pose graph: tables, measurements and poses [N, 3] (float64)."""

    def __init__(self, poses, edges, measurements, *, loss_native=("trivial", 0.0, 0.0, 1.0), fixed_node: int = 0,
                 device="cuda"):
        nat.load()
        self.device = torch.device(device)
        poses = torch.as_tensor(poses, dtype=torch.float64)
        self.num_nodes = int(poses.shape[0])
        edges = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
        self.num_edges = int(edges.shape[0])
        tables = build_pose_tables(edges, self.num_nodes)
        dev = self.device
        # one upload for the integer tables and one for the values (views of the two device arrays)
        ints = [("edges", np.ascontiguousarray(edges, dtype=np.int32).reshape(-1))]
        ints += [(k, v.reshape(-1)) for k, v in tables.items()]
        flat = np.concatenate([a for _, a in ints]) if ints else np.zeros(0, np.int32)
        dints = torch.from_numpy(np.ascontiguousarray(flat, dtype=np.int32)).to(dev)
        views, at = {}, 0
        for k, a in ints:
            views[k] = dints[at:at + a.size]
            at += a.size
        self.edges = views.pop("edges").view(-1, 2)
        self.tables = {k: views[k] for k in tables}
        meas = np.asarray(measurements, dtype=np.float64).reshape(-1)
        vals = torch.from_numpy(np.concatenate([meas, poses.detach().cpu().numpy().reshape(-1)])).to(dev)
        self.measurements = vals[:meas.size].view(-1, 3)
        self.poses = vals[meas.size:].view(-1, 3)
        self.num_pairs = int(tables["pair_nodes"].shape[0])
        self.loss_native = loss_native
        self.fixed_node = int(fixed_node)
        self._plans: dict = {}

    def problem_struct(self) -> nat.ct_pose2d_problem:
        p = nat.ct_pose2d_problem()
        p.num_nodes, p.num_edges, p.num_pairs, p.fixed_node = (self.num_nodes, self.num_edges, self.num_pairs,
                                                               self.fixed_node)
        p.edges = nat.ptr(self.edges)
        p.measurements = nat.ptr(self.measurements)
        for name in ("node_offsets", "node_edges", "pair_nodes", "pair_offsets", "pair_edges"):
            setattr(p, name, nat.ptr(self.tables[name]))
        kind, a, b, scale = self.loss_native
        p.loss = nat.ct_loss(LOSS_IDS[kind], float(a), float(b), float(scale))
        return p

    def set_poses(self, poses) -> None:
        self.poses.copy_(torch.as_tensor(poses, dtype=torch.float64).to(self.device))

    def _plan(self, opts: nat.ct_lm_options):
        lib = nat.load()
        key = tuple(getattr(opts, f) for f, _ in nat.ct_lm_options._fields_)
        prob = self.problem_struct()
        plan = self._plans.get(key)
        if plan is None:
            nbytes = ctypes.c_size_t(0)
            nat.check(lib.ct_pose2d_workspace_bytes(ctypes.byref(prob), ctypes.byref(opts), ctypes.byref(nbytes)),
                      "ct_pose2d_workspace_bytes")
            ws = torch.zeros(int(nbytes.value), dtype=torch.uint8, device=self.device)
            handle = ctypes.c_void_p()
            nat.check(lib.ct_pose2d_plan_create(ctypes.byref(prob), ctypes.byref(opts), nat.ptr(self.poses),
                                                nat.ptr(ws), ws.numel(), nat.current_stream_handle(),
                                                ctypes.byref(handle)),
                      "ct_pose2d_plan_create")
            plan = (handle, ws, prob)
            self._plans[key] = plan
        return plan

    def system_size(self, opts: nat.ct_lm_options) -> int:
        """Unknowns of the system the plan for ``opts`` factors densely per iteration."""
        n = ctypes.c_int32(0)
        nat.check(nat.load().ct_pose2d_plan_system_size(self._plan(opts)[0], ctypes.byref(n)),
                  "ct_pose2d_plan_system_size")
        return int(n.value)

    def run(self, opts: nat.ct_lm_options):
        """One LM solve in place on self.poses; returns (report, records)."""
        lib = nat.load()
        plan = self._plan(opts)
        rep = nat.ct_lm_report()
        rec = (ctypes.c_double * ((opts.max_num_iterations + 2) * nat.CT_REC_FIELDS))()
        nat.check(lib.ct_pose2d_plan_run(plan[0], ctypes.byref(rep), rec), "ct_pose2d_plan_run")
        records = np.frombuffer(rec, dtype=np.float64).reshape(-1, nat.CT_REC_FIELDS)[: rep.num_records].copy()
        return rep, records

    def close(self) -> None:
        lib = nat.load()
        for handle, _ws, _p in self._plans.values():
            lib.ct_pose2d_plan_destroy(handle)
        self._plans.clear()

    def __del__(self):  # pragma: no cover - interpreter shutdown order varies
        try:
            self.close()
        except Exception:
            pass


class _PoseExtract:
    def __init__(self, problem: Problem):
        rbs = problem._residual_blocks
        if not rbs or any(rb.cost_function.device_kind != "pose2d" for rb in rbs):
            raise NotImplementedError("not a pose-graph problem")
        losses = {native_loss(rb.loss_function) for rb in rbs}
        if len(losses) != 1:
            raise NotImplementedError("the device path needs one loss function for all residual blocks")
        self.loss = losses.pop()
        index, self.blocks, edges, meas = {}, [], [], []
        for rb in rbs:
            ids = []
            for blk in rb.parameter_blocks:
                if blk not in index:
                    index[blk] = len(self.blocks)
                    self.blocks.append(blk)
                ids.append(index[blk])
            edges.append(ids)
            meas.append(rb.cost_function.measurement)
        self.edges = np.asarray(edges, dtype=np.int64)
        self.meas = np.asarray(meas, dtype=np.float64)
        fixed = [i for i, b in enumerate(self.blocks) if b.constant]
        partial = [b for b in self.blocks if not b.constant and not all(b.free_mask())]
        if len(fixed) > 1 or partial:
            raise NotImplementedError("the device pose graph holds at most one whole pose constant")
        self.fixed = fixed[0] if fixed else -1
        for b in self.blocks:
            if b.lower_bound is not None or b.upper_bound is not None:
                raise NotImplementedError("parameter bounds are not on the MI355X device path")
            if b.tensor.dtype != torch.float64:
                raise NotImplementedError("pose parameters must be float64 tensors")

    def poses(self) -> torch.Tensor:
        return torch.stack([b.tensor.detach().reshape(-1).to("cpu") for b in self.blocks])

    def write_back(self, poses: torch.Tensor) -> None:
        host = poses.to("cpu")
        with torch.no_grad():
            for i, b in enumerate(self.blocks):
                b.tensor.reshape(-1).copy_(host[i].to(b.tensor.device))


def solve_pose_problem(options: SolverOptions, problem: Problem, start: float) -> SolverSummary:
    """``solve`` for a problem made of Pose2DRelativeError blocks."""
    from .ba import lm_options
    from .solver import DEVICE_SETTINGS, summary_from_report

    if options.linear_solver_type in (LinearSolverType.ITERATIVE_SCHUR, LinearSolverType.CGNR):
        raise NotImplementedError("the device pose graph solves the damped normal equations exactly")
    if options.use_mixed_precision_solves:
        raise NotImplementedError("the device pose graph runs in FP64")
    ex = _PoseExtract(problem)
    cached = getattr(problem, "_pose_device_cache", None)
    sig = (len(ex.blocks), ex.edges.shape[0], ex.fixed, ex.loss)
    if cached is not None and cached[0] == sig:
        dev = cached[1]
        dev.set_poses(ex.poses())
    else:
        dev = PoseGraphDevice(ex.poses(), ex.edges, ex.meas, loss_native=ex.loss, fixed_node=ex.fixed)
        problem._pose_device_cache = (sig, dev)
    opts = lm_options(options, precision="fp64", loop=DEVICE_SETTINGS["loop"], linear="dense_schur",
                      iterations_per_graph=DEVICE_SETTINGS["iterations_per_graph"])
    pre = time.perf_counter()
    rep, records = dev.run(opts)
    torch.cuda.synchronize()
    s = summary_from_report(options, problem, rep, records, start, "fp64")
    s.preprocessor_time_in_seconds = pre - start
    ex.write_back(dev.poses)
    s.total_time_in_seconds = time.perf_counter() - start
    return s
