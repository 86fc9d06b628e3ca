"""``solve(options, problem)`` (``src/ceres_torch/solver.py:28-32``) on the device.

The reference runs the trust-region loop of ``solver.py:35-357`` in Python
with one autograd call per residual block.  Here ``solve`` recognises a
problem made of typed residual blocks, moves it to the device once and runs
the whole LM loop in the HIP library; the summary it returns has the
reference's fields and counters.  Problems or options the device path does not
cover raise ``NotImplementedError`` naming what is unsupported; there is no
silent fallback.
"""
from __future__ import annotations

import time
from typing import Optional

import numpy as np
import torch

from . import _native as nat
from .ba import PRECONDITIONERS, BADevice, line_search_options, lm_options
from .losses import native_loss
from .problem import EvaluationResult, Problem
from .types import (
    IterationSummary,
    LinearSolverType,
    MinimizerType,
    SolverOptions,
    SolverSummary,
    TerminationType,
    TrustRegionStrategyType,
)

_TERMINATION = {0: TerminationType.CONVERGENCE, 1: TerminationType.NO_CONVERGENCE, 2: TerminationType.FAILURE}
_MESSAGES = {
    1: "Gradient tolerance reached.",
    2: "Parameter tolerance reached.",
    3: "Function tolerance reached.",
    4: "Minimum trust-region radius reached.",
    5: "Maximum iterations reached.",
    6: "Too many invalid trust-region steps.",
    7: "No active parameters.",
    8: "Invalid problem input.",
    9: "Maximum solver time reached.",
    10: "Projected gradient tolerance reached.",
    11: "Line search failed to find a decreasing step.",
}

# Device-path settings with no counterpart in SolverOptions.
DEVICE_SETTINGS = dict(loop="device", iterations_per_graph=8)


def set_device_loop(mode: str, iterations_per_graph: Optional[int] = None) -> None:
    """'device' (graph-captured, device-resident LM state) or 'host' (one sync per iteration)."""
    if mode not in ("device", "host"):
        raise ValueError("mode must be 'device' or 'host'")
    DEVICE_SETTINGS["loop"] = mode
    if iterations_per_graph is not None:
        DEVICE_SETTINGS["iterations_per_graph"] = int(iterations_per_graph)


def _check_supported(options: SolverOptions) -> None:
    if options.use_nonmonotonic_steps:
        raise NotImplementedError("non-monotonic steps are not on the MI355X device path")
    if options.use_inner_iterations:
        raise NotImplementedError("inner iterations are not on the MI355X device path")
    if options.callbacks:
        raise NotImplementedError("iteration callbacks are not on the MI355X device path")
    if options.check_gradients:
        raise NotImplementedError("check_gradients is not on the MI355X device path")
    if options.linear_solver_type is LinearSolverType.CGNR:
        raise NotImplementedError("CGNR is not on the MI355X device path")
    if (options.linear_solver_type is LinearSolverType.ITERATIVE_SCHUR
            and options.preconditioner_type.name not in PRECONDITIONERS):
        raise NotImplementedError(f"ITERATIVE_SCHUR on the MI355X device path takes the SCHUR_JACOBI, JACOBI or "
                                  f"IDENTITY preconditioner, not {options.preconditioner_type.name}")


def _linear_kind(options: SolverOptions) -> str:
    if options.linear_solver_type is LinearSolverType.ITERATIVE_SCHUR:
        return "iterative_schur"
    # DENSE_SCHUR / SPARSE_SCHUR, and the exact dense solvers, all produce the
    # exact solution of the damped normal equations: one Schur + Cholesky path.
    return "dense_schur"


def _effective_solver(t: LinearSolverType) -> LinearSolverType:
    """solver.py:788-795."""
    if t is LinearSolverType.SPARSE_NORMAL_CHOLESKY:
        return LinearSolverType.DENSE_NORMAL_CHOLESKY
    if t is LinearSolverType.SPARSE_SCHUR:
        return LinearSolverType.DENSE_SCHUR
    return t


# ---------------------------------------------------------------------------
# bundle adjustment extraction
# ---------------------------------------------------------------------------


class _QuaternionCamera:
    """The two parameter blocks of a quaternion camera (QuaternionReprojectionError): the rotation
    (4 values, QuaternionManifold or EigenQuaternionManifold) and t, f, k1, k2 (6 values).  To the
    device they are one camera of 10 stored values and 9 tangent coordinates (rotation first)."""

    def __init__(self, rotation, rest, order: str):
        from .manifolds import EigenQuaternionManifold, QuaternionManifold

        want = EigenQuaternionManifold if order == "eigen" else QuaternionManifold
        m = rotation.manifold
        if not isinstance(m, want) or (order == "ceres" and isinstance(m, EigenQuaternionManifold)):
            raise NotImplementedError(f"the rotation block of a QuaternionReprojectionError(order={order!r}) needs "
                                      f"a {want.__name__} on the MI355X device path")
        if rest.size != 6:
            raise NotImplementedError("the camera block of a QuaternionReprojectionError has 6 values")
        self.rotation, self.rest = rotation, rest

    @property
    def blocks(self):
        return (self.rotation, self.rest)

    @property
    def tangent_size(self) -> int:
        return self.rotation.tangent_size + self.rest.tangent_size

    def free_mask(self) -> list[bool]:
        """Tangent coordinates that move: the rotation's three, then the rest block's six."""
        return [not self.rotation.constant] * 3 + list(self.rest.free_mask())

    def values(self) -> torch.Tensor:
        return torch.cat([self.rotation.tensor.detach().reshape(-1), self.rest.tensor.detach().reshape(-1)])

    def write(self, v: torch.Tensor) -> None:
        self.rotation.tensor.reshape(-1).copy_(v[:4].to(self.rotation.tensor.device))
        self.rest.tensor.reshape(-1).copy_(v[4:].to(self.rest.tensor.device))


class _BAExtract:
    def __init__(self, problem: Problem):
        self.problem = problem
        self.camera_model = "angle_axis"
        if problem.bulk_ba and problem._residual_blocks:
            raise NotImplementedError("mixing bulk and per-block bundle adjustment residuals is not supported")
        if problem.bulk_ba:
            if len(problem.bulk_ba) != 1:
                raise NotImplementedError("one bulk bundle-adjustment group per problem")
            bulk = problem.bulk_ba[0]
            self.cam_blocks = [problem._row_blocks[(id(bulk.cameras), i)] for i in range(bulk.cameras.shape[0])]
            self.pt_blocks = [problem._row_blocks[(id(bulk.points), i)] for i in range(bulk.points.shape[0])]
            self.obs_camera = bulk.camera_index
            self.obs_point = bulk.point_index
            self.observations = bulk.observations
            self.loss = native_loss(bulk.loss)
            self.cam_base, self.pt_base = bulk.cameras, bulk.points
        else:
            rbs = problem._residual_blocks
            if not rbs or any(rb.cost_function.device_kind != "ba" for rb in rbs):
                raise NotImplementedError("not a bundle-adjustment problem")
            models = {getattr(rb.cost_function, "camera_model", "angle_axis") for rb in rbs}
            if len(models) != 1:
                raise NotImplementedError("the device path needs one camera model for all residual blocks")
            self.camera_model = models.pop()
            losses = {native_loss(rb.loss_function) for rb in rbs}
            if len(losses) != 1:
                raise NotImplementedError("the device path needs one loss function for all residual blocks")
            self.loss = losses.pop()
            cam_index, pt_index = {}, {}
            self.cam_blocks, self.pt_blocks = [], []
            oc, op, ob = [], [], []
            quaternion = self.camera_model != "angle_axis"
            rotation_of, rest_of = {}, {}
            for rb in rbs:
                if quaternion:
                    qb, cb, pb = rb.parameter_blocks
                    if rotation_of.setdefault(id(cb), qb) is not qb or rest_of.setdefault(id(qb), cb) is not cb:
                        raise NotImplementedError("each rotation block must pair with one camera block on the "
                                                  "MI355X device path")
                    key = id(qb)
                    if key not in cam_index:
                        cam_index[key] = len(self.cam_blocks)
                        self.cam_blocks.append(_QuaternionCamera(qb, cb, rb.cost_function.order))
                else:
                    cb, pb = rb.parameter_blocks
                    key = cb
                    if cb not in cam_index:
                        cam_index[cb] = len(self.cam_blocks)
                        self.cam_blocks.append(cb)
                if pb not in pt_index:
                    pt_index[pb] = len(self.pt_blocks)
                    self.pt_blocks.append(pb)
                oc.append(cam_index[key])
                op.append(pt_index[pb])
                ob.append((rb.cost_function.observed_x, rb.cost_function.observed_y))
            cam_ids = {id(b) for c in self.cam_blocks for b in (c.blocks if quaternion else (c,))}
            if cam_ids & {id(b) for b in pt_index}:
                raise NotImplementedError("a parameter block is used both as camera and point")
            self.obs_camera = np.asarray(oc)
            self.obs_point = np.asarray(op)
            self.observations = np.asarray(ob, dtype=np.float64)
            self.cam_base = self.pt_base = None
        for b in self.parameter_blocks():
            if b.lower_bound is not None or b.upper_bound is not None:
                raise NotImplementedError("parameter bounds are not on the MI355X device path")
            if b.tensor.dtype != torch.float64:
                raise NotImplementedError("bundle-adjustment parameters must be float64 tensors")

    def parameter_blocks(self) -> list:
        """Every parameter block of the cameras, then the points."""
        cams = [b for c in self.cam_blocks for b in (c.blocks if isinstance(c, _QuaternionCamera) else (c,))]
        return cams + list(self.pt_blocks)

    def stacked(self):
        if self.cam_base is not None:
            return self.cam_base.detach(), self.pt_base.detach()
        stride = 9 if self.camera_model == "angle_axis" else 10
        if not self.cam_blocks:
            cams = torch.zeros(0, stride, dtype=torch.float64)
        elif stride == 10:
            cams = torch.stack([c.values() for c in self.cam_blocks])
        else:
            cams = torch.stack([b.tensor.detach().reshape(-1) for b in self.cam_blocks])
        pts = torch.stack([b.tensor.detach().reshape(-1) for b in self.pt_blocks]) if self.pt_blocks else torch.zeros(0, 3, dtype=torch.float64)
        return cams, pts

    def masks(self):
        cm = np.array([b.free_mask() for b in self.cam_blocks], dtype=np.uint8).reshape(-1, 9)
        pm = np.array([b.free_mask() for b in self.pt_blocks], dtype=np.uint8).reshape(-1, 3)
        return (None if cm.all() else cm), (None if pm.all() else pm)

    def write_back(self, cams: torch.Tensor, pts: torch.Tensor) -> None:
        with torch.no_grad():
            if self.cam_base is not None:
                self.cam_base.copy_(cams.to(self.cam_base.device, self.cam_base.dtype))
                self.pt_base.copy_(pts.to(self.pt_base.device, self.pt_base.dtype))
                return
            cams_h = cams.to("cpu")
            pts_h = pts.to("cpu")
            for i, b in enumerate(self.cam_blocks):
                if isinstance(b, _QuaternionCamera):
                    b.write(cams_h[i])
                else:
                    b.tensor.reshape(-1).copy_(cams_h[i].to(b.tensor.device))
            for i, b in enumerate(self.pt_blocks):
                b.tensor.reshape(-1).copy_(pts_h[i].to(b.tensor.device))


def _device_for(problem: Problem) -> BADevice:
    cached = getattr(problem, "_device_cache", None)
    ex = _BAExtract(problem)
    cams, pts = ex.stacked()
    cm, pm = ex.masks()
    # constant blocks and manifolds may change between solves: the masks are part of the key
    key = (len(ex.cam_blocks), len(ex.pt_blocks), len(ex.obs_camera), ex.loss, ex.camera_model,
           None if cm is None else cm.tobytes(), None if pm is None else pm.tobytes())
    if cached is not None and cached[0] == key:
        dev = cached[1]
        dev.set_parameters(cams, pts)
        return dev, ex
    dev = BADevice(cams, pts, ex.obs_camera, ex.obs_point, ex.observations, loss_native=ex.loss,
                   camera_mask=cm, point_mask=pm, camera_model=ex.camera_model)
    problem._device_cache = (key, dev)
    return dev, ex


def _records_to_iterations(records: np.ndarray, eta: float) -> list[IterationSummary]:
    its = []
    prev_t = 0.0
    for row in records:
        t = float(row[nat.REC["timestamp"]])
        its.append(IterationSummary(
            iteration=int(row[0]),
            step_is_valid=bool(row[nat.REC["step_is_valid"]]),
            step_is_successful=bool(row[nat.REC["step_is_successful"]]),
            cost=float(row[nat.REC["cost"]]),
            cost_change=float(row[nat.REC["cost_change"]]),
            gradient_max_norm=float(row[nat.REC["gradient_max_norm"]]),
            gradient_norm=float(row[nat.REC["gradient_norm"]]),
            step_norm=float(row[nat.REC["step_norm"]]),
            relative_decrease=float(row[nat.REC["relative_decrease"]]),
            trust_region_radius=float(row[nat.REC["trust_region_radius"]]),
            eta=eta,
            linear_solver_iterations=int(row[nat.REC["linear_solver_iterations"]]),
            iteration_time_in_seconds=max(0.0, t - prev_t),
            cumulative_time_in_seconds=t,
        ))
        prev_t = t
    return its


def solve(options: SolverOptions, problem: Problem) -> SolverSummary:
    """Run the trust-region solve of ``problem`` on the MI355X."""
    options.validate()
    _check_supported(options)
    start = time.perf_counter()
    kinds = {rb.cost_function.device_kind for rb in problem._residual_blocks}
    if problem.bulk_ba:
        kinds.add("ba")
    if options.trust_region_strategy_type is TrustRegionStrategyType.DOGLEG and kinds != {"ba"}:
        raise NotImplementedError("DOGLEG is on the MI355X device path for bundle adjustment only")
    if options.minimizer_type is MinimizerType.LINE_SEARCH and kinds != {"ba"}:
        raise NotImplementedError("the LINE_SEARCH minimizer is on the MI355X device path for bundle adjustment only")
    if not torch.cuda.is_available():
        raise nat.NativeLibraryError("the MI355X solver needs a ROCm GPU (torch.cuda.is_available() is False)")
    if kinds == {"ba"} and options.minimizer_type is MinimizerType.LINE_SEARCH:
        return _solve_ba_line_search(options, problem, start)
    if kinds == {"ba"}:
        return _solve_ba(options, problem, start)
    if kinds == {"curve"}:
        from .curve import solve_curve_problem

        return solve_curve_problem(options, problem, start)
    if kinds == {"pose2d"}:
        from .pose2d import solve_pose_problem

        return solve_pose_problem(options, problem, start)
    raise NotImplementedError(
        "the MI355X device path solves problems made of SnavelyReprojectionError, ExpCurveResidual or "
        f"Pose2DRelativeError residual blocks; got {sorted(str(k) for k in kinds)}"
    )


def summary_from_report(options: SolverOptions, problem: Problem, rep, records, start: float,
                        precision: str) -> SolverSummary:
    s = _summary_base(options, problem)
    s.termination_type = _TERMINATION[int(rep.termination)]
    s.message = _MESSAGES.get(int(rep.reason), "Solver failed.")
    s.initial_cost = float(rep.initial_cost)
    s.final_cost = float(rep.final_cost)
    s.num_successful_steps = int(rep.num_successful_steps)
    s.num_unsuccessful_steps = int(rep.num_unsuccessful_steps)
    s.num_residual_evaluations = int(rep.num_residual_evaluations)
    s.num_jacobian_evaluations = int(rep.num_jacobian_evaluations)
    s.num_linear_solves = int(rep.num_linear_solves)
    s.iterations = _records_to_iterations(records, options.eta)
    s.device_time_in_seconds = float(rep.device_seconds)
    s.minimizer_time_in_seconds = float(rep.device_seconds)
    s.precision = precision
    s.total_time_in_seconds = time.perf_counter() - start
    return s


def _summary_base(options: SolverOptions, problem: Problem) -> SolverSummary:
    """The option and problem fields of a summary (solver.py _new_summary)."""
    return SolverSummary(
        minimizer_type=options.minimizer_type,
        num_parameter_blocks=problem.num_parameter_blocks(),
        num_parameters=problem.num_parameters(),
        num_effective_parameters=problem.num_effective_parameters(),
        num_residual_blocks=problem.num_residual_blocks(),
        num_residuals=problem.num_residuals(),
        linear_solver_type_given=options.linear_solver_type,
        linear_solver_type_used=_effective_solver(options.linear_solver_type),
        trust_region_strategy_type=options.trust_region_strategy_type,
        dogleg_type=options.dogleg_type,
        line_search_direction_type=options.line_search_direction_type,
        line_search_type=options.line_search_type,
        line_search_interpolation_type=options.line_search_interpolation_type,
        nonlinear_conjugate_gradient_type=options.nonlinear_conjugate_gradient_type,
        preconditioner_type=options.preconditioner_type,
        max_lbfgs_rank=options.max_lbfgs_rank,
        dense_linear_algebra_library_type=options.dense_linear_algebra_library_type,
        sparse_linear_algebra_library_type=options.sparse_linear_algebra_library_type,
    )


def _eliminated_blocks(problem: Problem) -> set:
    """Blocks the reference's Schur solvers eliminate: the first ordering group of the active blocks
    (_linear_solver_parameter_order and _num_eliminate_for_schur, solver.py:714-741), none when no
    group is set or all active blocks share one."""
    blocks = [b for b in problem.parameter_blocks if b.tangent_size > 0]
    if not blocks or not any(b.ordering_group is not None for b in blocks):
        return set()
    max_group = max((b.ordering_group for b in blocks if b.ordering_group is not None), default=0)
    index = {id(b): i for i, b in enumerate(blocks)}
    order = sorted(blocks, key=lambda b: (b.ordering_group if b.ordering_group is not None else max_group + 1,
                                          index[id(b)]))
    first = order[0].ordering_group
    if first is None or all(b.ordering_group == first for b in order):
        return set()
    return {id(b) for b in order if b.ordering_group == first}


def _check_elimination(problem: Problem, ex: "_BAExtract", options: SolverOptions) -> None:
    """ITERATIVE_SCHUR runs PCG on the Schur complement of the points, the device's one form; the
    reference eliminates its first ordering group, and with none runs CG on the normal equations
    (linear.py:191-257), a different iteration.  The exact solvers give the same step in any order."""
    if (options.linear_solver_type is not LinearSolverType.ITERATIVE_SCHUR
            or options.trust_region_strategy_type is TrustRegionStrategyType.DOGLEG):
        return
    points = {id(b) for b in ex.pt_blocks if b.tangent_size > 0}
    if not points or _eliminated_blocks(problem) != points:
        raise NotImplementedError("ITERATIVE_SCHUR on the MI355X device path eliminates exactly the point blocks: "
                                  "put every variable point (and nothing else) in the first ordering group")


def _solve_ba(options: SolverOptions, problem: Problem, start: float) -> SolverSummary:
    dev, ex = _device_for(problem)
    _check_elimination(problem, ex, options)
    precision = "mixed" if options.use_mixed_precision_solves else "fp64"
    opts = lm_options(options, precision=precision, loop=DEVICE_SETTINGS["loop"], linear=_linear_kind(options),
                      iterations_per_graph=DEVICE_SETTINGS["iterations_per_graph"])
    pre = time.perf_counter()
    rep, records = dev.run(opts)
    torch.cuda.synchronize()
    s = summary_from_report(options, problem, rep, records, start, precision)
    s.preprocessor_time_in_seconds = pre - start
    post = time.perf_counter()
    ex.write_back(dev.cameras, dev.points)
    s.postprocessor_time_in_seconds = time.perf_counter() - post
    s.total_time_in_seconds = time.perf_counter() - start
    return s


def _solve_ba_line_search(options: SolverOptions, problem: Problem, start: float) -> SolverSummary:
    """The LINE_SEARCH minimizer (solver.py:360-607) on the device: ct_ba_line_search_solve."""
    dev, ex = _device_for(problem)
    pre = time.perf_counter()
    rep, records = dev.line_search(line_search_options(options))
    torch.cuda.synchronize()
    s = _summary_base(options, problem)
    s.termination_type = _TERMINATION[int(rep.termination)]
    s.message = _MESSAGES.get(int(rep.reason), "Solver failed.")
    s.initial_cost = float(rep.initial_cost)
    s.final_cost = float(rep.final_cost)
    s.num_successful_steps = int(rep.num_successful_steps)
    s.num_unsuccessful_steps = int(rep.num_unsuccessful_steps)
    s.num_jacobian_evaluations = int(rep.num_jacobian_evaluations)
    s.num_line_search_steps = int(rep.num_line_search_steps)
    s.num_line_search_function_evaluations = int(rep.num_line_search_function_evaluations)
    s.num_line_search_gradient_evaluations = int(rep.num_line_search_gradient_evaluations)
    s.num_line_search_direction_restarts = int(rep.num_line_search_direction_restarts)
    s.iterations = _line_search_records_to_iterations(records)
    s.minimizer_time_in_seconds = float(rep.seconds)
    s.device_time_in_seconds = float(rep.seconds)
    s.precision = "fp64"
    s.preprocessor_time_in_seconds = pre - start
    post = time.perf_counter()
    ex.write_back(dev.cameras, dev.points)
    s.postprocessor_time_in_seconds = time.perf_counter() - post
    s.total_time_in_seconds = time.perf_counter() - start
    return s


def _line_search_records_to_iterations(records: np.ndarray) -> list[IterationSummary]:
    its = []
    prev_t = 0.0
    R = nat.LS_REC
    for row in records:
        t = float(row[R["timestamp"]])
        its.append(IterationSummary(
            iteration=int(row[R["iteration"]]),
            cost=float(row[R["cost"]]),
            gradient_norm=float(row[R["gradient_norm"]]),
            gradient_max_norm=float(row[R["gradient_max_norm"]]),
            step_size=float(row[R["step_size"]]),
            line_search_iterations=int(row[R["line_search_iterations"]]),
            line_search_function_evaluations=int(row[R["line_search_function_evaluations"]]),
            line_search_gradient_evaluations=int(row[R["line_search_gradient_evaluations"]]),
            line_search_direction_restarts=int(row[R["line_search_direction_restarts"]]),
            step_norm=float(row[R["step_norm"]]),
            cost_change=float(row[R["cost_change"]]),
            step_is_successful=bool(row[R["step_is_successful"]]),
            iteration_time_in_seconds=max(0.0, t - prev_t),
            cumulative_time_in_seconds=t,
        ))
        prev_t = t
    return its


# a dense Jacobian above this size is refused (the reference would build it on the host)
MAX_DENSE_JACOBIAN_BYTES = 8 << 30


def _column_maps(problem: Problem, ex: "_BAExtract", blocks):
    """Tangent column of every camera / point coordinate (-1: none), as _active_column_map
    (problem.py:522-535) orders them: the selected blocks in order, constant blocks skipped,
    a block's free coordinates in order."""
    from .problem import ParameterBlock

    selected = [problem._coerce(b) if not isinstance(b, ParameterBlock) else b for b in blocks]
    column = {}
    offset = 0
    for b in selected:
        if b.tangent_size == 0:
            continue
        # a quaternion rotation: its three tangent coordinates (no per-value mask)
        mask = [True] * 3 if getattr(b.manifold, "device_camera_model", None) else b.free_mask()
        cols = np.full(len(mask), -1, dtype=np.int64)
        free = np.nonzero(mask)[0]
        cols[free] = offset + np.arange(free.size)
        column[id(b)] = cols
        offset += b.tangent_size

    def cam_cols(c):
        if isinstance(c, _QuaternionCamera):
            return np.concatenate([column.get(id(c.rotation), np.full(3, -1, np.int64)),
                                   column.get(id(c.rest), np.full(6, -1, np.int64))])
        return column.get(id(c), np.full(9, -1, np.int64))

    cam = np.stack([cam_cols(c) for c in ex.cam_blocks]) if ex.cam_blocks else np.zeros((0, 9), np.int64)
    pt = np.stack([column.get(id(b), np.full(3, -1, np.int64)) for b in ex.pt_blocks]) if ex.pt_blocks else \
        np.zeros((0, 3), np.int64)
    return cam, pt, offset


def evaluate_problem(problem: Problem, options=None, compute_jacobian: bool = False) -> EvaluationResult:
    from .problem import CRSMatrix, EvaluateOptions

    options = options or EvaluateOptions()
    kinds = {rb.cost_function.device_kind for rb in problem._residual_blocks}
    if problem.bulk_ba:
        kinds.add("ba")
    if kinds != {"ba"}:
        raise NotImplementedError("device evaluation is implemented for bundle-adjustment problems")
    if options.residual_blocks is not None:
        raise NotImplementedError("device evaluation covers all residual blocks (EvaluateOptions.residual_blocks)")
    dev, ex = _device_for(problem)
    if not compute_jacobian and options.apply_loss_function:
        cost, res = dev.evaluate()
        return EvaluationResult(cost=torch.tensor(cost, dtype=torch.float64), residuals=res.reshape(-1))
    blocks = options.parameter_blocks if options.parameter_blocks is not None else problem.parameter_blocks
    cam_cols, pt_cols, ncols = _column_maps(problem, ex, blocks)
    if not compute_jacobian:
        cam_cols[:] = -1
        pt_cols[:] = -1
        ncols = 0
    need = 2 * dev.num_observations * ncols * 8
    if need > MAX_DENSE_JACOBIAN_BYTES:
        raise ValueError(f"the dense Jacobian ({2 * dev.num_observations} x {ncols}) would need {need / 2**30:.1f} GiB; "
                         "call evaluate(compute_jacobian=False)")
    cost, res, jac = dev.evaluate_jacobian(cam_cols, pt_cols, ncols, apply_loss=options.apply_loss_function)
    out = EvaluationResult(cost=torch.tensor(cost, dtype=torch.float64), residuals=res)
    if compute_jacobian:
        out.jacobian = jac
        out.gradient = jac.T @ res
        out.crs_jacobian = CRSMatrix.from_dense(jac)
    return out
