"""Bundle-adjustment device path: problem tables, plans and LM runs.

The reference evaluates every residual block with autograd, densifies the
Jacobian and solves the damped normal equations with a dense Schur
complement (``problem.py:377-413``, ``linear.py:667-699``).  Here the whole LM
solve (``solver.py:35-357``) runs in the HIP library: observations are grouped
by point, points are eliminated per point (3x3 blocks) and the reduced camera
system S (9C x 9C) is assembled per camera pair from a precomputed table of
observation pairs, then solved in FP64.
"""
from __future__ import annotations

import contextlib
import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _native as nat
from .losses import LOSS_IDS


@dataclass
class BATables:
    """Host-side structure tables (numpy int32), built once per problem."""

    num_cameras: int
    num_points: int
    order: np.ndarray            # device observation k = caller observation order[k]
    obs_camera: np.ndarray       # [N] (point-grouped)
    obs_point: np.ndarray        # [N] (non-decreasing)
    point_offsets: np.ndarray    # [P+1]
    camera_offsets: np.ndarray   # [C+1]
    camera_obs: np.ndarray       # [N]
    block_cameras: np.ndarray    # [B, 2] all pairs i <= j, row-major
    block_offsets: np.ndarray    # [B+1]
    block_terms: np.ndarray      # [T, 2]


def build_tables(obs_camera, obs_point, num_cameras: int, num_points: int) -> BATables:
    obs_camera = np.asarray(obs_camera, dtype=np.int64)
    obs_point = np.asarray(obs_point, dtype=np.int64)
    n = obs_camera.shape[0]
    order = np.lexsort((obs_camera, obs_point)) if n else np.zeros(0, np.int64)  # by point, then camera
    cam = obs_camera[order]
    pt = obs_point[order]
    counts = np.bincount(pt, minlength=num_points) if n else np.zeros(num_points, np.int64)
    point_offsets = np.zeros(num_points + 1, np.int64)
    point_offsets[1:] = np.cumsum(counts)
    corder = np.argsort(cam, kind="stable")
    ccounts = np.bincount(cam, minlength=num_cameras) if n else np.zeros(num_cameras, np.int64)
    camera_offsets = np.zeros(num_cameras + 1, np.int64)
    camera_offsets[1:] = np.cumsum(ccounts)
    # ordered observation pairs of a common point, kept when cam(a) <= cam(b)
    pa_list, pb_list = [], []
    maxk = int(counts.max()) if counts.size else 0
    for k in range(1, maxk + 1):
        pts = np.nonzero(counts == k)[0]
        if pts.size == 0:
            continue
        base = point_offsets[pts]
        ii, jj = np.meshgrid(np.arange(k), np.arange(k), indexing="ij")
        a = (base[:, None] + ii.reshape(1, -1)).reshape(-1)
        b = (base[:, None] + jj.reshape(1, -1)).reshape(-1)
        keep = cam[a] <= cam[b]
        pa_list.append(a[keep])
        pb_list.append(b[keep])
    pa = np.concatenate(pa_list) if pa_list else np.zeros(0, np.int64)
    pb = np.concatenate(pb_list) if pb_list else np.zeros(0, np.int64)
    key = cam[pa] * num_cameras + cam[pb]
    torder = np.argsort(key, kind="stable")
    pa, pb, key = pa[torder], pb[torder], key[torder]
    iu, ju = np.triu_indices(num_cameras)
    block_keys = iu * num_cameras + ju
    block_offsets = np.zeros(block_keys.shape[0] + 1, np.int64)
    block_offsets[1:] = np.searchsorted(key, block_keys, side="right")
    block_offsets[0] = 0
    starts = np.searchsorted(key, block_keys, side="left")
    assert np.array_equal(starts, block_offsets[:-1])
    return BATables(
        num_cameras=num_cameras,
        num_points=num_points,
        order=order,
        obs_camera=cam.astype(np.int32),
        obs_point=pt.astype(np.int32),
        point_offsets=point_offsets.astype(np.int32),
        camera_offsets=camera_offsets.astype(np.int32),
        camera_obs=corder.astype(np.int32),
        block_cameras=np.stack([iu, ju], axis=1).astype(np.int32),
        block_offsets=block_offsets.astype(np.int32),
        block_terms=np.stack([pa, pb], axis=1).astype(np.int32),
    )


def native_loss_struct(native: tuple) -> nat.ct_loss:
    kind, a, b, scale = native
    return nat.ct_loss(LOSS_IDS[kind], float(a), float(b), float(scale))


# ct_camera_model and the number of stored values per camera
CAMERA_MODELS = {"angle_axis": (0, 9), "quaternion": (1, 10), "eigen_quaternion": (2, 10)}


class BADevice:
    """Device-resident BA problem: tables, observations, masks and parameters.

    ``cameras`` [C, 9] (angle-axis) or [C, 10] (``camera_model`` "quaternion":
    q in w, x, y, z order, or "eigen_quaternion": x, y, z, w; then t, f, k1,
    k2); ``camera_mask`` [C, 9] is over the tangent coordinates.
    """

    def __init__(self, cameras: torch.Tensor, points: torch.Tensor, obs_camera, obs_point, observations, *,
                 loss_native=("trivial", 0.0, 0.0, 1.0), camera_mask=None, point_mask=None, device="cuda",
                 camera_model: str = "angle_axis"):
        nat.load()
        self.device = torch.device(device)
        if camera_model not in CAMERA_MODELS:
            raise ValueError(f"camera_model must be one of {sorted(CAMERA_MODELS)}")
        self.camera_model = camera_model
        stride = CAMERA_MODELS[camera_model][1]
        C, P = int(cameras.shape[0]), int(points.shape[0])
        if cameras.numel() != stride * C:
            raise ValueError(f"a {camera_model} camera has {stride} values")
        self.tables = build_tables(obs_camera, obs_point, C, P)
        t = self.tables
        obs = np.asarray(observations, dtype=np.float64).reshape(-1, 2)[t.order]
        dev = self.device

        def i32(a):
            return torch.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).to(dev)

        self.obs_camera = i32(t.obs_camera)
        self.obs_point = i32(t.obs_point)
        self.point_offsets = i32(t.point_offsets)
        self.observations = torch.from_numpy(np.ascontiguousarray(obs)).to(dev)
        self.camera_offsets = i32(t.camera_offsets)
        self.camera_obs = i32(t.camera_obs)
        self.block_cameras = i32(t.block_cameras)
        self.block_offsets = i32(t.block_offsets)
        self.block_terms = i32(t.block_terms)
        self.camera_mask = None if camera_mask is None else torch.as_tensor(
            np.asarray(camera_mask, np.uint8).reshape(-1)).to(dev)
        self.point_mask = None if point_mask is None else torch.as_tensor(
            np.asarray(point_mask, np.uint8).reshape(-1)).to(dev)
        self.cameras = cameras.detach().to(device=dev, dtype=torch.float64).contiguous().clone()
        self.points = points.detach().to(device=dev, dtype=torch.float64).contiguous().clone()
        self.loss_native = loss_native
        self.num_cameras, self.num_points = C, P
        self.num_observations = int(obs.shape[0])
        self._plans: dict = {}

    def problem_struct(self) -> nat.ct_ba_problem:
        p = nat.ct_ba_problem()
        p.num_cameras = self.num_cameras
        p.num_points = self.num_points
        p.num_observations = self.num_observations
        p.num_blocks = int(self.block_cameras.shape[0])
        p.num_terms = int(self.block_terms.shape[0])
        p.obs_camera = nat.ptr(self.obs_camera)
        p.obs_point = nat.ptr(self.obs_point)
        p.point_offsets = nat.ptr(self.point_offsets)
        p.observations = nat.ptr(self.observations)
        p.camera_offsets = nat.ptr(self.camera_offsets)
        p.camera_obs = nat.ptr(self.camera_obs)
        p.block_cameras = nat.ptr(self.block_cameras)
        p.block_offsets = nat.ptr(self.block_offsets)
        p.block_terms = nat.ptr(self.block_terms)
        p.camera_mask = nat.ptr(self.camera_mask)
        p.point_mask = nat.ptr(self.point_mask)
        p.loss = native_loss_struct(self.loss_native)
        p.camera_model = CAMERA_MODELS[self.camera_model][0]
        return p

    def set_parameters(self, cameras: torch.Tensor, points: torch.Tensor) -> None:
        self.cameras.copy_(cameras.detach().to(self.cameras.device, torch.float64))
        self.points.copy_(points.detach().to(self.points.device, torch.float64))

    def run(self, opts: nat.ct_lm_options):
        """Run one LM solve in place on self.cameras/self.points.

        Returns (report, records[num_records, CT_REC_FIELDS] numpy).
        """
        lib = nat.load()
        handle = self.plan(opts)
        rep = nat.ct_lm_report()
        nrec = (opts.max_num_iterations + 2) * nat.CT_REC_FIELDS
        rec = (ctypes.c_double * nrec)()
        nat.check(lib.ct_ba_plan_run(handle, ctypes.byref(rep), rec), "ct_ba_plan_run")
        records = np.frombuffer(rec, dtype=np.float64).reshape(-1, nat.CT_REC_FIELDS)[: rep.num_records].copy()
        return rep, records

    def plan(self, opts: nat.ct_lm_options):
        """The plan handle for these options (created on first use, kept until close())."""
        lib = nat.load()
        key = tuple(getattr(opts, f) for f, _ in nat.ct_lm_options._fields_)
        plan = self._plans.get(key)
        prob = self.problem_struct()
        if plan is None:
            nbytes = ctypes.c_size_t(0)
            nat.check(lib.ct_ba_workspace_bytes(ctypes.byref(prob), ctypes.byref(opts), ctypes.byref(nbytes)),
                      "ct_ba_workspace_bytes")
            ws = torch.zeros(int(nbytes.value), dtype=torch.uint8, device=self.device)
            handle = ctypes.c_void_p()
            nat.check(lib.ct_ba_plan_create(ctypes.byref(prob), ctypes.byref(opts), nat.ptr(self.cameras),
                                            nat.ptr(self.points), nat.ptr(ws), ws.numel(),
                                            nat.current_stream_handle(), ctypes.byref(handle)),
                      "ct_ba_plan_create")
            plan = (handle, ws, prob)
            self._plans[key] = plan
        return plan[0]

    def evaluate(self, cameras=None, points=None):
        """(cost, residuals[N, 2] in caller observation order) at the given parameters."""
        lib = nat.load()
        cams = self.cameras if cameras is None else cameras.detach().to(self.device, torch.float64).contiguous()
        pts = self.points if points is None else points.detach().to(self.device, torch.float64).contiguous()
        prob = self.problem_struct()
        res = torch.empty((self.num_observations, 2), dtype=torch.float64, device=self.device)
        part = torch.zeros(self.num_points // 256 + 2, dtype=torch.float64, device=self.device)
        cost = ctypes.c_double(0.0)
        nat.check(lib.ct_ba_evaluate(ctypes.byref(prob), nat.ptr(cams), nat.ptr(pts), ctypes.byref(cost),
                                     nat.ptr(res), nat.ptr(part), part.numel() * 8, nat.current_stream_handle()),
                  "ct_ba_evaluate")
        out = torch.empty_like(res)
        out[torch.from_numpy(self.tables.order).to(self.device)] = res
        return cost.value, out

    def evaluate_jacobian(self, camera_columns: np.ndarray, point_columns: np.ndarray, num_columns: int,
                          apply_loss: bool = True):
        """(cost, residuals [2N], dense Jacobian [2N, num_columns]) in caller observation order.

        camera_columns [C, 9] / point_columns [P, 3]: each coordinate's tangent
        column, -1 for none.
        """
        lib = nat.load()
        prob = self.problem_struct()
        N = self.num_observations
        jac = torch.zeros((2 * N, num_columns), dtype=torch.float64, device=self.device)
        res = torch.empty(2 * N, dtype=torch.float64, device=self.device)
        rows = torch.from_numpy(np.asarray(self.tables.order, dtype=np.int32)).to(self.device)
        cc = torch.from_numpy(np.asarray(camera_columns, dtype=np.int64).reshape(-1)).to(self.device)
        pc = torch.from_numpy(np.asarray(point_columns, dtype=np.int64).reshape(-1)).to(self.device)
        part = torch.zeros(N // 256 + 2, dtype=torch.float64, device=self.device)
        cost = ctypes.c_double(0.0)
        nat.check(lib.ct_ba_evaluate_jacobian(ctypes.byref(prob), nat.ptr(self.cameras), nat.ptr(self.points),
                                              nat.ptr(rows), nat.ptr(cc), nat.ptr(pc), int(num_columns),
                                              1 if apply_loss else 0, nat.ptr(jac), nat.ptr(res),
                                              ctypes.byref(cost), nat.ptr(part), part.numel() * 8,
                                              nat.current_stream_handle()),
                  "ct_ba_evaluate_jacobian")
        return cost.value, res, jac

    def line_search(self, opts: nat.ct_line_search_options):
        """The LINE_SEARCH minimizer in place on self.cameras/self.points (ct_ba_line_search_solve).

        Returns (report, records[num_records, CT_LS_REC_FIELDS] numpy).
        """
        lib = nat.load()
        prob = self.problem_struct()
        nbytes = ctypes.c_size_t(0)
        nat.check(lib.ct_ba_line_search_workspace_bytes(ctypes.byref(prob), ctypes.byref(opts), ctypes.byref(nbytes)),
                  "ct_ba_line_search_workspace_bytes")
        ws = torch.empty(int(nbytes.value), dtype=torch.uint8, device=self.device)
        rep = nat.ct_line_search_report()
        rec = np.zeros((opts.max_num_iterations + 1, nat.CT_LS_REC_FIELDS), dtype=np.float64)
        nat.check(lib.ct_ba_line_search_solve(ctypes.byref(prob), ctypes.byref(opts), nat.ptr(self.cameras),
                                              nat.ptr(self.points), nat.ptr(ws), ws.numel(),
                                              nat.current_stream_handle(), ctypes.byref(rep),
                                              rec.ctypes.data_as(ctypes.c_void_p)),
                  "ct_ba_line_search_solve")
        return rep, rec[: int(rep.num_records)].copy()

    def close(self) -> None:
        lib = nat.load()
        for handle, _ws, _p in self._plans.values():
            lib.ct_ba_plan_destroy(handle)
        self._plans.clear()

    def __del__(self):  # pragma: no cover - interpreter shutdown order varies
        try:
            self.close()
        except Exception:
            pass


_debug_invalid_steps = 0
_debug_fill_lds = 0

# How ITERATIVE_SCHUR forms S p: "auto" (from the evaluation's records when the
# plan allows it) or "explicit" (assembled S); see set_schur_operator.
SCHUR_OPERATOR = {"mode": "auto"}


def set_schur_operator(mode: str) -> None:
    """'auto' (S p from the per-observation records, no off-diagonal blocks of S) or 'explicit'."""
    if mode not in ("auto", "explicit"):
        raise ValueError("mode must be 'auto' or 'explicit'")
    SCHUR_OPERATOR["mode"] = mode


@contextlib.contextmanager
def debug_invalid_steps(count: int):
    """Test hook: inside the block, the first ``count`` trust-region steps of
    every bundle-adjustment solve are overwritten with NaN on the device.

    This is the device form of the reference's own test, which patches the
    linear solver to return NaN (``tests/test_problem_solver.py:637-664``): the
    invalid-step branch of ``solver.py:168-190`` then runs.  Not for normal use.
    """
    global _debug_invalid_steps
    if count < 0:
        raise ValueError("count must be >= 0")
    previous, _debug_invalid_steps = _debug_invalid_steps, int(count)
    try:
        yield
    finally:
        _debug_invalid_steps = previous


LS_DIRECTIONS = {"STEEPEST_DESCENT": 0, "NONLINEAR_CONJUGATE_GRADIENT": 1, "LBFGS": 2, "BFGS": 3}
LS_TYPES = {"ARMIJO": 0, "WOLFE": 1}
NCG_TYPES = {"FLETCHER_REEVES": 0, "POLAK_RIBIERE": 1, "HESTENES_STIEFEL": 2}
LS_INTERPOLATIONS = {"BISECTION": 0, "QUADRATIC": 1, "CUBIC": 2}


def line_search_options(options) -> nat.ct_line_search_options:
    """SolverOptions -> ct_line_search_options (the fields solver.py:360-607 and line_search.py read)."""
    o = nat.ct_line_search_options()
    o.max_num_iterations = int(options.max_num_iterations)
    o.direction = LS_DIRECTIONS[options.line_search_direction_type.name]
    o.line_search_type = LS_TYPES[options.line_search_type.name]
    o.nonlinear_cg_type = NCG_TYPES[options.nonlinear_conjugate_gradient_type.name]
    o.max_lbfgs_rank = int(options.max_lbfgs_rank)
    o.use_approximate_eigenvalue_bfgs_scaling = int(bool(options.use_approximate_eigenvalue_bfgs_scaling))
    o.interpolation = LS_INTERPOLATIONS[options.line_search_interpolation_type.name]
    o.max_num_step_size_iterations = int(options.max_num_line_search_step_size_iterations)
    o.max_num_direction_restarts = int(options.max_num_line_search_direction_restarts)
    o.min_line_search_step_size = float(options.min_line_search_step_size)
    o.sufficient_function_decrease = float(options.line_search_sufficient_function_decrease)
    o.max_step_contraction = float(options.max_line_search_step_contraction)
    o.min_step_contraction = float(options.min_line_search_step_contraction)
    o.sufficient_curvature_decrease = float(options.line_search_sufficient_curvature_decrease)
    o.function_tolerance = float(options.function_tolerance)
    o.gradient_tolerance = float(options.gradient_tolerance)
    o.parameter_tolerance = float(options.parameter_tolerance)
    o.max_solver_time_in_seconds = float(getattr(options, "max_solver_time_in_seconds", 1e9))
    return o


@contextlib.contextmanager
def debug_fill_lds():
    """Test hook: inside the block, every bundle-adjustment solve fills each compute unit's LDS
    with NaN right before each reduced-camera solve, so a solver kernel that reads LDS it has not
    written in its own launch fails its parity test.  Not for normal use."""
    global _debug_fill_lds
    previous, _debug_fill_lds = _debug_fill_lds, 1
    try:
        yield
    finally:
        _debug_fill_lds = previous


def lm_options(options, *, precision: str = "fp64", loop: str = "host", linear: str = "dense_schur",
               iterations_per_graph: int = 8, solver_workgroups: int = 0) -> nat.ct_lm_options:
    """Translate SolverOptions into the native option struct."""
    o = nat.ct_lm_options()
    o.max_num_iterations = int(options.max_num_iterations)
    o.max_num_consecutive_invalid_steps = int(options.max_num_consecutive_invalid_steps)
    o.linear_solver = 1 if linear == "iterative_schur" else 0
    o.precision = 1 if precision == "mixed" else 0
    o.loop_mode = 1 if loop == "device" else 0
    o.min_linear_solver_iterations = int(options.min_linear_solver_iterations)
    o.max_linear_solver_iterations = int(options.max_linear_solver_iterations)
    o.max_num_refinement_iterations = int(options.max_num_refinement_iterations)
    o.jacobi_scaling = 1 if options.jacobi_scaling else 0
    o.iterations_per_graph = int(iterations_per_graph)
    o.initial_trust_region_radius = float(options.initial_trust_region_radius)
    o.max_trust_region_radius = float(options.max_trust_region_radius)
    o.min_trust_region_radius = float(options.min_trust_region_radius)
    o.min_relative_decrease = float(options.min_relative_decrease)
    o.min_lm_diagonal = float(options.min_lm_diagonal)
    o.max_lm_diagonal = float(options.max_lm_diagonal)
    o.function_tolerance = float(options.function_tolerance)
    o.gradient_tolerance = float(options.gradient_tolerance)
    o.parameter_tolerance = float(options.parameter_tolerance)
    o.eta = float(options.eta)
    dogleg = getattr(options, "trust_region_strategy_type", None)
    o.trust_region_strategy = 1 if dogleg is not None and dogleg.name == "DOGLEG" else 0
    dtype = getattr(options, "dogleg_type", None)
    o.dogleg_type = 1 if dtype is not None and dtype.name == "SUBSPACE_DOGLEG" else 0
    o.max_solver_time_in_seconds = float(getattr(options, "max_solver_time_in_seconds", 1e9))
    pc = getattr(options, "preconditioner_type", None)
    o.preconditioner = PRECONDITIONERS.get(pc.name, 0) if pc is not None else 0
    o.debug_invalid_steps = _debug_invalid_steps
    o.debug_fill_lds = _debug_fill_lds
    o.schur_operator = 0 if SCHUR_OPERATOR["mode"] == "auto" else 1
    o.solver_workgroups = int(solver_workgroups)
    return o


# ITERATIVE_SCHUR preconditioners on the device (ct_preconditioner)
PRECONDITIONERS = {"SCHUR_JACOBI": 0, "JACOBI": 1, "IDENTITY": 2}
