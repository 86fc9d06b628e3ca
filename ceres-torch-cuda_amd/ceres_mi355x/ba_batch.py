"""Many small bundle adjustments in one persistent launch (``ct_ba_batch_solve``).

The reference solves one ``Problem`` per ``solve()`` call (``solver.py:28``);
a served stream of small bundle adjustments would pay a plan, its allocations
and a host-driven loop per request.  Here every problem of a batch (angle-axis
Snavely cameras, all parameters free, at most 21 cameras) is laid out in
concatenated device arrays with problem-local indices, and one workgroup per
problem runs the whole LM solve (``solver.py:35-357``, DENSE_SCHUR step of
``linear.py:667-699``) on the device.  A problem's arithmetic does not depend
on the rest of the batch, so its result is the same solved alone.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _native as nat
from .ba import native_loss_struct

MAX_CAMERAS = nat.CT_BA_BATCH_MAX_CAMERAS


@dataclass
class BAProblemInput:
    """One bundle adjustment: cameras [C, 9], points [P, 3], observations [N, 2]."""

    cameras: np.ndarray
    points: np.ndarray
    obs_camera: np.ndarray
    obs_point: np.ndarray
    observations: np.ndarray
    loss_native: tuple = ("trivial", 0.0, 0.0, 1.0)


def problem_bytes(num_cameras: int, num_points: int, num_observations: int, num_terms: int) -> int:
    """Device workspace of one problem in a batch (exact, from the library)."""
    lib = nat.load()
    out = ctypes.c_size_t(0)
    nat.check(lib.ct_ba_batch_problem_bytes(int(num_cameras), int(num_points), int(num_observations),
                                            int(num_terms), ctypes.byref(out)), "ct_ba_batch_problem_bytes")
    return int(out.value)


def device_bytes(num_cameras: int, num_points: int, num_observations: int, num_terms: int) -> int:
    """Every device byte a problem adds to a batch: its inputs, tables, parameters and workspace slice."""
    C, P, N = int(num_cameras), int(num_points), int(num_observations)
    B = C * (C + 1) // 2
    tables = 4 * (3 * N + (P + 1) + (C + 1) + (B + 1) + 2 * int(num_terms))
    return problem_bytes(C, P, N, num_terms) + tables + 16 * N + 8 * (9 * C + 3 * P) + 128


def _np_ptr(a: np.ndarray) -> int:
    return a.ctypes.data if a.size else None


def build_batch_tables(cams_per, pts_per, obs_camera: list, obs_point: list) -> dict:
    """Every problem's tables (those of ``ba.build_tables``, problem-local, concatenated in problem order) from
    the library's native builder (``ct_ba_batch_tables``: counting sorts, one pass per problem)."""
    lib = nat.load()
    C = np.ascontiguousarray(cams_per, dtype=np.int32)
    P = np.ascontiguousarray(pts_per, dtype=np.int32)
    N = np.array([np.asarray(o).size for o in obs_camera], dtype=np.int32)
    K = int(C.size)
    oc = np.ascontiguousarray(np.concatenate([np.asarray(o).reshape(-1) for o in obs_camera]) if K else [],
                              dtype=np.int32)
    op = np.ascontiguousarray(np.concatenate([np.asarray(o).reshape(-1) for o in obs_point]) if K else [],
                              dtype=np.int32)
    T = np.zeros(max(K, 1), dtype=np.int64)
    args = [K, _np_ptr(C), _np_ptr(P), _np_ptr(N), _np_ptr(oc), _np_ptr(op), T.ctypes.data]
    nat.check(lib.ct_ba_batch_tables(*args, *([None] * 8)), "ct_ba_batch_tables")
    Ntot = int(N.sum())
    B = C.astype(np.int64) * (C + 1) // 2
    out = dict(order=np.empty(Ntot, np.int32), obs_camera=np.empty(Ntot, np.int32), obs_point=np.empty(Ntot, np.int32),
               point_offsets=np.empty(int(P.sum()) + K, np.int32), camera_offsets=np.empty(int(C.sum()) + K, np.int32),
               camera_obs=np.empty(Ntot, np.int32), block_offsets=np.empty(int(B.sum()) + K, np.int32),
               block_terms=np.empty(2 * int(T[:K].sum()), np.int32))
    names = ["order", "obs_camera", "obs_point", "point_offsets", "camera_offsets", "camera_obs", "block_offsets",
             "block_terms"]
    # every output needs a valid host address, also an empty one
    bufs = [out[k] if out[k].size else np.zeros(1, np.int32) for k in names]
    nat.check(lib.ct_ba_batch_tables(*args, *[b.ctypes.data for b in bufs]), "ct_ba_batch_tables")
    T = T[:K]
    out["num_terms"] = T
    out["obase"] = np.cumsum(N, dtype=np.int64) - N
    out["cbase"] = np.cumsum(C, dtype=np.int64) - C
    out["pbase"] = np.cumsum(P, dtype=np.int64) - P
    out["tbase"] = np.cumsum(T) - T
    ar = np.arange(K, dtype=np.int64)
    out["poff_base"] = out["pbase"] + ar
    out["coff_base"] = out["cbase"] + ar
    out["boff_base"] = np.cumsum(B) - B + ar
    return out


class BABatch:
    """A batch of bundle adjustments resident on the device.

    ``launch`` enqueues the solve on the current stream without host
    synchronisation; ``solve`` runs it and returns per-problem reports,
    records and parameters on the host.
    """

    def __init__(self, problems: list[BAProblemInput], device="cuda", order: str = "largest"):
        nat.load()
        self.device = torch.device(device)
        self.num_problems = len(problems)
        cams = [np.asarray(pr.cameras, dtype=np.float64).reshape(-1, 9) for pr in problems]
        pts = [np.asarray(pr.points, dtype=np.float64).reshape(-1, 3) for pr in problems]
        C = np.array([c.shape[0] for c in cams], dtype=np.int64)
        P = np.array([p.shape[0] for p in pts], dtype=np.int64)
        if C.size and (C.min() < 1 or C.max() > MAX_CAMERAS):
            raise ValueError(f"a batched bundle adjustment has 1..{MAX_CAMERAS} cameras, got {int(C.max())}")
        ocs = [np.asarray(pr.obs_camera, dtype=np.int64).reshape(-1) for pr in problems]
        ops = [np.asarray(pr.obs_point, dtype=np.int64).reshape(-1) for pr in problems]
        for oc, op in zip(ocs, ops):
            if oc.shape != op.shape:
                raise ValueError("obs_camera and obs_point differ in length")
        tb = build_batch_tables(C, P, ocs, ops)  # checks every index
        N = np.array([oc.size for oc in ocs], dtype=np.int64)
        T = tb["num_terms"]
        xy = (np.concatenate([np.asarray(pr.observations, dtype=np.float64).reshape(-1, 2) for pr in problems])
              if problems else np.zeros((0, 2)))
        gorder = tb["order"] + np.repeat(tb["obase"], N).astype(np.int32)
        if not np.array_equal(gorder, np.arange(gorder.size, dtype=np.int32)):  # usually grouped by point already
            xy = xy[gorder]
        metas = []
        work = 256
        self.sizes = []
        for k, pr in enumerate(problems):
            m = nat.ct_ba_batch_problem()
            m.num_cameras, m.num_points, m.num_observations, m.num_terms = int(C[k]), int(P[k]), int(N[k]), int(T[k])
            m.camera_base, m.point_base, m.obs_base = int(tb["cbase"][k]), int(tb["pbase"][k]), int(tb["obase"][k])
            m.point_offsets_base, m.camera_offsets_base = int(tb["poff_base"][k]), int(tb["coff_base"][k])
            m.block_offsets_base, m.term_base = int(tb["boff_base"][k]), int(tb["tbase"][k])
            m.work_base = work
            m.loss = native_loss_struct(pr.loss_native)
            metas.append(m)
            work += problem_bytes(int(C[k]), int(P[k]), int(N[k]), int(T[k]))
            self.sizes.append((int(C[k]), int(P[k]), int(N[k]), int(T[k])))
        self.workspace_bytes = work
        dev = self.device

        def up(a, dtype):
            a = np.asarray(a, dtype=dtype).reshape(-1)
            if a.size == 0:
                a = np.zeros(1, dtype)  # a valid device address for empty tables
            return torch.from_numpy(np.ascontiguousarray(a)).to(dev)

        self.cameras = up(np.concatenate([c.reshape(-1) for c in cams]) if cams else [], np.float64)
        self.points = up(np.concatenate([p.reshape(-1) for p in pts]) if pts else [], np.float64)
        self.obs_camera = up(tb["obs_camera"], np.int32)
        self.obs_point = up(tb["obs_point"], np.int32)
        self.observations = up(xy, np.float64)
        self.point_offsets = up(tb["point_offsets"], np.int32)
        self.camera_offsets = up(tb["camera_offsets"], np.int32)
        self.camera_obs = up(tb["camera_obs"], np.int32)
        self.block_offsets = up(tb["block_offsets"], np.int32)
        self.block_terms = up(tb["block_terms"], np.int32)
        arr = (nat.ct_ba_batch_problem * max(1, self.num_problems))(*metas)
        raw = np.frombuffer(bytes(arr), dtype=np.uint8)
        self.problem_table = torch.from_numpy(raw.copy()).to(dev)
        # largest problems first, so the long ones start while the short ones fill in
        work_est = np.array([N + P + 20 * C * C for C, P, N, _ in self.sizes], dtype=np.int64)
        perm = np.argsort(-work_est, kind="stable") if order == "largest" else np.arange(self.num_problems)
        self.order = torch.from_numpy(np.ascontiguousarray(perm, dtype=np.int32)).to(dev)
        self.workspace = torch.empty(self.workspace_bytes, dtype=torch.uint8, device=dev)
        self.cam_slices = np.cumsum([0] + [s[0] * 9 for s in self.sizes])
        self.pt_slices = np.cumsum([0] + [s[1] * 3 for s in self.sizes])

    def batch_struct(self) -> nat.ct_ba_batch:
        b = nat.ct_ba_batch()
        b.num_problems = self.num_problems
        b.problems = nat.ptr(self.problem_table)
        b.obs_camera = nat.ptr(self.obs_camera)
        b.obs_point = nat.ptr(self.obs_point)
        b.observations = nat.ptr(self.observations)
        b.point_offsets = nat.ptr(self.point_offsets)
        b.camera_offsets = nat.ptr(self.camera_offsets)
        b.camera_obs = nat.ptr(self.camera_obs)
        b.block_offsets = nat.ptr(self.block_offsets)
        b.block_terms = nat.ptr(self.block_terms)
        b.order = nat.ptr(self.order)
        return b

    def launch(self, opts: nat.ct_lm_options, cameras: torch.Tensor, points: torch.Tensor, results=None,
               records=None) -> None:
        """Enqueue the batched solve on the current stream (parameters updated in place)."""
        lib = nat.load()
        b = self.batch_struct()
        nat.check(lib.ct_ba_batch_solve(ctypes.byref(b), ctypes.byref(opts), nat.ptr(cameras), nat.ptr(points),
                                        nat.ptr(results), nat.ptr(records), nat.ptr(self.workspace),
                                        self.workspace_bytes, nat.current_stream_handle()), "ct_ba_batch_solve")

    def solve(self, opts: nat.ct_lm_options, with_records: bool = False):
        """Solve every problem from its start values; returns (reports, records, cameras, points) on the host."""
        cams = self.cameras.clone()
        pts = self.points.clone()
        results = torch.zeros(max(1, self.num_problems) * nat.REPORT_DTYPE.itemsize, dtype=torch.uint8,
                              device=self.device)
        rec = None
        if with_records:
            rec = torch.zeros(max(1, self.num_problems) * (opts.max_num_iterations + 2) * nat.CT_REC_FIELDS,
                              dtype=torch.float64, device=self.device)
        self.launch(opts, cams, pts, results, rec)
        torch.cuda.current_stream().synchronize()
        reps = np.frombuffer(results.cpu().numpy().tobytes(), dtype=nat.REPORT_DTYPE)[: self.num_problems]
        recs = None
        if rec is not None:
            recs = rec.cpu().numpy().reshape(max(1, self.num_problems), opts.max_num_iterations + 2,
                                             nat.CT_REC_FIELDS)
        return reps, recs, cams.cpu().numpy(), pts.cpu().numpy()

    def split(self, flat_cameras: np.ndarray, flat_points: np.ndarray, k: int):
        """Problem k's cameras [C, 9] and points [P, 3] from the concatenated arrays."""
        c = flat_cameras[self.cam_slices[k]:self.cam_slices[k + 1]].reshape(-1, 9)
        p = flat_points[self.pt_slices[k]:self.pt_slices[k + 1]].reshape(-1, 3)
        return c, p
