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
from .ba import build_tables, native_loss_struct

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


def problem_bytes(num_cameras: int, num_points: int, num_observations: int) -> int:
    """Device workspace of one problem in a batch (exact, from the library)."""
    lib = nat.load()
    out = ctypes.c_size_t(0)
    nat.check(lib.ct_ba_batch_problem_bytes(int(num_cameras), int(num_points), int(num_observations),
                                            ctypes.byref(out)), "ct_ba_batch_problem_bytes")
    return int(out.value)


def device_bytes(num_cameras: int, num_points: int, num_observations: int, num_terms: int) -> int:
    """Every device byte a problem adds to a batch: its inputs, tables, parameters and workspace slice."""
    C, P, N = int(num_cameras), int(num_points), int(num_observations)
    B = C * (C + 1) // 2
    tables = 4 * (3 * N + (P + 1) + (C + 1) + (B + 1) + 2 * int(num_terms))
    return problem_bytes(C, P, N) + tables + 16 * N + 8 * (9 * C + 3 * P) + 128


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
        metas, cams, pts, ocam, opt, obs, poff, coff, cobs, boff, terms = ([] for _ in range(11))
        base = dict(camera=0, point=0, obs=0, point_offsets=0, camera_offsets=0, block_offsets=0, term=0)
        work = 256
        self.sizes = []
        for pr in problems:
            cam = np.asarray(pr.cameras, dtype=np.float64).reshape(-1, 9)
            pt = np.asarray(pr.points, dtype=np.float64).reshape(-1, 3)
            C, P = cam.shape[0], pt.shape[0]
            if not 1 <= C <= MAX_CAMERAS:
                raise ValueError(f"a batched bundle adjustment has 1..{MAX_CAMERAS} cameras, got {C}")
            oc = np.asarray(pr.obs_camera, dtype=np.int64).reshape(-1)
            op = np.asarray(pr.obs_point, dtype=np.int64).reshape(-1)
            if oc.shape != op.shape:
                raise ValueError("obs_camera and obs_point differ in length")
            if oc.size and (oc.min() < 0 or oc.max() >= C or op.min() < 0 or op.max() >= P):
                raise ValueError("observation index out of range")
            t = build_tables(oc, op, C, P)
            N = int(oc.size)
            xy = np.asarray(pr.observations, dtype=np.float64).reshape(-1, 2)[t.order]
            T = int(t.block_terms.shape[0])
            m = nat.ct_ba_batch_problem()
            m.num_cameras, m.num_points, m.num_observations = C, P, N
            m.camera_base, m.point_base, m.obs_base = base["camera"], base["point"], base["obs"]
            m.point_offsets_base, m.camera_offsets_base = base["point_offsets"], base["camera_offsets"]
            m.block_offsets_base, m.term_base = base["block_offsets"], base["term"]
            m.work_base = work
            m.loss = native_loss_struct(pr.loss_native)
            metas.append(m)
            cams.append(cam)
            pts.append(pt)
            ocam.append(t.obs_camera)
            opt.append(t.obs_point)
            obs.append(xy)
            poff.append(t.point_offsets)
            coff.append(t.camera_offsets)
            cobs.append(t.camera_obs)
            boff.append(t.block_offsets)
            terms.append(t.block_terms.reshape(-1))
            base["camera"] += C
            base["point"] += P
            base["obs"] += N
            base["point_offsets"] += P + 1
            base["camera_offsets"] += C + 1
            base["block_offsets"] += C * (C + 1) // 2 + 1
            base["term"] += T
            work += problem_bytes(C, P, N)
            self.sizes.append((C, P, N, T))
        self.workspace_bytes = work
        dev = self.device

        def cat(parts, dtype):
            a = np.concatenate(parts).astype(dtype) if parts else np.zeros(0, dtype)
            if a.size == 0:
                a = np.zeros(1, dtype)  # a valid device address for empty tables
            return torch.from_numpy(np.ascontiguousarray(a)).to(dev)

        self.cameras = cat([c.reshape(-1) for c in cams], np.float64)
        self.points = cat([p.reshape(-1) for p in pts], np.float64)
        self.obs_camera = cat(ocam, np.int32)
        self.obs_point = cat(opt, np.int32)
        self.observations = cat([o.reshape(-1) for o in obs], np.float64)
        self.point_offsets = cat(poff, np.int32)
        self.camera_offsets = cat(coff, np.int32)
        self.camera_obs = cat(cobs, np.int32)
        self.block_offsets = cat(boff, np.int32)
        self.block_terms = cat(terms, np.int32)
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
