"""Covariance of bundle-adjustment estimates on the device (``src/ceres_torch/covariance.py``).

The reference builds the dense Jacobian, takes its SVD (DENSE_SVD) or QR
(SPARSE_QR) and forms the whole tangent covariance ``(J^T J)^+``
(``covariance.py:86-276``), then slices the requested blocks.  For a
full-rank Jacobian both forms are ``(J^T J)^-1``.  Here nothing dense in the
problem size is formed: the HIP library assembles the undamped reduced camera
system at the user's parameters, factors it in FP64 and writes only the
requested blocks (``ct_ba_plan_covariance``):

* camera-camera blocks are blocks of ``S^-1``;
* point-camera and point-point blocks are short sums over the points'
  observations of the point factors, the camera Jacobians and ``S^-1``.

Rank: the SVD's threshold ``20 (m + n) eps`` (``covariance.py:374-377``) is
applied to the Cholesky pivots of S and of each point's 3x3 block; a rank
deficient Jacobian (a free gauge, a point seen by one camera) fails as in the
reference, ``"Jacobian is rank deficient."``, with rank = columns minus the
deficient pivots.  Not on the device path (``NotImplementedError``): the
pseudo-inverse forms (``null_space_rank != 0``), problems other than bundle
adjustment, and the spectrum fields of the summary (the device never forms
the singular values of J; they are reported as NaN, and
``min_reciprocal_condition_number`` is not tested).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np
import torch

from . import _native as nat
from .problem import ParameterBlock, Problem
from .types import CovarianceAlgorithmType, SparseLinearAlgebraLibraryType

_EPS = float(np.finfo(np.float64).eps)


@dataclass
class CovarianceOptions:
    """covariance.py:14-32."""

    sparse_linear_algebra_library_type: SparseLinearAlgebraLibraryType = SparseLinearAlgebraLibraryType.NO_SPARSE
    algorithm_type: CovarianceAlgorithmType = CovarianceAlgorithmType.DENSE_SVD
    column_pivot_threshold: float = -1.0
    min_reciprocal_condition_number: float = 1e-14
    null_space_rank: int = 0
    num_threads: int = 1
    apply_loss_function: bool = True

    def validate(self) -> None:
        for ok, message in (
            (self.column_pivot_threshold >= -1.0, "column_pivot_threshold must be >= -1"),
            (self.min_reciprocal_condition_number >= 0.0, "min_reciprocal_condition_number must be >= 0"),
            (self.null_space_rank >= -1, "null_space_rank must be >= -1"),
            (self.num_threads > 0, "num_threads must be > 0"),
        ):
            if not ok:
                raise ValueError(message)


@dataclass
class CovarianceSummary:
    """covariance.py:35-83 (same fields and reports)."""

    algorithm_type: CovarianceAlgorithmType
    success: bool = False
    message: str = "Covariance.compute was not called."
    rank: int = 0
    nullity: int = 0
    num_rows: int = 0
    num_columns: int = 0
    num_requested_blocks: int = 0
    num_computed_blocks: int = 0
    ambient_dimension: int = 0
    tangent_dimension: int = 0
    max_singular_value: float = 0.0
    min_retained_singular_value: float = 0.0
    reciprocal_condition_number: float = 0.0
    requested_null_space_rank: int = 0

    def BriefReport(self) -> str:
        return (f"ceres-torch Covariance Report: Algorithm: {self.algorithm_type.value}, Rank: {self.rank}, "
                f"Nullity: {self.nullity}, Success: {self.success}")

    def FullReport(self) -> str:
        return "\n".join([
            "Covariance Summary (ceres-torch)", "",
            f"Algorithm: {self.algorithm_type.value}",
            f"Success: {self.success}",
            f"Message: {self.message}",
            f"Rows: {self.num_rows}",
            f"Columns: {self.num_columns}",
            f"Ambient dimension: {self.ambient_dimension}",
            f"Tangent dimension: {self.tangent_dimension}",
            f"Requested blocks: {self.num_requested_blocks}",
            f"Computed blocks: {self.num_computed_blocks}",
            f"Rank: {self.rank}",
            f"Nullity: {self.nullity}",
            f"Requested null space rank: {self.requested_null_space_rank}",
            f"Max singular value: {self.max_singular_value:.12e}",
            f"Min retained singular value: {self.min_retained_singular_value:.12e}",
            f"Reciprocal condition number: {self.reciprocal_condition_number:.12e}",
        ])


def _copy_to_output(value: torch.Tensor, out: Optional[torch.Tensor]) -> torch.Tensor:
    """covariance.py:456-462."""
    if out is None:
        return value
    if out.numel() != value.numel():
        raise ValueError("Output tensor has the wrong number of elements")
    out.reshape(-1).copy_(value.reshape(-1).to(dtype=out.dtype, device=out.device))
    return out


class Covariance:
    """covariance.py:86-446 for bundle-adjustment problems, computed by the HIP library."""

    def __init__(self, options: Optional[CovarianceOptions] = None) -> None:
        self.options = options or CovarianceOptions()
        self._problem: Optional[Problem] = None
        self._tangent: dict = {}
        self._blocks: set = set()
        self.summary = CovarianceSummary(algorithm_type=self.options.algorithm_type,
                                         requested_null_space_rank=self.options.null_space_rank)

    # -- compute -------------------------------------------------------------
    def compute(self, covariance_blocks, problem: Problem) -> bool:
        from .solver import _BAExtract

        self.options.validate()
        self.summary = CovarianceSummary(algorithm_type=self.options.algorithm_type,
                                         requested_null_space_rank=self.options.null_space_rank)
        self._problem = problem
        self._tangent = {}
        if not covariance_blocks:
            self._blocks = set()
            self.summary.success = True
            self.summary.message = "No covariance blocks requested."
            self.summary.ambient_dimension = problem.num_parameters()
            self.summary.tangent_dimension = problem.num_effective_parameters()
            return True
        pairs = self._normalize_covariance_blocks(covariance_blocks, problem)
        self._blocks = set(pairs) | {(b, a) for a, b in pairs}
        dense_svd = self.options.algorithm_type is CovarianceAlgorithmType.DENSE_SVD
        if dense_svd and self.options.null_space_rank != 0:
            raise NotImplementedError("the pseudo-inverse covariance (null_space_rank != 0) is not on the MI355X "
                                      "device path")
        ex = _BAExtract(problem)  # raises NotImplementedError for other problems
        if ex.camera_model != "angle_axis":
            raise NotImplementedError("the covariance of quaternion cameras is not on the MI355X device path")
        s = self.summary
        s.num_requested_blocks = len(pairs)
        s.num_computed_blocks = len(self._blocks)
        s.ambient_dimension = problem.num_parameters()
        s.tangent_dimension = sum(b.tangent_size for b in ex.cam_blocks + ex.pt_blocks)
        s.num_rows = 2 * len(ex.obs_camera)
        s.num_columns = s.tangent_dimension
        s.max_singular_value = s.min_retained_singular_value = s.reciprocal_condition_number = math.nan
        if s.tangent_dimension == 0:
            s.success, s.message = True, "No active tangent parameters."
            return True
        if not dense_svd and s.num_rows < s.num_columns:
            s.success = False
            s.message = "Sparse QR covariance requires rows >= columns and cannot use null_space_rank."
            s.rank, s.nullity = s.num_rows, s.num_columns - s.num_rows
            return False
        where = {}
        for i, b in enumerate(ex.cam_blocks):
            where[b] = (0, i)
        for i, b in enumerate(ex.pt_blocks):
            where[b] = (1, i)
        unique = sorted({(a, b) if id(a) <= id(b) else (b, a) for a, b in pairs}, key=lambda ab: (id(ab[0]), id(ab[1])))
        unique = [ab for ab in unique if ab[0].tangent_size > 0 and ab[1].tangent_size > 0 and ab[0] in where
                  and ab[1] in where]
        blocks, deficient = self._device_blocks(problem, ex, [(where[a], where[b]) for a, b in unique])
        if blocks is None:
            s.rank = s.num_columns - deficient
            s.nullity = deficient
            s.success = False
            s.message = ("Jacobian is rank deficient." if dense_svd
                         else "Sparse QR detected a rank deficient Jacobian and cannot use null_space_rank.")
            return False
        for (a, b), blk in zip(unique, blocks):
            fa = np.nonzero(a.free_mask())[0]
            fb = np.nonzero(b.free_mask())[0]
            t = blk[: a.size, : b.size][np.ix_(fa, fb)]
            self._tangent[(a, b)] = t
            self._tangent[(b, a)] = t.T.contiguous()
        s.rank, s.nullity = s.num_columns, 0
        s.success = True
        s.message = "Computed covariance with dense SVD." if dense_svd else "Computed covariance with QR."
        return True

    Compute = compute

    def _device_blocks(self, problem, ex, requests):
        from .ba import BADevice, lm_options
        from .losses import native_loss
        from .types import SolverOptions, TrustRegionStrategyType

        cams, pts = ex.stacked()
        cm, pm = ex.masks()
        loss = ex.loss if self.options.apply_loss_function else native_loss(None)
        dev = BADevice(cams, pts, ex.obs_camera, ex.obs_point, ex.observations, loss_native=loss, camera_mask=cm,
                       point_mask=pm)
        lib = nat.load()
        try:
            opts = lm_options(SolverOptions(max_num_iterations=1,
                                            trust_region_strategy_type=TrustRegionStrategyType.DOGLEG),
                              precision="fp64", loop="host")
            handle = dev.plan(opts)
            R = len(requests)
            req = np.asarray([[ka, ia, kb, ib] for (ka, ia), (kb, ib) in requests], dtype=np.int32).reshape(-1)
            req = np.ascontiguousarray(req)
            out = torch.empty((max(R, 1), 81), dtype=torch.float64, device=dev.device)
            nbytes = ctypes.c_size_t(0)
            nat.check(lib.ct_ba_covariance_workspace_bytes(dev.num_cameras, R, ctypes.byref(nbytes)),
                      "ct_ba_covariance_workspace_bytes")
            ws = torch.empty(int(nbytes.value), dtype=torch.uint8, device=dev.device)
            rows = 2 * dev.num_observations
            cols = sum(b.tangent_size for b in ex.cam_blocks + ex.pt_blocks)
            thr = self.options.column_pivot_threshold
            rtol = 20.0 * (rows + cols) * _EPS if thr < 0 else min(thr, 0.5)
            deficient = ctypes.c_int64(0)
            rc = lib.ct_ba_plan_covariance(handle, req.ctypes.data_as(ctypes.c_void_p), R, rtol, nat.ptr(out),
                                           ctypes.byref(deficient), nat.ptr(ws), ws.numel())
            if rc == nat.CT_ERR_NOT_SPD:
                return None, int(deficient.value)
            nat.check(rc, "ct_ba_plan_covariance")
            blocks = []
            for k, ((ka, _), (kb, _)) in enumerate(requests):
                ra, rb = (9 if ka == 0 else 3), (9 if kb == 0 else 3)
                blocks.append(out[k, : ra * rb].reshape(ra, rb).cpu())
            return blocks, 0
        finally:
            dev.close()

    # -- queries (covariance.py:279-446) --------------------------------------
    def get_covariance_block(self, a, b, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        pa, pb = self._require(a), self._require(b)
        t = self._get_tangent_block(pa, pb)
        amb = torch.zeros((pa.size, pb.size), dtype=torch.float64)
        fa = torch.as_tensor(np.nonzero(pa.free_mask())[0]) if pa.tangent_size else None
        fb = torch.as_tensor(np.nonzero(pb.free_mask())[0]) if pb.tangent_size else None
        if fa is not None and fb is not None:
            amb[fa[:, None], fb[None, :]] = t
        return _copy_to_output(amb, out)

    def get_covariance_block_in_tangent_space(self, a, b, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        return _copy_to_output(self._get_tangent_block(self._require(a), self._require(b)), out)

    def get_covariance_matrix(self, parameter_blocks: Sequence, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        rows = [torch.cat([self.get_covariance_block(a, b) for b in parameter_blocks], dim=1) for a in parameter_blocks]
        return _copy_to_output(torch.cat(rows, dim=0), out)

    def get_covariance_matrix_in_tangent_space(self, parameter_blocks: Sequence,
                                               out: Optional[torch.Tensor] = None) -> torch.Tensor:
        rows = [torch.cat([self.get_covariance_block_in_tangent_space(a, b) for b in parameter_blocks], dim=1)
                for a in parameter_blocks]
        return _copy_to_output(torch.cat(rows, dim=0), out)

    GetCovarianceBlock = get_covariance_block
    GetCovarianceBlockInTangentSpace = get_covariance_block_in_tangent_space
    GetCovarianceMatrix = get_covariance_matrix
    GetCovarianceMatrixInTangentSpace = get_covariance_matrix_in_tangent_space

    def has_covariance_block(self, a, b) -> bool:
        if self._problem is None:
            return False
        try:
            pa, pb = self._require(a), self._require(b)
        except KeyError:
            return False
        return (pa, pb) in self._blocks

    HasCovarianceBlock = has_covariance_block

    def rank(self) -> int:
        return self.summary.rank

    def nullity(self) -> int:
        return self.summary.nullity

    def reciprocal_condition_number(self) -> float:
        return self.summary.reciprocal_condition_number

    Rank = rank
    Nullity = nullity
    ReciprocalConditionNumber = reciprocal_condition_number

    # -- helpers -------------------------------------------------------------
    def _require(self, value) -> ParameterBlock:
        if self._problem is None:
            raise RuntimeError("Covariance.compute must be called first")
        return self._problem._require(value)

    def _get_tangent_block(self, pa: ParameterBlock, pb: ParameterBlock) -> torch.Tensor:
        """covariance.py:354-364."""
        if self._problem is None:
            raise RuntimeError("Covariance.compute must be called first")
        if not self.summary.success:
            raise RuntimeError(f"Covariance.compute did not succeed: {self.summary.message}")
        if (pa, pb) not in self._blocks:
            raise KeyError("Requested covariance block was not computed")
        if pa.tangent_size == 0 or pb.tangent_size == 0:
            return torch.zeros((pa.tangent_size, pb.tangent_size), dtype=torch.float64)
        return self._tangent[(pa, pb)]

    @staticmethod
    def _normalize_covariance_blocks(covariance_blocks, problem: Problem):
        """covariance.py:424-446."""
        first = covariance_blocks[0]
        if isinstance(first, tuple):
            pairs = [(problem._require(a), problem._require(b)) for a, b in covariance_blocks]
            seen = set()
            for a, b in pairs:
                key = frozenset((a, b))
                if key in seen:
                    raise ValueError("covariance_blocks cannot contain duplicate block pairs")
                seen.add(key)
            return pairs
        params = [problem._require(p) for p in covariance_blocks]
        if len(set(params)) != len(params):
            raise ValueError("parameter_blocks cannot contain duplicates")
        return [(a, b) for a in params for b in params]
