"""Problem graph (``src/ceres_torch/problem.py:43-452``).

Parameter blocks are user tensors, aliased by ``id()`` and mutated in place by
``solve`` (the final state is the best state, ``problem.py:457-460``).
Residual blocks hold a cost function, an optional loss and their parameter
blocks in insertion order.  Besides the reference's one-block-at-a-time API,
``add_bundle_adjustment_observations`` adds many Snavely residual blocks
from index arrays without creating one Python object per observation.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np
import torch

from .costs import CallableCostFunction, CostFunction, SnavelyReprojectionError
from .losses import LossFunction
from .manifolds import EuclideanManifold, Manifold


@dataclass(eq=False)
class ParameterBlock:
    tensor: torch.Tensor
    manifold: Optional[Manifold] = None
    constant: bool = False
    name: Optional[str] = None
    lower_bound: Optional[torch.Tensor] = None
    upper_bound: Optional[torch.Tensor] = None
    ordering_group: Optional[int] = None
    has_explicit_manifold: bool = field(init=False)

    def __post_init__(self) -> None:
        if not isinstance(self.tensor, torch.Tensor):
            raise TypeError("ParameterBlock tensor must be a torch.Tensor")
        self.has_explicit_manifold = self.manifold is not None
        if self.manifold is None:
            self.manifold = EuclideanManifold(self.tensor.numel())
        if self.manifold.ambient_size != self.tensor.numel():
            raise ValueError("Manifold ambient size must equal parameter tensor numel")

    @property
    def size(self) -> int:
        return self.tensor.numel()

    @property
    def tangent_size(self) -> int:
        return 0 if self.constant else self.manifold.tangent_size

    def free_mask(self) -> list[bool]:
        if self.constant:
            return [False] * self.size
        return self.manifold.free_mask()


@dataclass(eq=False)
class ResidualBlock:
    cost_function: CostFunction
    loss_function: Optional[LossFunction]
    parameter_blocks: list[ParameterBlock]
    name: Optional[str] = None


@dataclass
class BulkBundleAdjustment:
    """Snavely residual blocks added from arrays (observation order kept)."""

    cameras: torch.Tensor          # [C, 9] base tensor of the camera blocks
    points: torch.Tensor           # [P, 3]
    camera_index: np.ndarray       # [N]
    point_index: np.ndarray        # [N]
    observations: np.ndarray       # [N, 2]
    loss: Optional[LossFunction]


@dataclass
class CRSMatrix:
    """Compressed-row copy of a dense matrix (``problem.py:15-41`` of the reference): the nonzero
    entries row by row, ``rows`` the row starts."""

    num_rows: int
    num_cols: int
    rows: list[int]
    cols: list[int]
    values: list[float]

    @classmethod
    def from_dense(cls, matrix: torch.Tensor) -> "CRSMatrix":
        m = matrix.detach()
        nz = torch.nonzero(m != 0, as_tuple=False)          # row-major order
        counts = torch.bincount(nz[:, 0], minlength=m.shape[0]) if nz.numel() else torch.zeros(
            m.shape[0], dtype=torch.int64, device=m.device)
        rows = [0] + torch.cumsum(counts, 0).cpu().tolist()
        vals = m[nz[:, 0], nz[:, 1]] if nz.numel() else m.new_zeros(0)
        return cls(int(m.shape[0]), int(m.shape[1]), rows, nz[:, 1].cpu().tolist(), vals.cpu().tolist())

    def to_dense(self, *, dtype: torch.dtype = torch.float64, device: torch.device | str = "cpu") -> torch.Tensor:
        matrix = torch.zeros((self.num_rows, self.num_cols), dtype=dtype)
        if self.cols:
            r = torch.repeat_interleave(torch.arange(self.num_rows), torch.diff(torch.tensor(self.rows)))
            matrix[r, torch.tensor(self.cols)] = torch.tensor(self.values, dtype=dtype)
        return matrix.to(device)


@dataclass
class EvaluateOptions:
    """``problem.py:100-105`` of the reference."""

    parameter_blocks: Optional[Sequence] = None
    residual_blocks: Optional[Sequence] = None
    apply_loss_function: bool = True
    num_threads: int = 1
    new_evaluation_point: bool = True


@dataclass
class EvaluationResult:
    cost: torch.Tensor
    residuals: torch.Tensor
    gradient: Optional[torch.Tensor] = None
    jacobian: Optional[torch.Tensor] = None
    crs_jacobian: Optional[CRSMatrix] = None


class Problem:
    def __init__(self, options=None, *, evaluation_callback=None) -> None:
        self.options = options
        self.parameter_blocks: list[ParameterBlock] = []
        self._residual_blocks: list[ResidualBlock] = []
        self._tensor_to_block: dict[int, ParameterBlock] = {}
        self.bulk_ba: list[BulkBundleAdjustment] = []
        self._row_blocks: dict[tuple[int, int], ParameterBlock] = {}

    # -- parameter blocks -------------------------------------------------
    def add_parameter_block(self, values: torch.Tensor, size: Optional[int] = None,
                            manifold: Optional[Manifold] = None, *, name: Optional[str] = None) -> ParameterBlock:
        if size is not None and values.numel() != size:
            raise ValueError(f"Expected parameter block size {size}, got {values.numel()}")
        block = self._tensor_to_block.get(id(values))
        if block is not None:
            if manifold is not None:
                block.manifold = manifold
                block.has_explicit_manifold = True
            if name is not None:
                block.name = name
            return block
        block = ParameterBlock(values, manifold=manifold, name=name)
        self.parameter_blocks.append(block)
        self._tensor_to_block[id(values)] = block
        return block

    AddParameterBlock = add_parameter_block

    def _coerce(self, value) -> ParameterBlock:
        if isinstance(value, ParameterBlock):
            return value
        block = self._tensor_to_block.get(id(value))
        return block if block is not None else self.add_parameter_block(value)

    def _require(self, value) -> ParameterBlock:
        if isinstance(value, ParameterBlock):
            if value not in self.parameter_blocks:
                raise KeyError("Parameter block is not part of this problem")
            return value
        block = self._tensor_to_block.get(id(value))
        if block is None:
            raise KeyError("Parameter block is not part of this problem")
        return block

    def set_parameter_block_constant(self, block) -> None:
        self._coerce(block).constant = True

    SetParameterBlockConstant = set_parameter_block_constant

    def set_parameter_block_variable(self, block) -> None:
        self._coerce(block).constant = False

    SetParameterBlockVariable = set_parameter_block_variable

    def is_parameter_block_constant(self, block) -> bool:
        b = self._coerce(block)
        return b.constant or b.manifold.tangent_size == 0

    IsParameterBlockConstant = is_parameter_block_constant

    def set_manifold(self, block, manifold: Optional[Manifold]) -> None:
        b = self._coerce(block)
        b.manifold = manifold or EuclideanManifold(b.size)
        b.has_explicit_manifold = manifold is not None
        if b.manifold.ambient_size != b.size:
            raise ValueError("Manifold ambient size must equal parameter block size")

    SetManifold = set_manifold

    def get_manifold(self, block) -> Optional[Manifold]:
        b = self._require(block)
        return b.manifold if b.has_explicit_manifold else None

    GetManifold = get_manifold

    def has_manifold(self, block) -> bool:
        return self._require(block).has_explicit_manifold

    HasManifold = has_manifold

    def set_parameter_block_ordering_group(self, block, group: Optional[int]) -> None:
        self._coerce(block).ordering_group = group

    SetParameterBlockOrderingGroup = set_parameter_block_ordering_group

    def get_parameter_block_ordering_group(self, block) -> Optional[int]:
        return self._require(block).ordering_group

    GetParameterBlockOrderingGroup = get_parameter_block_ordering_group

    def set_bounds(self, block, lower=None, upper=None) -> None:
        b = self._coerce(block)
        if lower is not None:
            b.lower_bound = torch.as_tensor(lower, dtype=b.tensor.dtype).broadcast_to(b.tensor.shape).clone()
        if upper is not None:
            b.upper_bound = torch.as_tensor(upper, dtype=b.tensor.dtype).broadcast_to(b.tensor.shape).clone()

    SetBounds = set_bounds

    # -- residual blocks --------------------------------------------------
    def add_residual_block(self, cost_function, loss_function: Optional[LossFunction],
                           parameter_blocks: Sequence, *, name: Optional[str] = None) -> ResidualBlock:
        if not isinstance(cost_function, CostFunction):
            cost_function = CallableCostFunction(cost_function)
        blocks = [self._coerce(p) for p in parameter_blocks]
        if cost_function.parameter_block_sizes is not None:
            expected = list(cost_function.parameter_block_sizes)
            got = [b.size for b in blocks]
            if expected != got:
                raise ValueError(f"Cost function parameter sizes {expected} do not match blocks {got}")
        rb = ResidualBlock(cost_function, loss_function, blocks, name=name)
        self._residual_blocks.append(rb)
        return rb

    AddResidualBlock = add_residual_block

    def add_bundle_adjustment_observations(self, cameras: torch.Tensor, points: torch.Tensor,
                                           camera_index, point_index, observations,
                                           loss: Optional[LossFunction] = None,
                                           point_ordering_group: Optional[int] = 0) -> None:
        """Add one SnavelyReprojectionError block per observation, in order.

        ``cameras`` [C, 9] and ``points`` [P, 3] are registered row by row as
        parameter blocks (cameras first), like
        ``examples/simple_bundle_adjuster.py:82-112``; points go to
        ``point_ordering_group``.
        """
        if cameras.ndim != 2 or cameras.shape[1] != 9 or points.ndim != 2 or points.shape[1] != 3:
            raise ValueError("cameras must be [C, 9] and points [P, 3]")
        for i in range(cameras.shape[0]):
            self._row_blocks[(id(cameras), i)] = self.add_parameter_block(cameras[i])
        for p in range(points.shape[0]):
            pb = self.add_parameter_block(points[p])
            self._row_blocks[(id(points), p)] = pb
            if point_ordering_group is not None:
                pb.ordering_group = point_ordering_group
        ci = np.asarray(camera_index, dtype=np.int64).reshape(-1)
        pi = np.asarray(point_index, dtype=np.int64).reshape(-1)
        obs = np.asarray(observations, dtype=np.float64).reshape(-1, 2)
        if not (ci.shape[0] == pi.shape[0] == obs.shape[0]):
            raise ValueError("camera_index, point_index and observations must have the same length")
        if ci.size and (ci.min() < 0 or ci.max() >= cameras.shape[0] or pi.min() < 0 or pi.max() >= points.shape[0]):
            raise ValueError("observation index out of range")
        self.bulk_ba.append(BulkBundleAdjustment(cameras, points, ci, pi, obs, loss))

    @property
    def residual_blocks(self) -> list[ResidualBlock]:
        if self.bulk_ba:
            self._materialize_bulk()
        return self._residual_blocks

    def _materialize_bulk(self) -> None:
        for bulk in self.bulk_ba:
            for k in range(bulk.camera_index.shape[0]):
                cost = SnavelyReprojectionError(bulk.observations[k, 0], bulk.observations[k, 1])
                self._residual_blocks.append(ResidualBlock(
                    cost, bulk.loss,
                    [self._row_blocks[(id(bulk.cameras), int(bulk.camera_index[k]))],
                     self._row_blocks[(id(bulk.points), int(bulk.point_index[k]))]]))
        self.bulk_ba = []

    def remove_residual_block(self, rb: ResidualBlock) -> None:
        self.residual_blocks.remove(rb)

    RemoveResidualBlock = remove_residual_block

    def remove_parameter_block(self, block) -> None:
        b = self._coerce(block)
        self._residual_blocks = [rb for rb in self.residual_blocks if b not in rb.parameter_blocks]
        self.parameter_blocks.remove(b)
        self._tensor_to_block.pop(id(b.tensor), None)

    RemoveParameterBlock = remove_parameter_block

    # -- introspection ----------------------------------------------------
    def get_parameter_blocks(self) -> list[ParameterBlock]:
        return list(self.parameter_blocks)

    GetParameterBlocks = get_parameter_blocks

    def get_residual_blocks(self) -> list[ResidualBlock]:
        return list(self.residual_blocks)

    GetResidualBlocks = get_residual_blocks

    def num_parameter_blocks(self) -> int:
        return len(self.parameter_blocks)

    NumParameterBlocks = num_parameter_blocks

    def num_parameters(self) -> int:
        return sum(b.size for b in self.parameter_blocks)

    NumParameters = num_parameters

    def num_effective_parameters(self) -> int:
        return sum(b.tangent_size for b in self.parameter_blocks)

    NumEffectiveParameters = num_effective_parameters

    def num_residual_blocks(self) -> int:
        return len(self._residual_blocks) + sum(int(b.camera_index.shape[0]) for b in self.bulk_ba)

    NumResidualBlocks = num_residual_blocks

    def num_residuals(self) -> int:
        total = 2 * sum(int(b.camera_index.shape[0]) for b in self.bulk_ba)
        for rb in self._residual_blocks:
            if rb.cost_function.num_residuals is None:
                params = [b.tensor.detach() for b in rb.parameter_blocks]
                total += rb.cost_function.residuals(*params).numel()
            else:
                total += rb.cost_function.num_residuals
        return int(total)

    NumResiduals = num_residuals

    def has_parameter_block(self, block) -> bool:
        if isinstance(block, ParameterBlock):
            return block in self.parameter_blocks
        return id(block) in self._tensor_to_block

    HasParameterBlock = has_parameter_block

    def parameter_block_size(self, block) -> int:
        return self._require(block).size

    ParameterBlockSize = parameter_block_size

    def parameter_block_tangent_size(self, block) -> int:
        return self._require(block).manifold.tangent_size

    ParameterBlockTangentSize = parameter_block_tangent_size

    # -- state ------------------------------------------------------------
    def snapshot(self) -> list[torch.Tensor]:
        return [b.tensor.detach().clone() for b in self.parameter_blocks]

    def restore(self, snapshot: Sequence[torch.Tensor]) -> None:
        with torch.no_grad():
            for b, v in zip(self.parameter_blocks, snapshot):
                b.tensor.reshape(-1).copy_(v.to(dtype=b.tensor.dtype, device=b.tensor.device).reshape(-1))

    def evaluate(self, options: Optional[EvaluateOptions] = None, *, compute_jacobian: bool = True) -> EvaluationResult:
        """``Problem.evaluate`` (``problem.py:377-413``) on the device: cost, loss-corrected residuals in
        residual-block order and, with ``compute_jacobian``, the dense Jacobian over the active tangent
        columns (blocks in problem order, or ``options.parameter_blocks``), ``g = J^T r`` and its CRS copy."""
        from .solver import evaluate_problem

        return evaluate_problem(self, options or EvaluateOptions(), compute_jacobian)

    Evaluate = evaluate
