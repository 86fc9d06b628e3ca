"""Cost functions (``src/ceres_torch/costs.py:21-95``) and the typed residuals
the device path evaluates with hand-written HIP kernels.

``CostFunction``/``AutoDiffCostFunction``/``AnalyticCostFunction`` keep the
reference's interface (``residuals``, ``parameter_block_sizes``,
``num_residuals``) so problems can be described the same way, but only the
typed residuals below have device kernels:

* ``SnavelyReprojectionError``  BAL camera (9) + point (3) -> 2 residuals
  (``examples/simple_bundle_adjuster.py:73-79``)
* ``ExpCurveResidual``          m (1), c (1) -> ``y - exp(m x + c)``
  (``examples/robust_curve_fitting.py:87``)
* ``Pose2DRelativeError``       SE2 pose a (3), b (3) -> relative pose - measurement
  (``examples/pose_graph_2d.py:10-15``)
"""
from __future__ import annotations

from typing import Callable, Optional, Sequence

import torch

from .rotation import angle_axis_rotate_point, convert_eigen_to_ceres_quaternion, quaternion_rotate_point


def as_residual_tensor(value, like: Optional[torch.Tensor] = None) -> torch.Tensor:
    if isinstance(value, torch.Tensor):
        return value.reshape(-1)
    if like is not None:
        return torch.as_tensor(value, dtype=like.dtype, device=like.device).reshape(-1)
    return torch.as_tensor(value, dtype=torch.float64).reshape(-1)


class CostFunction:
    parameter_block_sizes: Optional[list[int]] = None
    num_residuals: Optional[int] = None
    device_kind: Optional[str] = None   # set by typed residuals with device kernels

    def residuals(self, *parameters: torch.Tensor) -> torch.Tensor:
        raise NotImplementedError

    def __call__(self, *parameters: torch.Tensor) -> torch.Tensor:
        return self.residuals(*parameters)


class CallableCostFunction(CostFunction):
    def __init__(self, functor: Callable[..., torch.Tensor], parameter_block_sizes: Optional[Sequence[int]] = None,
                 num_residuals: Optional[int] = None) -> None:
        self.functor = functor
        self.parameter_block_sizes = list(parameter_block_sizes) if parameter_block_sizes is not None else None
        self.num_residuals = num_residuals

    def residuals(self, *parameters):
        return as_residual_tensor(self.functor(*parameters), parameters[0] if parameters else None)


class AutoDiffCostFunction(CallableCostFunction):
    pass


class DynamicAutoDiffCostFunction(AutoDiffCostFunction):
    pass


class AnalyticCostFunction(CostFunction):
    def __init__(self, residual_fun, jacobian_fun, parameter_block_sizes=None, num_residuals=None) -> None:
        self.residual_fun = residual_fun
        self.jacobian_fun = jacobian_fun
        self.parameter_block_sizes = list(parameter_block_sizes) if parameter_block_sizes is not None else None
        self.num_residuals = num_residuals

    def residuals(self, *parameters):
        return as_residual_tensor(self.residual_fun(*parameters), parameters[0] if parameters else None)


def snavely_reprojection(camera: torch.Tensor, point: torch.Tensor) -> torch.Tensor:
    """Predicted pixel of ``point`` in a BAL camera (simple_bundle_adjuster.py:73-79)."""
    p = angle_axis_rotate_point(camera[:3], point) + camera[3:6]
    xp = -p[0] / p[2]
    yp = -p[1] / p[2]
    r2 = xp * xp + yp * yp
    distortion = 1.0 + r2 * (camera[7] + camera[8] * r2)
    return camera[6] * distortion * torch.stack([xp, yp])


class SnavelyReprojectionError(CostFunction):
    """Reprojection residual of one observation; blocks [camera(9), point(3)]."""

    parameter_block_sizes = [9, 3]
    num_residuals = 2
    device_kind = "ba"
    camera_model = "angle_axis"

    def __init__(self, observed_x: float, observed_y: float) -> None:
        self.observed_x = float(observed_x)
        self.observed_y = float(observed_y)

    def residuals(self, camera, point):
        obs = torch.tensor([self.observed_x, self.observed_y], dtype=camera.dtype, device=camera.device)
        return snavely_reprojection(camera, point) - obs


def quaternion_reprojection(rotation: torch.Tensor, camera: torch.Tensor, point: torch.Tensor,
                            order: str = "ceres") -> torch.Tensor:
    """Predicted pixel of ``point`` in a camera whose rotation is a quaternion.

    The BAL projection of simple_bundle_adjuster.py:73-79 with the rotation
    quaternion_rotate_point(q, X) (rotation.py:217-218) in place of the
    angle-axis one; ``camera`` is (t(3), f, k1, k2), ``order`` the storage
    order of ``rotation`` ("ceres": w, x, y, z; "eigen": x, y, z, w).
    """
    q = rotation.reshape(4)
    if order == "eigen":
        q = convert_eigen_to_ceres_quaternion(q)
    elif order != "ceres":
        raise ValueError("order must be 'ceres' or 'eigen'")
    p = quaternion_rotate_point(q, point) + camera[:3]
    xp = -p[0] / p[2]
    yp = -p[1] / p[2]
    r2 = xp * xp + yp * yp
    distortion = 1.0 + r2 * (camera[4] + camera[5] * r2)
    return camera[3] * distortion * torch.stack([xp, yp])


class QuaternionReprojectionError(CostFunction):
    """Reprojection residual of one observation; blocks [rotation(4), camera(6): t, f, k1, k2, point(3)].

    The rotation block carries a QuaternionManifold (``order="ceres"``) or an
    EigenQuaternionManifold (``order="eigen"``); on the device the two camera
    blocks form one quaternion camera of 10 values and 9 tangent coordinates.
    """

    parameter_block_sizes = [4, 6, 3]
    num_residuals = 2
    device_kind = "ba"

    def __init__(self, observed_x: float, observed_y: float, order: str = "ceres") -> None:
        if order not in ("ceres", "eigen"):
            raise ValueError("order must be 'ceres' or 'eigen'")
        self.observed_x = float(observed_x)
        self.observed_y = float(observed_y)
        self.order = order

    @property
    def camera_model(self) -> str:
        return "quaternion" if self.order == "ceres" else "eigen_quaternion"

    def residuals(self, rotation, camera, point):
        obs = torch.tensor([self.observed_x, self.observed_y], dtype=camera.dtype, device=camera.device)
        return quaternion_reprojection(rotation, camera, point, self.order) - obs


class ExpCurveResidual(CostFunction):
    """``y - exp(m x + c)`` for one sample; blocks [m(1), c(1)]."""

    parameter_block_sizes = [1, 1]
    num_residuals = 1
    device_kind = "curve"

    def __init__(self, x: float, y: float) -> None:
        self.x = float(x)
        self.y = float(y)

    def residuals(self, m, c):
        return (self.y - torch.exp(m[0] * self.x + c[0])).reshape(1)


def relative_pose(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """Pose of b in a's frame, angle wrapped (pose_graph_2d.py:10-15)."""
    ca, sa = torch.cos(a[2]), torch.sin(a[2])
    d = b[:2] - a[:2]
    local = torch.stack([ca * d[0] + sa * d[1], -sa * d[0] + ca * d[1]])
    dth = b[2] - a[2]
    return torch.cat([local, torch.atan2(torch.sin(dth), torch.cos(dth)).reshape(1)])


class Pose2DRelativeError(CostFunction):
    """``relative_pose(a, b) - measurement``; blocks [pose a(3), pose b(3)]."""

    parameter_block_sizes = [3, 3]
    num_residuals = 3
    device_kind = "pose2d"

    def __init__(self, dx: float, dy: float, dtheta: float) -> None:
        self.measurement = (float(dx), float(dy), float(dtheta))

    def residuals(self, a, b):
        m = torch.tensor(self.measurement, dtype=a.dtype, device=a.device)
        return relative_pose(a, b) - m
