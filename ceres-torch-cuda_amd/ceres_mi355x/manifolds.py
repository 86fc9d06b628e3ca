"""Manifolds on the device path (``src/ceres_torch/manifolds.py:17-171``).

Euclidean Plus is ``x + delta`` (``manifolds.py:69-91``).  A SubsetManifold
holds chosen coordinates constant (``manifolds.py:94-124``); on the device it
becomes a per-parameter mask that zeroes those Jacobian columns, which leaves
their step at exactly zero - the same update the reduced tangent space gives.
QuaternionManifold / EigenQuaternionManifold (``manifolds.py:127-171``) move a
4-value rotation along its 3-dimensional tangent; on the device they are the
rotation of the quaternion camera models (the HIP kernels apply the same Plus
and the Jacobian ``J_ambient @ plus_jacobian``).
"""
from __future__ import annotations

from typing import Sequence

import torch

from .rotation import (convert_ceres_to_eigen_quaternion, convert_eigen_to_ceres_quaternion, normalize_quaternion,
                       quaternion_conjugate, quaternion_product)


class Manifold:
    ambient_size: int
    tangent_size: int

    def AmbientSize(self) -> int:
        return self.ambient_size

    def TangentSize(self) -> int:
        return self.tangent_size

    def plus(self, x: torch.Tensor, delta: torch.Tensor) -> torch.Tensor:
        raise NotImplementedError

    def Plus(self, x, delta):
        return self.plus(x, delta)

    def Minus(self, y, x):
        return self.minus(y, x)

    def PlusJacobian(self, x):
        return self.plus_jacobian(x)

    def free_mask(self) -> list[bool]:
        """Ambient coordinates that move under Plus (device form)."""
        raise NotImplementedError(f"{type(self).__name__} has no device form as a per-coordinate mask")


class EuclideanManifold(Manifold):
    def __init__(self, size: int) -> None:
        self.size = int(size)

    @property
    def ambient_size(self) -> int:
        return self.size

    @property
    def tangent_size(self) -> int:
        return self.size

    def plus(self, x, delta):
        return x.reshape(-1) + delta.reshape(-1)

    def minus(self, y, x):
        return y.reshape(-1) - x.reshape(-1)

    def free_mask(self) -> list[bool]:
        return [True] * self.size


class SubsetManifold(Manifold):
    def __init__(self, size: int, constant_parameters: Sequence[int]) -> None:
        self.size = int(size)
        self.constant_parameters = {int(i) for i in constant_parameters}
        if any(i < 0 or i >= self.size for i in self.constant_parameters):
            raise ValueError("constant parameter index out of range")
        self.variable_indices = [i for i in range(self.size) if i not in self.constant_parameters]

    @property
    def ambient_size(self) -> int:
        return self.size

    @property
    def tangent_size(self) -> int:
        return len(self.variable_indices)

    def plus(self, x, delta):
        out = x.reshape(-1).clone()
        out[self.variable_indices] = out[self.variable_indices] + delta.reshape(-1)
        return out

    def free_mask(self) -> list[bool]:
        return [i not in self.constant_parameters for i in range(self.size)]


class QuaternionManifold(Manifold):
    """Unit quaternions (w, x, y, z): Plus(x, d) = normalize(q_d * normalize(x)) with
    q_d = (cos|d|, sin|d| d / |d|) (manifolds.py:127-158)."""

    ambient_size = 4
    tangent_size = 3
    device_camera_model = "quaternion"

    def plus(self, x, delta):
        x = normalize_quaternion(x.reshape(4))
        delta = delta.reshape(3)
        nd = torch.linalg.norm(delta)
        eps = torch.finfo(delta.dtype).eps
        scale = torch.sin(nd) / nd if float(nd) > eps else 1.0 - nd * nd / 6.0
        q_delta = torch.cat([torch.cos(nd).reshape(1), scale * delta])
        return normalize_quaternion(quaternion_product(q_delta, x))

    def minus(self, y, x):
        q = quaternion_product(normalize_quaternion(y.reshape(4)), quaternion_conjugate(normalize_quaternion(x.reshape(4))))
        u = q[1:]
        un = torch.linalg.norm(u)
        if float(un) > torch.finfo(q.dtype).eps:
            return torch.atan2(un, q[0]) / un * u
        return u

    def plus_jacobian(self, x):
        w, a, b, c = normalize_quaternion(x.reshape(4))
        return torch.stack([torch.stack([-a, -b, -c]), torch.stack([w, c, -b]),
                            torch.stack([-c, w, a]), torch.stack([b, -a, w])])


class EigenQuaternionManifold(QuaternionManifold):
    """QuaternionManifold for quaternions stored (x, y, z, w) (manifolds.py:161-171)."""

    device_camera_model = "eigen_quaternion"

    def plus(self, x, delta):
        return convert_ceres_to_eigen_quaternion(super().plus(convert_eigen_to_ceres_quaternion(x.reshape(4)), delta))

    def minus(self, y, x):
        return super().minus(convert_eigen_to_ceres_quaternion(y.reshape(4)),
                             convert_eigen_to_ceres_quaternion(x.reshape(4)))

    def plus_jacobian(self, x):
        J = super().plus_jacobian(convert_eigen_to_ceres_quaternion(x.reshape(4)))
        return torch.cat([J[1:], J[:1]], dim=0)
