"""Manifolds on the device path (``src/ceres_torch/manifolds.py:17-124``).

Euclidean Plus is ``x + delta`` (``manifolds.py:69-91``).  A SubsetManifold
holds chosen coordinates constant (``manifolds.py:94-124``); on the device it
becomes a per-parameter mask that zeroes those Jacobian columns, which leaves
their step at exactly zero - the same update the reduced tangent space gives.
"""
from __future__ import annotations

from typing import Sequence

import torch


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

    def free_mask(self) -> list[bool]:
        """Ambient coordinates that move under Plus (device form)."""
        raise NotImplementedError(f"{type(self).__name__} has no device form")


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
