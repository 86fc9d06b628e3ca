"""Robust loss functions (``src/ceres_torch/losses.py:13-144``).

Each loss keeps the reference's ``evaluate(sq_norm) -> (rho0, rho1, rho2)``
method (torch, used for introspection and by tests) and adds ``native()``:
the ``ct_loss`` descriptor that the device kernels evaluate.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import torch

LOSS_IDS = dict(trivial=0, huber=1, softlone=2, cauchy=3, arctan=4, tolerant=5, tukey=6)


def _like(v: float, t: torch.Tensor) -> torch.Tensor:
    return torch.as_tensor(v, dtype=t.dtype, device=t.device)


class LossFunction:
    kind = "trivial"

    def evaluate(self, sq_norm: torch.Tensor):
        raise NotImplementedError

    def __call__(self, sq_norm: torch.Tensor):
        return self.evaluate(sq_norm)

    def native(self) -> tuple[str, float, float, float]:
        """(kind, a, b, scale) for the device kernels."""
        raise NotImplementedError(f"{type(self).__name__} has no device form")


class TrivialLoss(LossFunction):
    def evaluate(self, s):
        return s, torch.ones_like(s), torch.zeros_like(s)

    def native(self):
        return ("trivial", 0.0, 0.0, 1.0)


@dataclass(frozen=True)
class HuberLoss(LossFunction):
    a: float

    def evaluate(self, s):
        b, a = _like(self.a * self.a, s), _like(self.a, s)
        inlier = s <= b
        root = torch.sqrt(torch.clamp(s, min=torch.finfo(s.dtype).tiny))
        return (
            torch.where(inlier, s, 2.0 * a * root - b),
            torch.where(inlier, torch.ones_like(s), a / root),
            torch.where(inlier, torch.zeros_like(s), -0.5 * a / (root * s)),
        )

    def native(self):
        return ("huber", float(self.a), 0.0, 1.0)


@dataclass(frozen=True)
class SoftLOneLoss(LossFunction):
    a: float

    def evaluate(self, s):
        b = _like(self.a * self.a, s)
        z = 1.0 + s / b
        rz = torch.sqrt(z)
        return 2.0 * b * (rz - 1.0), 1.0 / rz, -0.5 / (b * z * rz)

    def native(self):
        return ("softlone", float(self.a), 0.0, 1.0)


@dataclass(frozen=True)
class CauchyLoss(LossFunction):
    a: float

    def evaluate(self, s):
        b = _like(self.a * self.a, s)
        z = 1.0 + s / b
        return b * torch.log(z), 1.0 / z, -1.0 / (b * z * z)

    def native(self):
        return ("cauchy", float(self.a), 0.0, 1.0)


@dataclass(frozen=True)
class ArctanLoss(LossFunction):
    a: float

    def evaluate(self, s):
        a = _like(self.a, s)
        z = s / a
        d = 1.0 + z * z
        return a * torch.atan(z), 1.0 / d, -2.0 * z / (a * d * d)

    def native(self):
        return ("arctan", float(self.a), 0.0, 1.0)


@dataclass(frozen=True)
class TolerantLoss(LossFunction):
    a: float
    b: float

    def evaluate(self, s):
        a, b = _like(self.a, s), _like(self.b, s)
        z = (s - a) / b
        c = b * torch.nn.functional.softplus(-a / b)
        rho1 = torch.sigmoid(z)
        return b * torch.nn.functional.softplus(z) - c, rho1, rho1 * (1.0 - rho1) / b

    def native(self):
        return ("tolerant", float(self.a), float(self.b), 1.0)


@dataclass(frozen=True)
class TukeyLoss(LossFunction):
    a: float

    def evaluate(self, s):
        a2 = _like(self.a * self.a, s)
        inner = 1.0 - s / a2
        inside = s <= a2
        return (
            torch.where(inside, (a2 / 3.0) * (1.0 - inner**3), a2 / 3.0),
            torch.where(inside, inner**2, torch.zeros_like(s)),
            torch.where(inside, -2.0 * inner / a2, torch.zeros_like(s)),
        )

    def native(self):
        return ("tukey", float(self.a), 0.0, 1.0)


@dataclass
class ScaledLoss(LossFunction):
    rho: Optional[LossFunction]
    scale: float

    def evaluate(self, s):
        r0, r1, r2 = (self.rho or TrivialLoss()).evaluate(s)
        k = _like(self.scale, s)
        return k * r0, k * r1, k * r2

    def native(self):
        kind, a, b, scale = (self.rho or TrivialLoss()).native()
        return (kind, a, b, scale * float(self.scale))


@dataclass
class ComposedLoss(LossFunction):
    outer: LossFunction
    inner: LossFunction

    def evaluate(self, s):
        g0, g1, g2 = self.inner.evaluate(s)
        f0, f1, f2 = self.outer.evaluate(g0)
        return f0, f1 * g1, f2 * g1 * g1 + f1 * g2


class LossFunctionWrapper(LossFunction):
    def __init__(self, rho: Optional[LossFunction]) -> None:
        self.rho = rho

    def reset(self, rho: Optional[LossFunction]) -> None:
        self.rho = rho

    def evaluate(self, s):
        return (self.rho or TrivialLoss()).evaluate(s)

    def native(self):
        return (self.rho or TrivialLoss()).native()


def native_loss(loss: Optional[LossFunction]) -> tuple[str, float, float, float]:
    return ("trivial", 0.0, 0.0, 1.0) if loss is None else loss.native()


def oracle_loss(loss: Optional[LossFunction]) -> tuple:
    """The same loss in the tuple form the test oracle uses."""
    kind, a, b, scale = native_loss(loss)
    base = {"trivial": ("trivial",), "tolerant": ("tolerant", a, b)}.get(kind, (kind, a))
    return base if scale == 1.0 else ("scaled", base, scale)
