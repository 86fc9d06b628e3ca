"""Robust losses and the Ceres corrector, vectorised over residual blocks.

Follows ``src/ceres_torch/losses.py``: each loss maps the squared block norm
``s`` to ``(rho0, rho1, rho2)`` (``losses.py:21-144``) and
``robustify_residual_and_jacobian`` (``losses.py:158-193``) rescales the block
residual and Jacobian.  Losses are given as tuples so that both the oracle and
the device path can share one description:

    ("trivial",) ("huber", a) ("softlone", a) ("cauchy", a) ("arctan", a)
    ("tolerant", a, b) ("tukey", a) ("scaled", inner_tuple_or_None, scale)
"""
from __future__ import annotations

import numpy as np

_TINY = np.finfo(np.float64).tiny


def _softplus(x: np.ndarray) -> np.ndarray:
    return np.where(x > 20.0, x, np.log1p(np.exp(np.minimum(x, 20.0))))


def loss_rho(loss: tuple | None, s: np.ndarray) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Return (rho0, rho1, rho2) for squared norms ``s``."""
    s = np.asarray(s, dtype=np.float64)
    if loss is None or loss[0] == "trivial":
        # losses.py:21-23
        return s, np.ones_like(s), np.zeros_like(s)
    kind = loss[0]
    if kind == "huber":
        # losses.py:26-38
        a = float(loss[1])
        b = a * a
        mask = s <= b
        sqrt_s = np.sqrt(np.maximum(s, _TINY))
        rho0 = np.where(mask, s, 2.0 * a * sqrt_s - b)
        rho1 = np.where(mask, 1.0, a / sqrt_s)
        rho2 = np.where(mask, 0.0, -0.5 * a / (sqrt_s * s))
        return rho0, rho1, rho2
    if kind == "softlone":
        # losses.py:41-52
        b = float(loss[1]) ** 2
        z = 1.0 + s / b
        sz = np.sqrt(z)
        return 2.0 * b * (sz - 1.0), 1.0 / sz, -0.5 / (b * z * sz)
    if kind == "cauchy":
        # losses.py:55-65
        b = float(loss[1]) ** 2
        z = 1.0 + s / b
        return b * np.log(z), 1.0 / z, -1.0 / (b * z * z)
    if kind == "arctan":
        # losses.py:68-79
        a = float(loss[1])
        z = s / a
        d = 1.0 + z * z
        return a * np.arctan(z), 1.0 / d, -2.0 * z / (a * d * d)
    if kind == "tolerant":
        # losses.py:82-95
        a, b = float(loss[1]), float(loss[2])
        z = (s - a) / b
        c = b * _softplus(np.asarray(-a / b))
        rho1 = 1.0 / (1.0 + np.exp(-z))
        return b * _softplus(z) - c, rho1, rho1 * (1.0 - rho1) / b
    if kind == "tukey":
        # losses.py:98-110
        a2 = float(loss[1]) ** 2
        z = s / a2
        inner = 1.0 - z
        mask = s <= a2
        return (
            np.where(mask, (a2 / 3.0) * (1.0 - inner**3), a2 / 3.0),
            np.where(mask, inner**2, 0.0),
            np.where(mask, -2.0 * inner / a2, 0.0),
        )
    if kind == "scaled":
        # losses.py:113-122
        inner, scale = loss[1], float(loss[2])
        r0, r1, r2 = loss_rho(inner, s)
        return scale * r0, scale * r1, scale * r2
    raise ValueError(f"unknown loss {loss!r}")


def robustify(loss: tuple | None, r: np.ndarray, J: np.ndarray | None):
    """Corrected (cost, r, J) for a batch of blocks.

    ``r`` has shape [B, m]; ``J`` has shape [B, m, n] or is None.  Mirrors
    ``losses.py:158-193``: residuals scale by sqrt(rho1) and, where s > 0 and
    rho2 > 0, by the Triggs factor 1/(1-alpha) with the rank-one Jacobian
    correction ``J <- sqrt(rho1) (J - alpha/s r r^T J)``.
    """
    s = np.sum(r * r, axis=1)
    rho0, rho1, rho2 = loss_rho(loss, s)
    sqrt_rho1 = np.sqrt(np.maximum(rho1, 0.0))
    scale = sqrt_rho1.copy()
    alpha_sq = np.zeros_like(s)
    triggs = (s > 0.0) & (rho2 > 0.0)
    if np.any(triggs):
        if np.any(rho1[triggs] <= 0.0):
            raise ValueError("Loss function has rho'' > 0 but rho' <= 0")
        disc = 1.0 + 2.0 * s[triggs] * rho2[triggs] / rho1[triggs]
        alpha = 1.0 - np.sqrt(disc)
        scale[triggs] = sqrt_rho1[triggs] / (1.0 - alpha)
        alpha_sq[triggs] = alpha / s[triggs]
    r_out = scale[:, None] * r
    J_out = None
    if J is not None:
        rTJ = np.einsum("bm,bmn->bn", r, J)
        J_out = sqrt_rho1[:, None, None] * (J - alpha_sq[:, None, None] * r[:, :, None] * rTJ[:, None, :])
    return 0.5 * rho0, r_out, J_out
