"""Curve-fit and 2D pose-graph models for the oracle's trust-region loop.

Test infrastructure only.

* Curve fit: one residual block per sample ``y - exp(m x + c)`` with a robust
  loss on each block and parameter blocks ``m`` then ``c``
  (``examples/robust_curve_fitting.py:80-97``).  The reference's default
  linear solver is DENSE_QR: a damped least-squares solve of
  ``[J; sqrt(D)] dx = [-r; 0]`` (``linear.py:359-371``).
* Pose graph: SE2 poses (x, y, theta), residual ``relative_pose(a, b) - m``
  (``examples/pose_graph_2d.py:6-15``), first pose constant
  (``pose_graph_2d.py:34``), DENSE_QR (``pose_graph_2d.py:42-45``).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .lm import _solve_square, dense_linear_solve
from .losses import robustify


@dataclass
class DenseEval:
    cost: float
    r: np.ndarray          # flat corrected residuals
    J: np.ndarray | None   # dense corrected Jacobian (rows x active columns)


class CurveModel:
    def __init__(self, x: np.ndarray, y: np.ndarray, loss=("cauchy", 0.5), m0: float = 0.0, c0: float = 0.0,
                 linear_solver: str = "dense_qr"):
        self.xs = np.asarray(x, dtype=np.float64)
        self.ys = np.asarray(y, dtype=np.float64)
        self.loss = loss
        self.start = np.array([m0, c0], dtype=np.float64)
        self.linear_solver = linear_solver

    def x0(self):
        return self.start.copy()

    def num_active_columns(self) -> int:
        return 2

    def state_norm(self, x) -> float:
        return float(np.sqrt(np.sum(x * x)))

    def apply(self, x, step):
        return x + step

    def evaluate(self, x, jacobian: bool):
        e = np.exp(x[0] * self.xs + x[1])
        res = (self.ys - e)[:, None]
        if jacobian:
            J = np.stack([-self.xs * e, -e], axis=-1)[:, None, :]
            cost_b, r, Jc = robustify(self.loss, res, J)
            ev = DenseEval(float(np.sum(cost_b)), r.reshape(-1), Jc.reshape(-1, 2))
            return ev, ev.J.T @ ev.r
        cost_b, r, _ = robustify(self.loss, res, None)
        return DenseEval(float(np.sum(cost_b)), r.reshape(-1), None), None

    def column_sq(self, ev) -> np.ndarray:
        return np.sum(ev.J * ev.J, axis=0)

    def solve(self, ev, damping, opts):
        return dense_linear_solve(ev.J, -ev.r, damping, 0, self.linear_solver, opts)

    def model_decrease(self, ev, step) -> float:
        after = ev.r + ev.J @ step
        return max(0.0, float(0.5 * (ev.r @ ev.r) - 0.5 * (after @ after)))

    def gauss_newton_step(self, ev) -> np.ndarray:
        """linear.py:570-571: solve(J^T J, -J^T r)."""
        return _solve_square(ev.J.T @ ev.J, -(ev.J.T @ ev.r))

    def jmul(self, ev, v) -> np.ndarray:
        return ev.J @ v


def relative_pose_residual(a: np.ndarray, b: np.ndarray, meas: np.ndarray, jacobian: bool):
    """Residual [E,3] and Jacobians [E,3,3] wrt a and b (pose_graph_2d.py:10-15)."""
    ca, sa = np.cos(a[:, 2]), np.sin(a[:, 2])
    dx = b[:, 0] - a[:, 0]
    dy = b[:, 1] - a[:, 1]
    dth = b[:, 2] - a[:, 2]
    r = np.stack([ca * dx + sa * dy, -sa * dx + ca * dy, np.arctan2(np.sin(dth), np.cos(dth))], axis=-1) - meas
    if not jacobian:
        return r, None, None
    E = a.shape[0]
    Ja = np.zeros((E, 3, 3))
    Jb = np.zeros((E, 3, 3))
    Ja[:, 0, 0], Ja[:, 0, 1], Ja[:, 0, 2] = -ca, -sa, -sa * dx + ca * dy
    Ja[:, 1, 0], Ja[:, 1, 1], Ja[:, 1, 2] = sa, -ca, -ca * dx - sa * dy
    Ja[:, 2, 2] = -1.0
    Jb[:, 0, 0], Jb[:, 0, 1] = ca, sa
    Jb[:, 1, 0], Jb[:, 1, 1] = -sa, ca
    Jb[:, 2, 2] = 1.0
    return r, Ja, Jb


class PoseGraphModel:
    def __init__(self, poses: np.ndarray, edges: np.ndarray, measurements: np.ndarray, loss=None,
                 linear_solver: str = "dense_qr"):
        self.start = np.asarray(poses, dtype=np.float64)
        self.edges = np.asarray(edges, dtype=np.int64)
        self.meas = np.asarray(measurements, dtype=np.float64)
        self.loss = loss
        self.linear_solver = linear_solver
        self.n = self.start.shape[0]

    def x0(self):
        return self.start.copy()

    def num_active_columns(self) -> int:
        return 3 * (self.n - 1)

    def state_norm(self, x) -> float:
        return float(np.sqrt(np.sum(x * x)))

    def apply(self, x, step):
        out = x.copy()
        out[1:] += step.reshape(-1, 3)
        return out

    def evaluate(self, x, jacobian: bool):
        a = x[self.edges[:, 0]]
        b = x[self.edges[:, 1]]
        res, Ja, Jb = relative_pose_residual(a, b, self.meas, jacobian)
        if not jacobian:
            cost_b, r, _ = robustify(self.loss, res, None)
            return DenseEval(float(np.sum(cost_b)), r.reshape(-1), None), None
        E = self.edges.shape[0]
        J = np.zeros((E, 3, 3 * self.n))
        rows = np.arange(E)
        for k in range(3):
            for j in range(3):
                J[rows, k, 3 * self.edges[:, 0] + j] += Ja[:, k, j]
                J[rows, k, 3 * self.edges[:, 1] + j] += Jb[:, k, j]
        J = J[:, :, 3:]
        cost_b, r, Jc = robustify(self.loss, res, J)
        ev = DenseEval(float(np.sum(cost_b)), r.reshape(-1), Jc.reshape(3 * E, -1))
        return ev, ev.J.T @ ev.r

    def column_sq(self, ev) -> np.ndarray:
        return np.sum(ev.J * ev.J, axis=0)

    def solve(self, ev, damping, opts):
        return dense_linear_solve(ev.J, -ev.r, damping, 0, self.linear_solver, opts)

    def model_decrease(self, ev, step) -> float:
        after = ev.r + ev.J @ step
        return max(0.0, float(0.5 * (ev.r @ ev.r) - 0.5 * (after @ after)))

    def gauss_newton_step(self, ev) -> np.ndarray:
        """linear.py:570-571: solve(J^T J, -J^T r)."""
        return _solve_square(ev.J.T @ ev.J, -(ev.J.T @ ev.r))

    def jmul(self, ev, v) -> np.ndarray:
        return ev.J @ v
