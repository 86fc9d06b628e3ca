"""Rotation helpers on the BA path (``src/ceres_torch/rotation.py:9-31, 68-110, 211-222``).

Torch versions for API users and tests; the device kernels implement the same
angle-axis -> unit quaternion -> rotate formula, and the quaternion rotation of
the quaternion cameras, in HIP.  Quaternions are (w, x, y, z) ("ceres" order)
unless stated; the "eigen" order is (x, y, z, w).
"""
from __future__ import annotations

import torch


def angle_axis_to_quaternion(angle_axis: torch.Tensor) -> torch.Tensor:
    theta = torch.linalg.norm(angle_axis, dim=-1, keepdim=True)
    half = 0.5 * theta
    small = theta <= torch.finfo(theta.dtype).eps
    k = torch.where(small, 0.5 - theta * theta / 48.0, torch.sin(half) / torch.where(small, torch.ones_like(theta), theta))
    q = torch.cat([torch.cos(half), k * angle_axis], dim=-1)
    return q / torch.linalg.norm(q, dim=-1, keepdim=True).clamp_min(torch.finfo(q.dtype).tiny)


def unit_quaternion_rotate_point(q: torch.Tensor, point: torch.Tensor) -> torch.Tensor:
    u = q[..., 1:]
    uv = 2.0 * torch.linalg.cross(u, point, dim=-1)
    return point + q[..., :1] * uv + torch.linalg.cross(u, uv, dim=-1)


def angle_axis_rotate_point(angle_axis: torch.Tensor, point: torch.Tensor) -> torch.Tensor:
    return unit_quaternion_rotate_point(angle_axis_to_quaternion(angle_axis), point)


def normalize_quaternion(q: torch.Tensor) -> torch.Tensor:
    """q / |q| (rotation.py:68-69)."""
    return q / torch.linalg.norm(q, dim=-1, keepdim=True).clamp_min(torch.finfo(q.dtype).tiny)


def quaternion_product(z: torch.Tensor, w: torch.Tensor) -> torch.Tensor:
    """Hamilton product z w (rotation.py:72-83)."""
    z0, z1, z2, z3 = z.unbind(-1)
    w0, w1, w2, w3 = w.unbind(-1)
    return torch.stack([z0 * w0 - z1 * w1 - z2 * w2 - z3 * w3,
                        z0 * w1 + z1 * w0 + z2 * w3 - z3 * w2,
                        z0 * w2 - z1 * w3 + z2 * w0 + z3 * w1,
                        z0 * w3 + z1 * w2 - z2 * w1 + z3 * w0], dim=-1)


def quaternion_conjugate(q: torch.Tensor) -> torch.Tensor:
    """(w, -x, -y, -z) (rotation.py:86-87)."""
    return torch.cat([q[..., :1], -q[..., 1:]], dim=-1)


def quaternion_rotate_point(q: torch.Tensor, point: torch.Tensor) -> torch.Tensor:
    """Rotation by q / |q| (rotation.py:217-218)."""
    return unit_quaternion_rotate_point(normalize_quaternion(q), point)


def convert_eigen_to_ceres_quaternion(q: torch.Tensor) -> torch.Tensor:
    """(x, y, z, w) -> (w, x, y, z)."""
    return torch.cat([q[..., 3:4], q[..., 0:3]], dim=-1)


def convert_ceres_to_eigen_quaternion(q: torch.Tensor) -> torch.Tensor:
    """(w, x, y, z) -> (x, y, z, w)."""
    return torch.cat([q[..., 1:4], q[..., 0:1]], dim=-1)
