"""Rotation helpers on the BA path (``src/ceres_torch/rotation.py:104-110, 211-222``).

Torch versions for API users and tests; the device kernels implement the same
angle-axis -> unit quaternion -> rotate formula in HIP.
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
