"""Bundle Adjustment in the Large (BAL) files for the device path.

``BALProblem`` keeps the reference's reader (``examples/simple_bundle_adjuster.py:9-70``):
the header ``num_cameras num_points num_observations``, one line
``camera_index point_index x y`` per observation, then the 9 parameters of
every camera (angle-axis rotation, translation, focal length, k1, k2) and the
3 coordinates of every point; any value past the last point is an error.  The
text is tokenised once and converted with numpy (a million observations parse
in about a second), not value by value.

``build_problem`` is the reference's ``build_problem``
(``simple_bundle_adjuster.py:82-112``) on the bulk bundle-adjustment API: one
Snavely residual per observation, points in elimination group 0, cameras
optionally constant.  The camera and point tensors are the problem's parameter
blocks (row by row) and receive the solution in place.
"""
from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path

import numpy as np
import torch

from .problem import Problem


@dataclass
class BALProblem:
    camera_indices: list[int]
    point_indices: list[int]
    observations: torch.Tensor   # [N, 2]
    cameras: torch.Tensor        # [C, 9]
    points: torch.Tensor         # [P, 3]

    @property
    def num_cameras(self) -> int:
        return int(self.cameras.shape[0])

    @property
    def num_points(self) -> int:
        return int(self.points.shape[0])

    @property
    def num_observations(self) -> int:
        return int(self.observations.shape[0])

    @classmethod
    def from_file(cls, path: str | Path, *, dtype: torch.dtype = torch.float64) -> "BALProblem":
        return cls.from_text(Path(path).read_text(), dtype=dtype)

    @classmethod
    def from_text(cls, text: str, *, dtype: torch.dtype = torch.float64) -> "BALProblem":
        tokens = text.split()
        if len(tokens) < 3:
            raise ValueError("BAL input has no header")
        try:
            num_cameras, num_points, num_observations = (int(t) for t in tokens[:3])
        except ValueError as exc:
            raise ValueError("BAL header must hold three integers") from exc
        if min(num_cameras, num_points, num_observations) < 0:
            raise ValueError("BAL header counts must be non-negative")
        need = 3 + 4 * num_observations + 9 * num_cameras + 3 * num_points
        if len(tokens) < need:
            raise ValueError(f"BAL input ends early: {len(tokens)} values, {need} expected")
        if len(tokens) > need:
            raise ValueError("BAL input contains trailing values")
        obs = np.array(tokens[3: 3 + 4 * num_observations], dtype=np.float64).reshape(num_observations, 4)
        idx = obs[:, :2]
        if not np.all(idx == np.round(idx)):
            raise ValueError("BAL camera and point indices must be integers")
        cam_idx = idx[:, 0].astype(np.int64)
        pt_idx = idx[:, 1].astype(np.int64)
        if num_observations and (cam_idx.min() < 0 or cam_idx.max() >= num_cameras or pt_idx.min() < 0
                                 or pt_idx.max() >= num_points):
            raise ValueError("BAL observation refers to a missing camera or point")
        values = np.array(tokens[3 + 4 * num_observations:], dtype=np.float64)
        cameras = values[: 9 * num_cameras].reshape(num_cameras, 9)
        points = values[9 * num_cameras:].reshape(num_points, 3)
        return cls(
            camera_indices=cam_idx.tolist(),
            point_indices=pt_idx.tolist(),
            observations=torch.tensor(obs[:, 2:], dtype=dtype),
            cameras=torch.tensor(cameras, dtype=dtype),
            points=torch.tensor(points, dtype=dtype),
        )

    def to_text(self) -> str:
        """BAL text of this problem (17 significant digits, so from_text round-trips exactly)."""
        lines = [f"{self.num_cameras} {self.num_points} {self.num_observations}"]
        obs = self.observations.detach().cpu().numpy()
        lines.extend(f"{c} {p} {x:.17g} {y:.17g}" for c, p, (x, y) in
                     zip(self.camera_indices, self.point_indices, obs.tolist()))
        lines.extend(f"{v:.17g}" for v in self.cameras.detach().cpu().reshape(-1).tolist())
        lines.extend(f"{v:.17g}" for v in self.points.detach().cpu().reshape(-1).tolist())
        return "\n".join(lines) + "\n"


def build_problem(bal: BALProblem, *, fix_cameras: bool = False, loss=None,
                  device: str | torch.device | None = None) -> tuple[Problem, torch.Tensor, torch.Tensor]:
    """A Problem with one Snavely reprojection residual per observation.

    ``device`` moves the camera and point tensors (default: where they are);
    the returned tensors are the ones the solve updates in place.
    """
    cams = bal.cameras if device is None else bal.cameras.to(device)
    pts = bal.points if device is None else bal.points.to(device)
    problem = Problem()
    problem.add_bundle_adjustment_observations(cams, pts, np.asarray(bal.camera_indices),
                                               np.asarray(bal.point_indices),
                                               bal.observations.detach().cpu().numpy(), loss=loss)
    if fix_cameras:
        for i in range(bal.num_cameras):
            problem.SetParameterBlockConstant(problem.parameter_blocks[i])
    return problem, cams, pts
