"""Seeded synthetic workloads (test infrastructure only).

These generators define the inputs of the five workloads the solver core is
measured on.  The reference ships no generators; the problem definitions follow
its examples:

* bundle adjustment uses the Snavely/BAL camera model of
  ``examples/simple_bundle_adjuster.py:73-79`` (9 camera parameters: angle-axis
  rotation, translation, focal length f, radial k1 and k2; the camera looks down
  -z, so points in front have p_z < 0), with cameras added before points and
  points placed in ordering group 0 (``simple_bundle_adjuster.py:82-112``);
* curve fitting uses one residual block per sample ``y - exp(m x + c)`` with a
  Cauchy loss, as in ``examples/robust_curve_fitting.py:80-97``;
* 2D pose graphs use the SE2 relative-pose residual of
  ``examples/pose_graph_2d.py:6-15`` with the first pose held constant
  (``pose_graph_2d.py:34``).

Everything is numpy float64 driven by ``numpy.random.default_rng(seed)`` so the
same seed gives the same problem on every machine.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


# ----------------------------------------------------------------------------
# bundle adjustment
# ----------------------------------------------------------------------------


@dataclass
class BAProblemData:
    """A bundle-adjustment problem in observation-major arrays.

    Observations are grouped by point (point-major, cameras ascending within a
    point); this is also the residual-block insertion order.
    """

    cameras: np.ndarray          # [C, 9] float64 (angle-axis) or [C, 10] (quaternion models), initial estimate
    points: np.ndarray           # [P, 3] float64, initial estimate
    obs_camera: np.ndarray       # [N] int32
    obs_point: np.ndarray        # [N] int32 (non-decreasing)
    observations: np.ndarray     # [N, 2] float64 pixels
    true_cameras: np.ndarray | None = None
    true_points: np.ndarray | None = None
    outlier_mask: np.ndarray | None = None
    loss: tuple = ("trivial",)   # ("trivial",) | ("huber", a) | ("cauchy", a) ...
    meta: dict = field(default_factory=dict)
    camera_model: str = "angle_axis"  # | "quaternion" (q = w, x, y, z) | "eigen_quaternion" (x, y, z, w)

    @property
    def num_cameras(self) -> int:
        return int(self.cameras.shape[0])

    @property
    def num_points(self) -> int:
        return int(self.points.shape[0])

    @property
    def num_observations(self) -> int:
        return int(self.observations.shape[0])


def _circle_cameras(num_cameras: int, radius: float, focal: float) -> np.ndarray:
    """Cameras on a circle in the x-z plane, each looking at the origin.

    Camera i sits at angle phi_i = 2*pi*i/C at centre 10*(sin phi, 0, cos phi).
    The world-to-camera rotation is a rotation about y by -phi, whose
    angle-axis is (0, -phi, 0) with phi wrapped into (-pi, pi].  The camera z
    axis points away from the origin, so the origin has p_z = -radius.
    """
    cams = np.zeros((num_cameras, 9), dtype=np.float64)
    for i in range(num_cameras):
        phi = 2.0 * np.pi * i / num_cameras
        wrapped = (phi + np.pi) % (2.0 * np.pi) - np.pi
        if wrapped <= -np.pi:
            wrapped += 2.0 * np.pi
        centre = radius * np.array([np.sin(phi), 0.0, np.cos(phi)])
        rot = np.array(
            [
                [np.cos(phi), 0.0, -np.sin(phi)],
                [0.0, 1.0, 0.0],
                [np.sin(phi), 0.0, np.cos(phi)],
            ]
        )
        cams[i, 0:3] = (0.0, -wrapped, 0.0)
        cams[i, 3:6] = -rot @ centre
        cams[i, 6] = focal
    return cams


def make_bundle_adjustment(
    *,
    num_cameras: int,
    num_points: int,
    obs_per_point: tuple[int, int],
    seed: int,
    pixel_noise: float = 0.5,
    rotation_noise: float = 0.01,
    translation_noise: float = 0.05,
    point_noise: float = 0.05,
    outlier_fraction: float = 0.0,
    outlier_range: float = 1000.0,
    loss: tuple = ("trivial",),
    point_box: float = 2.0,
    radius: float = 10.0,
    focal: float = 500.0,
) -> BAProblemData:
    """Synthetic BA problem: circle cameras, uniform points, noisy pixels.

    ``obs_per_point`` is the inclusive range of distinct cameras observing each
    point.  Pixel observations are the true projections plus N(0, noise^2); a
    fraction ``outlier_fraction`` of them is then replaced by U[0, range]^2.
    """
    from .ba import project  # local import keeps generators importable alone

    rng = np.random.default_rng(seed)
    true_cams = _circle_cameras(num_cameras, radius, focal)
    true_pts = rng.uniform(-point_box, point_box, size=(num_points, 3))
    lo, hi = obs_per_point
    if lo == hi:
        counts = np.full(num_points, lo, dtype=np.int64)
    else:
        counts = rng.integers(lo, hi + 1, size=num_points)
    counts = np.minimum(counts, num_cameras)
    obs_cam = []
    for p in range(num_points):
        chosen = np.sort(rng.choice(num_cameras, size=int(counts[p]), replace=False))
        obs_cam.append(chosen)
    obs_camera = np.concatenate(obs_cam).astype(np.int32)
    obs_point = np.repeat(np.arange(num_points, dtype=np.int32), counts)
    clean = project(true_cams[obs_camera], true_pts[obs_point])
    observations = clean + rng.normal(0.0, pixel_noise, size=clean.shape)
    outlier_mask = np.zeros(observations.shape[0], dtype=bool)
    if outlier_fraction > 0.0:
        n_out = int(round(outlier_fraction * observations.shape[0]))
        idx = rng.choice(observations.shape[0], size=n_out, replace=False)
        observations[idx] = rng.uniform(0.0, outlier_range, size=(n_out, 2))
        outlier_mask[idx] = True
    cams = true_cams.copy()
    cams[:, 0:3] += rng.normal(0.0, rotation_noise, size=(num_cameras, 3))
    cams[:, 3:6] += rng.normal(0.0, translation_noise, size=(num_cameras, 3))
    pts = true_pts + rng.normal(0.0, point_noise, size=true_pts.shape)
    return BAProblemData(
        cameras=cams,
        points=pts,
        obs_camera=obs_camera,
        obs_point=obs_point,
        observations=observations,
        true_cameras=true_cams,
        true_points=true_pts,
        outlier_mask=outlier_mask,
        loss=loss,
        meta=dict(seed=seed, obs_per_point=obs_per_point, outlier_fraction=outlier_fraction),
    )


def with_quaternion_cameras(data: BAProblemData, order: str = "ceres", scale=1.0) -> BAProblemData:
    """The same problem with each camera's rotation as a quaternion [q(4), t(3), f, k1, k2].

    q = angle_axis_to_quaternion(aa) (rotation.py:104-110) times ``scale`` (a
    scalar or one factor per camera): quaternion_rotate_point normalises, so
    the projections are unchanged while a non-unit start exercises the 1/|q|
    of the rotation Jacobian.  ``order`` "eigen" stores q as (x, y, z, w).
    """
    from .ba import angle_axis_to_quaternion

    def convert(cams):
        if cams is None:
            return None
        q = angle_axis_to_quaternion(cams[:, 0:3]) * np.reshape(np.asarray(scale, dtype=np.float64), (-1, 1))
        if order == "eigen":
            q = q[:, [1, 2, 3, 0]]
        return np.concatenate([q, cams[:, 3:9]], axis=1)

    out = BAProblemData(cameras=convert(data.cameras), points=data.points.copy(), obs_camera=data.obs_camera,
                        obs_point=data.obs_point, observations=data.observations,
                        true_cameras=convert(data.true_cameras), true_points=data.true_points,
                        outlier_mask=data.outlier_mask, loss=data.loss, meta=dict(data.meta),
                        camera_model="quaternion" if order == "ceres" else "eigen_quaternion")
    return out


def ba_config(name: str, seed: int | None = None) -> BAProblemData:
    """Named BA workloads.

    * ``ba16``      16 cameras, 2,000 points, 4 obs/point (8,000 obs), trivial loss
    * ``ba200``     200 cameras, 100,000 points, 3..10 obs/point (~650k obs)
    * ``ba200_huber`` ba200 with 5% uniform outliers and Huber(1.0)
    * ``ba16_huber``  ba16 with 5% uniform outliers and Huber(1.0)
    * ``ba4``       4 cameras, 60 points, 3 obs/point: a reduced case the dense
                    reference algebra finishes in well under a second
    * ``ba6_huber`` 6 cameras, 120 points, 2..4 obs/point, 5% outliers, Huber(1.0)
    """
    if name == "ba16":
        return make_bundle_adjustment(num_cameras=16, num_points=2000, obs_per_point=(4, 4), seed=0 if seed is None else seed)
    if name == "ba16_huber":
        return make_bundle_adjustment(
            num_cameras=16, num_points=2000, obs_per_point=(4, 4), seed=2 if seed is None else seed,
            outlier_fraction=0.05, loss=("huber", 1.0),
        )
    if name == "ba200":
        return make_bundle_adjustment(num_cameras=200, num_points=100_000, obs_per_point=(3, 10), seed=1 if seed is None else seed)
    if name == "ba200_huber":
        return make_bundle_adjustment(
            num_cameras=200, num_points=100_000, obs_per_point=(3, 10), seed=2 if seed is None else seed,
            outlier_fraction=0.05, loss=("huber", 1.0),
        )
    if name == "ba4":
        return make_bundle_adjustment(num_cameras=4, num_points=60, obs_per_point=(3, 3), seed=10 if seed is None else seed)
    if name == "ba6_huber":
        return make_bundle_adjustment(
            num_cameras=6, num_points=120, obs_per_point=(2, 4), seed=11 if seed is None else seed,
            outlier_fraction=0.05, loss=("huber", 1.0),
        )
    raise KeyError(name)


# ----------------------------------------------------------------------------
# batched curve fits  y = exp(m x + c)
# ----------------------------------------------------------------------------


@dataclass
class CurveBatchData:
    offsets: np.ndarray   # [B+1] int64 CSR offsets into x/y
    x: np.ndarray         # [total] float64
    y: np.ndarray         # [total] float64
    true_m: np.ndarray    # [B]
    true_c: np.ndarray    # [B]
    loss: tuple = ("cauchy", 0.5)

    @property
    def num_problems(self) -> int:
        return int(self.offsets.shape[0] - 1)


def make_curve_batch(
    *,
    num_problems: int,
    seed: int,
    n_range: tuple[int, int] = (50, 2000),
    x_range: tuple[float, float] = (0.0, 5.0),
    m_range: tuple[float, float] = (0.1, 0.4),
    c_range: tuple[float, float] = (0.0, 0.3),
    noise: float = 0.2,
    loss: tuple = ("cauchy", 0.5),
) -> CurveBatchData:
    rng = np.random.default_rng(seed)
    counts = rng.integers(n_range[0], n_range[1] + 1, size=num_problems)
    m = rng.uniform(*m_range, size=num_problems)
    c = rng.uniform(*c_range, size=num_problems)
    offsets = np.zeros(num_problems + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(counts)
    total = int(offsets[-1])
    x = rng.uniform(*x_range, size=total)
    owner = np.repeat(np.arange(num_problems), counts)
    y = np.exp(m[owner] * x + c[owner]) + rng.normal(0.0, noise, size=total)
    return CurveBatchData(offsets=offsets, x=x, y=y, true_m=m, true_c=c, loss=loss)


# ----------------------------------------------------------------------------
# 2D pose graphs
# ----------------------------------------------------------------------------


@dataclass
class PoseGraphData:
    poses: np.ndarray        # [N, 3] initial estimate (x, y, theta)
    edges: np.ndarray        # [E, 2] int32 (a, b)
    measurements: np.ndarray  # [E, 3] relative pose of b in a's frame
    true_poses: np.ndarray


def _relative_pose(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    ca, sa = np.cos(a[..., 2]), np.sin(a[..., 2])
    dx = b[..., 0] - a[..., 0]
    dy = b[..., 1] - a[..., 1]
    dth = np.arctan2(np.sin(b[..., 2] - a[..., 2]), np.cos(b[..., 2] - a[..., 2]))
    return np.stack([ca * dx + sa * dy, -sa * dx + ca * dy, dth], axis=-1)


def _compose(a: np.ndarray, rel: np.ndarray) -> np.ndarray:
    ca, sa = np.cos(a[2]), np.sin(a[2])
    return np.array([a[0] + ca * rel[0] - sa * rel[1], a[1] + sa * rel[0] + ca * rel[1], a[2] + rel[2]])


def make_pose_graph(
    *,
    num_nodes: int,
    seed: int,
    loop_fraction: float = 0.10,
    sigma_xy: float = 0.05,
    sigma_theta: float = 0.01,
    radius: float | None = None,
) -> PoseGraphData:
    """Nodes on a noisy loop; odometry edges plus random loop closures.

    The true trajectory walks once round a circle.  Odometry connects i -> i+1;
    ``loop_fraction * N`` extra edges join random pairs at least two apart.
    Every measurement is the true relative pose plus N(0, sigma^2) noise, and
    the initial estimate is dead reckoning from the noisy odometry.
    """
    rng = np.random.default_rng(seed)
    if radius is None:
        radius = num_nodes / (2.0 * np.pi)  # ~1 unit between consecutive poses
    ang = 2.0 * np.pi * np.arange(num_nodes) / num_nodes
    true = np.stack([radius * np.cos(ang), radius * np.sin(ang), np.arctan2(np.cos(ang), -np.sin(ang))], axis=1)
    true[:, :2] += rng.normal(0.0, 0.1, size=(num_nodes, 2))
    edges = [(i, i + 1) for i in range(num_nodes - 1)]
    n_loop = int(round(loop_fraction * num_nodes))
    seen = set(edges)
    while len(edges) < (num_nodes - 1) + n_loop:
        a, b = (int(v) for v in rng.integers(0, num_nodes, size=2))
        if abs(a - b) < 2:
            continue
        a, b = min(a, b), max(a, b)
        if (a, b) in seen:
            continue
        seen.add((a, b))
        edges.append((a, b))
    edges_arr = np.asarray(edges, dtype=np.int32)
    meas = _relative_pose(true[edges_arr[:, 0]], true[edges_arr[:, 1]])
    meas[:, :2] += rng.normal(0.0, sigma_xy, size=(meas.shape[0], 2))
    meas[:, 2] += rng.normal(0.0, sigma_theta, size=meas.shape[0])
    init = np.zeros_like(true)
    init[0] = true[0]
    for i in range(num_nodes - 1):
        init[i + 1] = _compose(init[i], meas[i])
    return PoseGraphData(poses=init, edges=edges_arr, measurements=meas, true_poses=true)
