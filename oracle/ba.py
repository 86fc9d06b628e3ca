"""Bundle adjustment in float64 numpy: residuals, analytic Jacobians, Schur solve.

Test infrastructure only.  Restates, in block-sparse form, what the reference
computes densely for a Snavely/BAL bundle-adjustment problem:

* rotation: ``angle_axis_to_quaternion`` + ``unit_quaternion_rotate_point``
  (``src/ceres_torch/rotation.py:104-110, 211-222``);
* residual: ``examples/simple_bundle_adjuster.py:73-79``;
* Jacobian: the reference differentiates the residual with autograd
  (``costs.py:31-44``); here it is the closed-form derivative of the same
  function (Rodrigues form of the rotation derivative);
* loss correction: ``losses.py:158-193`` (see ``oracle/losses.py``);
* damping: ``linear.py:539-551``;
* Schur elimination: ``schur_solve_dense`` (``linear.py:667-699``) with the
  eliminated block ``Haa`` inverted point by point (it is block diagonal);
* explicit-S PCG with the SCHUR_JACOBI preconditioner: ``iterative_schur_solve``
  (``linear.py:702-776``), ``_build_matrix_preconditioner``
  (``linear.py:1156-1182``) and ``_preconditioned_conjugate_gradient``
  (``linear.py:1216-1256``), one diagonal block per parameter block
  (``solver.py:153``, ``linear.py:1269-1289``);
* quaternion cameras [q(4), t(3), f, k1, k2]: the rotation is
  ``quaternion_rotate_point`` (``rotation.py:217-218``), q moves by
  ``QuaternionManifold.plus`` (``manifolds.py:130-137``; Eigen order
  ``manifolds.py:161-171``) and its tangent Jacobian is
  ``J_ambient @ plus_jacobian`` (``problem.py:505-514``), here in closed form;
  the rotation and the other six values are separate parameter blocks, so the
  preconditioner has a 3x3 and a 6x6 block per camera.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .losses import robustify

_EPS = np.finfo(np.float64).eps
_TINY = np.finfo(np.float64).tiny


# ----------------------------------------------------------------------------
# rotation and projection
# ----------------------------------------------------------------------------


def angle_axis_to_quaternion(aa: np.ndarray) -> np.ndarray:
    """rotation.py:104-110 (normalised [w, x, y, z])."""
    theta = np.sqrt(np.sum(aa * aa, axis=-1, keepdims=True))
    half = 0.5 * theta
    small = theta <= _EPS
    safe = np.where(small, 1.0, theta)
    k = np.where(small, 0.5 - theta * theta / 48.0, np.sin(half) / safe)
    q = np.concatenate([np.cos(half), k * aa], axis=-1)
    return q / np.maximum(np.sqrt(np.sum(q * q, axis=-1, keepdims=True)), _TINY)


def angle_axis_rotate_point(aa: np.ndarray, x: np.ndarray) -> np.ndarray:
    """rotation.py:211-222: p + 2w(u x p) + u x (2 u x p)."""
    q = angle_axis_to_quaternion(aa)
    u = q[..., 1:]
    uv = 2.0 * np.cross(u, x)
    return x + q[..., :1] * uv + np.cross(u, uv)


def project(cams: np.ndarray, pts: np.ndarray) -> np.ndarray:
    """Snavely camera projection (simple_bundle_adjuster.py:73-79)."""
    p = angle_axis_rotate_point(cams[..., 0:3], pts) + cams[..., 3:6]
    xp = -p[..., 0] / p[..., 2]
    yp = -p[..., 1] / p[..., 2]
    r2 = xp * xp + yp * yp
    d = 1.0 + r2 * (cams[..., 7] + cams[..., 8] * r2)
    return cams[..., 6:7] * d[..., None] * np.stack([xp, yp], axis=-1)


def _skew(v: np.ndarray) -> np.ndarray:
    z = np.zeros(v.shape[:-1])
    return np.stack(
        [
            np.stack([z, -v[..., 2], v[..., 1]], -1),
            np.stack([v[..., 2], z, -v[..., 0]], -1),
            np.stack([-v[..., 1], v[..., 0], z], -1),
        ],
        -2,
    )


def rotation_point_jacobians(aa: np.ndarray, x: np.ndarray):
    """d(R(aa) x)/d aa and R(aa), from Rodrigues' formula.

    R x = x + A (w x x) + B (w x (w x x)) with A = sin t / t, B = (1 - cos t)/t^2.
    d/dw = -A [x]_x + C1 (w x x) w^T + B((w.x) I + w x^T - 2 x w^T)
           + C2 (w x (w x x)) w^T,
    C1 = (t cos t - sin t)/t^3, C2 = (t sin t - 2(1 - cos t))/t^4, with Taylor
    series below t = 1e-4.
    """
    t2 = np.sum(aa * aa, axis=-1)
    t = np.sqrt(t2)
    small = t < 1e-4
    ts = np.where(small, 1.0, t)
    A = np.where(small, 1.0 - t2 / 6.0 + t2 * t2 / 120.0, np.sin(ts) / ts)
    B = np.where(small, 0.5 - t2 / 24.0 + t2 * t2 / 720.0, (1.0 - np.cos(ts)) / (ts * ts))
    C1 = np.where(small, -1.0 / 3.0 + t2 / 30.0 - t2 * t2 / 840.0, (ts * np.cos(ts) - np.sin(ts)) / ts**3)
    C2 = np.where(
        small,
        -1.0 / 12.0 + t2 / 180.0 - t2 * t2 / 6720.0,
        (ts * np.sin(ts) - 2.0 * (1.0 - np.cos(ts))) / ts**4,
    )
    wx = np.cross(aa, x)
    wwx = np.cross(aa, wx)
    wdotx = np.sum(aa * x, axis=-1)
    eye = np.eye(3)
    dR = (
        -A[..., None, None] * _skew(x)
        + C1[..., None, None] * wx[..., :, None] * aa[..., None, :]
        + B[..., None, None]
        * (wdotx[..., None, None] * eye + aa[..., :, None] * x[..., None, :] - 2.0 * x[..., :, None] * aa[..., None, :])
        + C2[..., None, None] * wwx[..., :, None] * aa[..., None, :]
    )
    W = _skew(aa)
    R = eye + A[..., None, None] * W + B[..., None, None] * (W @ W)
    return dR, R


def snavely_residual_jacobian(cams: np.ndarray, pts: np.ndarray, obs: np.ndarray):
    """Residual (N,2), camera Jacobian (N,2,9), point Jacobian (N,2,3)."""
    aa = cams[:, 0:3]
    p = angle_axis_rotate_point(aa, pts) + cams[:, 3:6]
    dR, R = rotation_point_jacobians(aa, pts)
    f, k1, k2 = cams[:, 6], cams[:, 7], cams[:, 8]
    xp = -p[:, 0] / p[:, 2]
    yp = -p[:, 1] / p[:, 2]
    r2 = xp * xp + yp * yp
    d = 1.0 + r2 * (k1 + k2 * r2)
    res = np.stack([f * d * xp, f * d * yp], axis=-1) - obs
    # d(xp, yp)/dp
    inv_z = 1.0 / p[:, 2]
    dxy_dp = np.zeros((p.shape[0], 2, 3))
    dxy_dp[:, 0, 0] = -inv_z
    dxy_dp[:, 0, 2] = -xp * inv_z
    dxy_dp[:, 1, 1] = -inv_z
    dxy_dp[:, 1, 2] = -yp * inv_z
    dd = k1 + 2.0 * k2 * r2  # dd/dr2
    duv_dxy = np.empty((p.shape[0], 2, 2))
    duv_dxy[:, 0, 0] = f * (d + 2.0 * xp * xp * dd)
    duv_dxy[:, 0, 1] = f * 2.0 * xp * yp * dd
    duv_dxy[:, 1, 0] = duv_dxy[:, 0, 1]
    duv_dxy[:, 1, 1] = f * (d + 2.0 * yp * yp * dd)
    duv_dp = duv_dxy @ dxy_dp
    Jc = np.empty((p.shape[0], 2, 9))
    Jc[:, :, 0:3] = duv_dp @ dR
    Jc[:, :, 3:6] = duv_dp
    Jc[:, 0, 6] = d * xp
    Jc[:, 1, 6] = d * yp
    Jc[:, 0, 7] = f * xp * r2
    Jc[:, 1, 7] = f * yp * r2
    Jc[:, 0, 8] = f * xp * r2 * r2
    Jc[:, 1, 8] = f * yp * r2 * r2
    Jp = duv_dp @ R
    return res, Jc, Jp


# ----------------------------------------------------------------------------
# quaternion cameras
# ----------------------------------------------------------------------------


def normalize_quaternion(q: np.ndarray) -> np.ndarray:
    """rotation.py:68-69."""
    return q / np.maximum(np.sqrt(np.sum(q * q, axis=-1, keepdims=True)), _TINY)


def quaternion_product(z: np.ndarray, w: np.ndarray) -> np.ndarray:
    """rotation.py:72-83."""
    z0, z1, z2, z3 = np.moveaxis(z, -1, 0)
    w0, w1, w2, w3 = np.moveaxis(w, -1, 0)
    return np.stack([z0 * w0 - z1 * w1 - z2 * w2 - z3 * w3,
                     z0 * w1 + z1 * w0 + z2 * w3 - z3 * w2,
                     z0 * w2 - z1 * w3 + z2 * w0 + z3 * w1,
                     z0 * w3 + z1 * w2 - z2 * w1 + z3 * w0], axis=-1)


def quaternion_rotate_point(q: np.ndarray, x: np.ndarray) -> np.ndarray:
    """rotation.py:211-218: unit_quaternion_rotate_point(q / |q|, x)."""
    qn = normalize_quaternion(q)
    u = qn[..., 1:]
    uv = 2.0 * np.cross(u, x)
    return x + qn[..., :1] * uv + np.cross(u, uv)


def quaternion_plus(q: np.ndarray, delta: np.ndarray) -> np.ndarray:
    """QuaternionManifold.plus (manifolds.py:130-137), row-wise: normalize(q_delta * normalize(q))."""
    x = normalize_quaternion(q)
    nd = np.sqrt(np.sum(delta * delta, axis=-1, keepdims=True))
    big = nd > _EPS
    scale = np.where(big, np.sin(nd) / np.where(big, nd, 1.0), 1.0 - nd * nd / 6.0)
    q_delta = np.concatenate([np.cos(nd), scale * delta], axis=-1)
    return normalize_quaternion(quaternion_product(q_delta, x))


def quaternion_plus_jacobian(q: np.ndarray) -> np.ndarray:
    """QuaternionManifold.plus_jacobian (manifolds.py:148-158): [4, 3] at normalize(q)."""
    w, a, b, c = normalize_quaternion(q)
    return np.array([[-a, -b, -c], [w, c, -b], [-c, w, a], [b, -a, w]])


def ceres_order(cams: np.ndarray, camera_model: str) -> np.ndarray:
    """Quaternion cameras with q in (w, x, y, z) order."""
    if camera_model == "eigen_quaternion":
        return np.concatenate([cams[..., 3:4], cams[..., 0:3], cams[..., 4:]], axis=-1)
    return cams


def storage_order(cams: np.ndarray, camera_model: str) -> np.ndarray:
    """Inverse of ceres_order."""
    if camera_model == "eigen_quaternion":
        return np.concatenate([cams[..., 1:4], cams[..., 0:1], cams[..., 4:]], axis=-1)
    return cams


def quaternion_project(cams: np.ndarray, pts: np.ndarray) -> np.ndarray:
    """The Snavely projection with the rotation quaternion_rotate_point(q, X); cams in Ceres order."""
    p = quaternion_rotate_point(cams[..., 0:4], pts) + cams[..., 4:7]
    xp = -p[..., 0] / p[..., 2]
    yp = -p[..., 1] / p[..., 2]
    r2 = xp * xp + yp * yp
    d = 1.0 + r2 * (cams[..., 8] + cams[..., 9] * r2)
    return cams[..., 7:8] * d[..., None] * np.stack([xp, yp], axis=-1)


def quaternion_residual_jacobian(cams: np.ndarray, pts: np.ndarray, obs: np.ndarray):
    """Residual (N,2), tangent camera Jacobian (N,2,9), point Jacobian (N,2,3); cams in Ceres order.

    Rotation columns: d(R X)/d delta = -2 [R X]_x / |q| (q_delta rotates by 2|delta|; the
    normalisation of quaternion_rotate_point divides by |q|).  Point columns: the derivative of
    p + 2w(u x p) + u x (2u x p) in p at the unit quaternion, I + 2w[u]_x + 2(u u^T - |u|^2 I).
    """
    qn = normalize_quaternion(cams[:, 0:4])
    inv = 1.0 / np.maximum(np.sqrt(np.sum(cams[:, 0:4] ** 2, axis=-1)), _TINY)
    pr = quaternion_rotate_point(cams[:, 0:4], pts)
    p = pr + cams[:, 4:7]
    f, k1, k2 = cams[:, 7], cams[:, 8], cams[:, 9]
    xp = -p[:, 0] / p[:, 2]
    yp = -p[:, 1] / p[:, 2]
    r2 = xp * xp + yp * yp
    d = 1.0 + r2 * (k1 + k2 * r2)
    res = np.stack([f * d * xp, f * d * yp], axis=-1) - obs
    inv_z = 1.0 / p[:, 2]
    dxy_dp = np.zeros((p.shape[0], 2, 3))
    dxy_dp[:, 0, 0] = -inv_z
    dxy_dp[:, 0, 2] = -xp * inv_z
    dxy_dp[:, 1, 1] = -inv_z
    dxy_dp[:, 1, 2] = -yp * inv_z
    dd = k1 + 2.0 * k2 * r2
    duv_dxy = np.empty((p.shape[0], 2, 2))
    duv_dxy[:, 0, 0] = f * (d + 2.0 * xp * xp * dd)
    duv_dxy[:, 0, 1] = f * 2.0 * xp * yp * dd
    duv_dxy[:, 1, 0] = duv_dxy[:, 0, 1]
    duv_dxy[:, 1, 1] = f * (d + 2.0 * yp * yp * dd)
    duv_dp = duv_dxy @ dxy_dp
    w, u = qn[:, 0], qn[:, 1:]
    eye = np.eye(3)
    M = (eye + 2.0 * w[:, None, None] * _skew(u) + 2.0 * (u[:, :, None] * u[:, None, :]
                                                          - np.sum(u * u, axis=-1)[:, None, None] * eye))
    Jc = np.empty((p.shape[0], 2, 9))
    Jc[:, :, 0:3] = duv_dp @ (-2.0 * inv[:, None, None] * _skew(pr))
    Jc[:, :, 3:6] = duv_dp
    Jc[:, 0, 6] = d * xp
    Jc[:, 1, 6] = d * yp
    Jc[:, 0, 7] = f * xp * r2
    Jc[:, 1, 7] = f * yp * r2
    Jc[:, 0, 8] = f * xp * r2 * r2
    Jc[:, 1, 8] = f * yp * r2 * r2
    Jp = duv_dp @ M
    return res, Jc, Jp


def camera_plus(cams: np.ndarray, dc: np.ndarray, camera_model: str, active=None) -> np.ndarray:
    """Plus on every camera: Euclidean for angle-axis cameras; QuaternionManifold.plus on the
    rotation and Euclidean on t, f, k1, k2 for quaternion cameras.  ``active`` (per camera, default
    all): apply_delta (problem.py:462-474) leaves constant blocks untouched, so a constant
    quaternion keeps its stored (possibly non-unit) values."""
    if camera_model == "angle_axis":
        return cams + dc
    c = ceres_order(cams, camera_model)
    q = quaternion_plus(c[:, 0:4], dc[:, 0:3])
    if active is not None:
        q = np.where(np.asarray(active, bool)[:, None], q, c[:, 0:4])
    out = np.concatenate([q, c[:, 4:] + dc[:, 3:]], axis=1)
    return storage_order(out, camera_model)


# ----------------------------------------------------------------------------
# problem structure
# ----------------------------------------------------------------------------


@dataclass
class BAStructure:
    num_cameras: int
    num_points: int
    obs_camera: np.ndarray
    obs_point: np.ndarray
    point_offsets: np.ndarray      # [P+1]
    camera_active: np.ndarray      # [C] bool
    point_active: np.ndarray       # [P] bool
    pair_a: np.ndarray             # ordered observation pairs of a common point,
    pair_b: np.ndarray             # sorted (stably) by camera block
    pair_key: np.ndarray           # cam(a) * C + cam(b), non-decreasing

    @classmethod
    def build(cls, obs_camera, obs_point, num_cameras, num_points, camera_active=None, point_active=None):
        obs_point = np.asarray(obs_point)
        if np.any(np.diff(obs_point) < 0):
            raise ValueError("observations must be grouped by point")
        counts = np.bincount(obs_point, minlength=num_points)
        offsets = np.zeros(num_points + 1, dtype=np.int64)
        offsets[1:] = np.cumsum(counts)
        # all ordered (a, b) pairs of observations of the same point
        pa, pb = [], []
        maxk = int(counts.max()) if counts.size else 0
        for k in range(1, maxk + 1):
            pts = np.nonzero(counts == k)[0]
            if pts.size == 0:
                continue
            base = offsets[pts]
            ii, jj = np.meshgrid(np.arange(k), np.arange(k), indexing="ij")
            pa.append((base[:, None] + ii.reshape(1, -1)).reshape(-1))
            pb.append((base[:, None] + jj.reshape(1, -1)).reshape(-1))
        pair_a = np.concatenate(pa) if pa else np.zeros(0, np.int64)
        pair_b = np.concatenate(pb) if pb else np.zeros(0, np.int64)
        cam = np.asarray(obs_camera, dtype=np.int64)
        key = cam[pair_a] * num_cameras + cam[pair_b]
        order = np.argsort(key, kind="stable")
        pair_a, pair_b, key = pair_a[order], pair_b[order], key[order]
        return cls(
            num_cameras=num_cameras,
            num_points=num_points,
            obs_camera=np.asarray(obs_camera, dtype=np.int64),
            obs_point=obs_point.astype(np.int64),
            point_offsets=offsets,
            camera_active=np.ones(num_cameras, bool) if camera_active is None else np.asarray(camera_active, bool),
            point_active=np.ones(num_points, bool) if point_active is None else np.asarray(point_active, bool),
            pair_a=pair_a,
            pair_b=pair_b,
            pair_key=key,
        )


@dataclass
class BAEvaluation:
    cost: float
    r: np.ndarray        # [N, 2] corrected residuals
    Jc: np.ndarray | None  # [N, 2, 9] corrected (zero for constant cameras)
    Jp: np.ndarray | None  # [N, 2, 3] corrected (zero for constant points)


def ba_evaluate(struct: BAStructure, cams, pts, obs, loss, jacobian: bool,
                camera_model: str = "angle_axis") -> BAEvaluation:
    c = cams[struct.obs_camera]
    p = pts[struct.obs_point]
    quaternion = camera_model != "angle_axis"
    if quaternion:
        c = ceres_order(c, camera_model)
    if jacobian:
        res, Jc, Jp = quaternion_residual_jacobian(c, p, obs) if quaternion else snavely_residual_jacobian(c, p, obs)
        Jc = Jc * struct.camera_active[struct.obs_camera][:, None, None]
        Jp = Jp * struct.point_active[struct.obs_point][:, None, None]
        J = np.concatenate([Jc, Jp], axis=2)
        cost_b, r, J = robustify(loss, res, J)
        return BAEvaluation(float(np.sum(cost_b)), r, J[:, :, :9], J[:, :, 9:])
    res = (quaternion_project(c, p) if quaternion else project(c, p)) - obs
    cost_b, r, _ = robustify(loss, res, None)
    return BAEvaluation(float(np.sum(cost_b)), r, None, None)


def _segment_sum(values: np.ndarray, offsets: np.ndarray) -> np.ndarray:
    """Sum rows of ``values`` over CSR segments; an empty segment (a point no
    camera observes) sums to zero.  ``np.add.reduceat`` alone would return the
    row at an empty segment's offset instead, so it runs on the non-empty
    segments only (each then ends where the next non-empty one starts)."""
    out = np.zeros((offsets.shape[0] - 1,) + values.shape[1:], dtype=values.dtype)
    nonempty = offsets[1:] > offsets[:-1]
    if nonempty.any():
        out[nonempty] = np.add.reduceat(values, offsets[:-1][nonempty], axis=0)
    return out


# ----------------------------------------------------------------------------
# block-sparse normal-equation quantities in the reference's column order:
#   active points first (ordering group 0), then active cameras
# ----------------------------------------------------------------------------


def gradient_blocks(struct: BAStructure, ev: BAEvaluation):
    gp = _segment_sum(np.einsum("nij,ni->nj", ev.Jp, ev.r), struct.point_offsets)
    gc = np.zeros((struct.num_cameras, 9))
    np.add.at(gc, struct.obs_camera, np.einsum("nij,ni->nj", ev.Jc, ev.r))
    return gp, gc


def column_vector(struct: BAStructure, vp: np.ndarray, vc: np.ndarray) -> np.ndarray:
    return np.concatenate([vp[struct.point_active].reshape(-1), vc[struct.camera_active].reshape(-1)])


def split_columns(struct: BAStructure, v: np.ndarray):
    npa = int(struct.point_active.sum())
    vp = np.zeros((struct.num_points, 3))
    vc = np.zeros((struct.num_cameras, 9))
    vp[struct.point_active] = v[: 3 * npa].reshape(-1, 3)
    vc[struct.camera_active] = v[3 * npa :].reshape(-1, 9)
    return vp, vc


def column_sq_norms(struct: BAStructure, ev: BAEvaluation):
    """colsum(J o J) per point and per camera column (linear.py:549)."""
    dp = _segment_sum(np.sum(ev.Jp * ev.Jp, axis=1), struct.point_offsets)
    dc = np.zeros((struct.num_cameras, 9))
    np.add.at(dc, struct.obs_camera, np.sum(ev.Jc * ev.Jc, axis=1))
    return dp, dc


def schur_system(struct: BAStructure, ev: BAEvaluation, Dp: np.ndarray, Dc: np.ndarray, *, chunk: int = 400_000):
    """Reduced camera system S, its right-hand side and the point factors.

    With b = -r (solver.py:143-145): H = J^T J + diag(D), rhs = -J^T r and
    S = Hbb - Hba Haa^-1 Hab, rhs_S = gb - Hba Haa^-1 ga (linear.py:681-699).
    Returns S over ALL cameras (constant cameras are dropped by the caller),
    rhs_S, Vinv [P,3,3], W [N,9,3] and the point right-hand sides ga [P,3].
    """
    C = struct.num_cameras
    # constant points have zero Jacobian columns; a unit diagonal keeps their
    # (unused) 3x3 block invertible without touching any active quantity
    Dp = np.where(struct.point_active[:, None], Dp, 1.0)
    V = _segment_sum(np.matmul(ev.Jp.transpose(0, 2, 1), ev.Jp), struct.point_offsets)
    V = V + Dp[:, :, None] * np.eye(3)[None]
    ga = -_segment_sum(np.einsum("nij,ni->nj", ev.Jp, ev.r), struct.point_offsets)
    gb = np.zeros((C, 9))
    np.add.at(gb, struct.obs_camera, -np.einsum("nij,ni->nj", ev.Jc, ev.r))
    U = np.zeros((C, 9, 9))
    np.add.at(U, struct.obs_camera, np.matmul(ev.Jc.transpose(0, 2, 1), ev.Jc))
    U = U + Dc[:, :, None] * np.eye(9)[None]
    Vinv = np.linalg.inv(V)
    W = np.matmul(ev.Jc.transpose(0, 2, 1), ev.Jp)              # [N, 9, 3]
    Y = np.matmul(W, Vinv[struct.obs_point])                    # W V^-1
    Wt = np.ascontiguousarray(W.transpose(0, 2, 1))             # [N, 3, 9]
    S = np.zeros((C * C, 9, 9))
    for s in range(0, struct.pair_a.shape[0], chunk):
        a = struct.pair_a[s : s + chunk]
        b = struct.pair_b[s : s + chunk]
        blk = struct.pair_key[s : s + chunk]
        contrib = np.matmul(Y[a], Wt[b])
        starts = np.nonzero(np.r_[True, blk[1:] != blk[:-1]])[0]
        S[blk[starts]] -= np.add.reduceat(contrib, starts, axis=0)
    S = S.reshape(C, C, 9, 9).transpose(0, 2, 1, 3).reshape(9 * C, 9 * C)
    for i in range(C):
        S[9 * i : 9 * i + 9, 9 * i : 9 * i + 9] += U[i]
    ga_obs = ga[struct.obs_point]
    corr = np.zeros((C, 9))
    np.add.at(corr, struct.obs_camera, np.einsum("nij,nj->ni", Y, ga_obs))
    rhs = (gb - corr).reshape(-1)
    return S, rhs, Vinv, W, ga


def schur_back_substitute(struct: BAStructure, Vinv, W, ga, dxc: np.ndarray) -> np.ndarray:
    """xa = Haa^-1 (ga - Hab xb) per point (linear.py:698)."""
    t = np.einsum("nij,ni->nj", W, dxc[struct.obs_camera])     # W^T dx_c
    tp = _segment_sum(t, struct.point_offsets)
    return np.einsum("pij,pj->pi", Vinv, ga - tp)


def block_jacobi_pcg(S, rhs, block, *, tol, max_iter, min_iter=0, preconditioner="schur_jacobi"):
    """``block``: one size for every diagonal block, or the list of block sizes."""
    """_preconditioned_conjugate_gradient (linear.py:1216-1256) on the explicit Schur complement.

    ``preconditioner`` follows build_schur_complement_preconditioner (linear.py:462-537):
    "schur_jacobi" - Cholesky (LU if not SPD) of the diagonal blocks of ``block``;
    "jacobi" - the inverse diagonal, 1 / max(S_kk, eps) (linear.py:531-537);
    "identity" - none (linear.py:480-481, 54-55).
    """
    n = rhs.shape[0]
    sizes = [block] * (n // block) if np.isscalar(block) else list(block)
    if preconditioner == "identity":
        def apply(r):
            return r
    elif preconditioner == "jacobi":
        dinv = 1.0 / np.maximum(np.diag(S), _EPS)

        def apply(r):
            return dinv * r
    elif preconditioner == "schur_jacobi":
        apply = _block_apply(S, sizes)
    else:
        raise ValueError(preconditioner)
    return _pcg(S, rhs, apply, tol=tol, max_iter=max_iter, min_iter=min_iter)


def _block_apply(S, sizes):
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    factors = []
    for i in range(len(sizes)):
        blk = S[offsets[i]:offsets[i + 1], offsets[i]:offsets[i + 1]]
        try:
            factors.append((np.linalg.cholesky(blk), True))
        except np.linalg.LinAlgError:
            factors.append((blk, False))

    def apply(r):
        z = np.empty_like(r)
        for i, (f, chol) in enumerate(factors):
            sl = slice(offsets[i], offsets[i + 1])
            if chol:
                z[sl] = np.linalg.solve(f.T, np.linalg.solve(f, r[sl]))
            else:
                z[sl] = np.linalg.solve(f, r[sl])
        return z

    return apply


def _pcg(S, rhs, apply, *, tol, max_iter, min_iter):
    n = rhs.shape[0]
    x = np.zeros(n)
    r = rhs - S @ x
    z = apply(r)
    p = z.copy()
    rz_old = float(r @ z)
    rhs_norm = max(float(np.linalg.norm(rhs)), _EPS)
    if min_iter <= 0 and np.linalg.norm(r) <= tol * rhs_norm:
        return x, 0, True
    iterations = 0
    success = False
    for k in range(max_iter):
        Ap = S @ p
        denom = float(p @ Ap)
        if abs(denom) <= _EPS:
            success = bool(np.linalg.norm(r) <= tol * rhs_norm)
            break
        alpha = rz_old / denom
        x = x + alpha * p
        r = r - alpha * Ap
        iterations = k + 1
        if iterations >= min_iter and np.linalg.norm(r) <= tol * rhs_norm:
            success = True
            break
        z = apply(r)
        rz_new = float(r @ z)
        if abs(rz_old) <= _EPS:
            break
        beta = rz_new / rz_old
        p = z + beta * p
        rz_old = rz_new
    return x, iterations, success


# ----------------------------------------------------------------------------
# the BA model used by the trust-region loop
# ----------------------------------------------------------------------------


class BAModel:
    """Block-sparse restatement of ``Problem.evaluate`` + the Schur solvers.

    ``linear_solver`` is "dense_schur" or "iterative_schur".  With
    ``dense_faithful=True`` the solve is instead the reference's dense algebra
    on the full Jacobian (only for small problems).
    """

    def __init__(self, data, *, linear_solver: str = "dense_schur", dense_faithful: bool = False,
                 camera_active=None, point_active=None):
        self.data = data
        self.struct = BAStructure.build(
            data.obs_camera, data.obs_point, data.num_cameras, data.num_points,
            camera_active=camera_active, point_active=point_active,
        )
        self.obs = np.asarray(data.observations, dtype=np.float64)
        self.loss = data.loss
        self.camera_model = getattr(data, "camera_model", "angle_axis")
        self.linear_solver = linear_solver
        self.dense_faithful = dense_faithful

    # -- state -------------------------------------------------------------
    def x0(self):
        return (np.array(self.data.cameras, dtype=np.float64), np.array(self.data.points, dtype=np.float64))

    def num_active_columns(self) -> int:
        return int(3 * self.struct.point_active.sum() + 9 * self.struct.camera_active.sum())

    def num_eliminate(self) -> int:
        """solver.py:729-741: points form the first ordering group."""
        if self.linear_solver not in ("dense_schur", "iterative_schur"):
            return 0
        npa = int(self.struct.point_active.sum())
        nca = int(self.struct.camera_active.sum())
        if npa == 0 or nca == 0:
            return 0
        return 3 * npa

    def state_norm(self, x) -> float:
        cams, pts = x
        return float(np.sqrt(np.sum(cams * cams) + np.sum(pts * pts)))

    def apply(self, x, step):
        cams, pts = x
        dp, dc = split_columns(self.struct, step)
        return (camera_plus(cams, dc, self.camera_model, self.struct.camera_active), pts + dp)

    # -- evaluation ----------------------------------------------------------
    def evaluate(self, x, jacobian: bool):
        cams, pts = x
        ev = ba_evaluate(self.struct, cams, pts, self.obs, self.loss, jacobian, self.camera_model)
        if not jacobian:
            return ev, None
        gp, gc = gradient_blocks(self.struct, ev)
        return ev, column_vector(self.struct, gp, gc)

    def column_sq(self, ev) -> np.ndarray:
        dp, dc = column_sq_norms(self.struct, ev)
        return column_vector(self.struct, dp, dc)

    def dense_jacobian(self, ev) -> np.ndarray:
        st = self.struct
        n_obs = ev.r.shape[0]
        npa = int(st.point_active.sum())
        pcol = -np.ones(st.num_points, np.int64)
        pcol[st.point_active] = 3 * np.arange(npa)
        ccol = -np.ones(st.num_cameras, np.int64)
        ccol[st.camera_active] = 3 * npa + 9 * np.arange(int(st.camera_active.sum()))
        J = np.zeros((2 * n_obs, self.num_active_columns()))
        rows = np.arange(n_obs)
        for k in range(2):
            for j in range(3):
                ok = pcol[st.obs_point] >= 0
                J[2 * rows[ok] + k, pcol[st.obs_point[ok]] + j] = ev.Jp[ok, k, j]
            for j in range(9):
                ok = ccol[st.obs_camera] >= 0
                J[2 * rows[ok] + k, ccol[st.obs_camera[ok]] + j] = ev.Jc[ok, k, j]
        return J

    def model_decrease(self, ev, step) -> float:
        """solver.py:798-802: max(0, 1/2 r.r - 1/2 |r + J step|^2)."""
        dp, dc = split_columns(self.struct, step)
        jd = np.einsum("nij,nj->ni", ev.Jc, dc[self.struct.obs_camera]) + np.einsum(
            "nij,nj->ni", ev.Jp, dp[self.struct.obs_point]
        )
        r = ev.r.reshape(-1)
        after = r + jd.reshape(-1)
        return max(0.0, float(0.5 * (r @ r) - 0.5 * (after @ after)))

    # -- dogleg ----------------------------------------------------------------
    def jmul(self, ev, v: np.ndarray) -> np.ndarray:
        """J v for a vector in the active column order (flattened per residual)."""
        dp, dc = split_columns(self.struct, v)
        jd = np.einsum("nij,nj->ni", ev.Jc, dc[self.struct.obs_camera]) + np.einsum(
            "nij,nj->ni", ev.Jp, dp[self.struct.obs_point]
        )
        return jd.reshape(-1)

    def gauss_newton_step(self, ev) -> np.ndarray:
        """solve(J^T J, -g) (linear.py:570-571) by eliminating the points with no
        damping: the exact solution whenever J^T J is nonsingular (which needs
        the gauge of the reconstruction fixed, e.g. by constant cameras)."""
        if self.dense_faithful:
            from .lm import _solve_square

            J = self.dense_jacobian(ev)
            return _solve_square(J.T @ J, -(J.T @ ev.r.reshape(-1)))
        st = self.struct
        zero_p = np.zeros((st.num_points, 3))
        zero_c = np.zeros((st.num_cameras, 9))
        S, rhs, Vinv, W, ga = schur_system(st, ev, zero_p, zero_c)
        act = np.repeat(st.camera_active, 9)
        dxc = np.zeros(9 * st.num_cameras)
        if act.any():
            dxc[act] = np.linalg.solve(S[np.ix_(act, act)], rhs[act])
        dxc = dxc.reshape(-1, 9)
        dxp = schur_back_substitute(st, Vinv, W, ga, dxc)
        return column_vector(st, dxp, dxc)

    # -- linear solve ----------------------------------------------------------
    def solve(self, ev, damping: np.ndarray, opts) -> tuple[np.ndarray, int]:
        if self.dense_faithful:
            from .lm import dense_linear_solve

            J = self.dense_jacobian(ev)
            return dense_linear_solve(J, -ev.r.reshape(-1), damping, self.num_eliminate(), self.linear_solver, opts,
                                      block_sizes=self._block_sizes())
        st = self.struct
        Dp, Dc = split_columns(st, damping)
        ne = self.num_eliminate()
        if ne == 0:
            # damped least squares on the point-only or camera-only system
            if st.point_active.any():
                S, rhs, Vinv, W, ga = schur_system(st, ev, Dp, Dc)
                dxp = np.einsum("pij,pj->pi", Vinv, ga)
            else:
                dxp = np.zeros((st.num_points, 3))
            cam_idx = np.nonzero(st.camera_active)[0]
            if cam_idx.size:
                # no points active: solve the camera normal equations directly
                U = np.zeros((st.num_cameras, 9, 9))
                np.add.at(U, st.obs_camera, np.einsum("nki,nkj->nij", ev.Jc, ev.Jc))
                gb = np.zeros((st.num_cameras, 9))
                np.add.at(gb, st.obs_camera, -np.einsum("nij,ni->nj", ev.Jc, ev.r))
                dxc = np.zeros((st.num_cameras, 9))
                Hc = np.zeros((cam_idx.size * 9, cam_idx.size * 9))
                for a, i in enumerate(cam_idx):
                    Hc[9 * a : 9 * a + 9, 9 * a : 9 * a + 9] = U[i] + np.diag(Dc[i])
                # cameras couple only through points, which are all constant here
                sol = np.linalg.solve(Hc, gb[cam_idx].reshape(-1))
                dxc[cam_idx] = sol.reshape(-1, 9)
            else:
                dxc = np.zeros((st.num_cameras, 9))
            return column_vector(st, dxp, dxc), 1
        S, rhs, Vinv, W, ga = schur_system(st, ev, Dp, Dc)
        act = np.repeat(st.camera_active, 9)
        S_a = S[np.ix_(act, act)]
        rhs_a = rhs[act]
        iters = 1
        if self.linear_solver == "iterative_schur":
            xb, iters, _ = block_jacobi_pcg(S_a, rhs_a, self._camera_block_sizes() * int(st.camera_active.sum()),
                                            tol=opts.eta, max_iter=opts.max_linear_solver_iterations,
                                            preconditioner=opts.preconditioner_type,
                                            min_iter=opts.min_linear_solver_iterations)
        else:
            xb = np.linalg.solve(S_a, rhs_a)
        dxc = np.zeros(9 * st.num_cameras)
        dxc[act] = xb
        dxc = dxc.reshape(-1, 9)
        dxp = schur_back_substitute(st, Vinv, W, ga, dxc)
        return column_vector(st, dxp, dxc), iters

    def _camera_block_sizes(self) -> list:
        """Parameter-block tangent sizes of one camera: the rotation and the rest are separate
        blocks for quaternion cameras."""
        return [9] if self.camera_model == "angle_axis" else [3, 6]

    def _block_sizes(self):
        st = self.struct
        return [3] * int(st.point_active.sum()) + self._camera_block_sizes() * int(st.camera_active.sum())
