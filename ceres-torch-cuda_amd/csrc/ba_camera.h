// Camera models and the robust corrector of a 2-residual block, shared by the
// bundle-adjustment kernels (LM and line search).
#pragma once

#include "ba_device.h"

namespace ct {

// ---------------------------------------------------------------------------
// Snavely camera model (simple_bundle_adjuster.py:73-79).  The rotation is the
// reference's angle-axis -> unit quaternion -> rotate formula
// (rotation.py:104-110, 211-222); its derivative uses Rodrigues' form, which
// is the same function.
// ---------------------------------------------------------------------------
// p + 2w(u x p) + u x (2 u x p) for a unit quaternion (w, u) (rotation.py:211-214).
__device__ __forceinline__ void unit_quat_rotate(double w, double u0, double u1, double u2, const double* X,
                                                 double* out) {
  const double v0 = 2.0 * (u1 * X[2] - u2 * X[1]);
  const double v1 = 2.0 * (u2 * X[0] - u0 * X[2]);
  const double v2 = 2.0 * (u0 * X[1] - u1 * X[0]);
  out[0] = X[0] + w * v0 + (u1 * v2 - u2 * v1);
  out[1] = X[1] + w * v1 + (u2 * v0 - u0 * v2);
  out[2] = X[2] + w * v2 + (u0 * v1 - u1 * v0);
}

__device__ __forceinline__ void aa_rotate(const double* aa, const double* X, double* out) {
  const double t2 = aa[0] * aa[0] + aa[1] * aa[1] + aa[2] * aa[2];
  const double theta = sqrt(t2);
  const double half = 0.5 * theta;
  double s, c;
  sincos(half, &s, &c);
  const double k = theta <= 2.220446049250313e-16 ? 0.5 - t2 / 48.0 : s / theta;
  double w = c, u0 = k * aa[0], u1 = k * aa[1], u2 = k * aa[2];
  const double nrm = fmax(sqrt(w * w + u0 * u0 + u1 * u1 + u2 * u2), 2.2250738585072014e-308);
  w /= nrm; u0 /= nrm; u1 /= nrm; u2 /= nrm;
  unit_quat_rotate(w, u0, u1, u2, X, out);
}

__device__ __forceinline__ void snavely_residual(const double* cam, const double* X, const double* obs,
                                                 double* r) {
  double p[3];
  aa_rotate(cam, X, p);
  p[0] += cam[3]; p[1] += cam[4]; p[2] += cam[5];
  const double xp = -p[0] / p[2];
  const double yp = -p[1] / p[2];
  const double r2 = xp * xp + yp * yp;
  const double d = 1.0 + r2 * (cam[7] + cam[8] * r2);
  const double fd = cam[6] * d;
  r[0] = fd * xp - obs[0];
  r[1] = fd * yp - obs[1];
}

// Residual and full Jacobian of one observation.
__device__ __forceinline__ void snavely_jacobian(const double* cam, const double* X, const double* obs, double (&r)[2],
                                                 double (&Jc)[2][9], double (&Jp)[2][3]) {
  const double a0 = cam[0], a1 = cam[1], a2 = cam[2];
  double p[3];
  aa_rotate(cam, X, p);
  p[0] += cam[3]; p[1] += cam[4]; p[2] += cam[5];
  // Rodrigues coefficients (series below theta = 1e-4)
  const double t2 = a0 * a0 + a1 * a1 + a2 * a2;
  double A, B, C1, C2;
  if (t2 < 1e-8) {
    A = 1.0 - t2 / 6.0 + t2 * t2 / 120.0;
    B = 0.5 - t2 / 24.0 + t2 * t2 / 720.0;
    C1 = -1.0 / 3.0 + t2 / 30.0 - t2 * t2 / 840.0;
    C2 = -1.0 / 12.0 + t2 / 180.0 - t2 * t2 / 6720.0;
  } else {
    const double t = sqrt(t2);
    double st, cth;
    sincos(t, &st, &cth);
    A = st / t;
    B = (1.0 - cth) / t2;
    C1 = (t * cth - st) / (t2 * t);
    C2 = (t * st - 2.0 * (1.0 - cth)) / (t2 * t2);
  }
  const double x0 = X[0], x1 = X[1], x2 = X[2];
  const double wx0 = a1 * x2 - a2 * x1, wx1 = a2 * x0 - a0 * x2, wx2 = a0 * x1 - a1 * x0;      // w x X
  const double ww0 = a1 * wx2 - a2 * wx1, ww1 = a2 * wx0 - a0 * wx2, ww2 = a0 * wx1 - a1 * wx0;  // w x (w x X)
  const double wd = a0 * x0 + a1 * x1 + a2 * x2;
  const double aa[3] = {a0, a1, a2};
  const double xx[3] = {x0, x1, x2};
  const double wx[3] = {wx0, wx1, wx2};
  const double ww[3] = {ww0, ww1, ww2};
  // dR[i][j] = d(Rx)_i / d w_j
  double dR[3][3];
  // -A [x]_x
  const double sk[3][3] = {{0.0, -x2, x1}, {x2, 0.0, -x0}, {-x1, x0, 0.0}};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      dR[i][j] = -A * sk[i][j] + C1 * wx[i] * aa[j] +
                 B * ((i == j ? wd : 0.0) + aa[i] * xx[j] - 2.0 * xx[i] * aa[j]) + C2 * ww[i] * aa[j];
    }
  }
  // R = I + A W + B W^2, W = [w]_x
  const double W[3][3] = {{0.0, -a2, a1}, {a2, 0.0, -a0}, {-a1, a0, 0.0}};
  double R[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double w2 = W[i][0] * W[0][j] + W[i][1] * W[1][j] + W[i][2] * W[2][j];
      R[i][j] = (i == j ? 1.0 : 0.0) + A * W[i][j] + B * w2;
    }
  }
  const double f = cam[6], k1 = cam[7], k2 = cam[8];
  const double iz = 1.0 / p[2];
  const double xp = -p[0] * iz;
  const double yp = -p[1] * iz;
  const double r2 = xp * xp + yp * yp;
  const double d = 1.0 + r2 * (k1 + k2 * r2);
  const double fd = f * d;
  r[0] = fd * xp - obs[0];
  r[1] = fd * yp - obs[1];
  const double dd = k1 + 2.0 * k2 * r2;
  const double u_x = f * (d + 2.0 * xp * xp * dd);
  const double u_y = f * 2.0 * xp * yp * dd;
  const double v_y = f * (d + 2.0 * yp * yp * dd);
  // d(xp,yp)/dp = [[-iz, 0, -xp iz], [0, -iz, -yp iz]]
  double G[2][3];
  G[0][0] = -u_x * iz; G[0][1] = -u_y * iz; G[0][2] = -(u_x * xp + u_y * yp) * iz;
  G[1][0] = -u_y * iz; G[1][1] = -v_y * iz; G[1][2] = -(u_y * xp + v_y * yp) * iz;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      Jc[k][j] = G[k][0] * dR[0][j] + G[k][1] * dR[1][j] + G[k][2] * dR[2][j];
      Jc[k][3 + j] = G[k][j];
      Jp[k][j] = G[k][0] * R[0][j] + G[k][1] * R[1][j] + G[k][2] * R[2][j];
    }
  }
  Jc[0][6] = d * xp; Jc[1][6] = d * yp;
  Jc[0][7] = f * xp * r2; Jc[1][7] = f * yp * r2;
  Jc[0][8] = f * xp * r2 * r2; Jc[1][8] = f * yp * r2 * r2;
}

// ---------------------------------------------------------------------------
// Quaternion cameras [q(4), t(3), f, k1, k2], held here in Ceres order
// (w, x, y, z; the Eigen order x, y, z, w is permuted on load and store).  The
// rotation is quaternion_rotate_point(q, X) = unit_quaternion_rotate_point(
// q / |q|, X) (rotation.py:68-69, 211-218).  The camera's tangent is 9-wide:
// the 3-vector of QuaternionManifold.plus (manifolds.py:127-137), then t, f,
// k1, k2.  The reference's tangent Jacobian is J_ambient(q) plus_jacobian(q)
// (problem.py:505-514); along the manifold the rotated point R X moves by
// d(R X)/d delta = -2 [R X]_x (q_delta = (cos|d|, sin|d| d/|d|) rotates by
// 2|d|), and the normalisation in quaternion_rotate_point divides by |q|.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load_quat_camera(const double* __restrict__ src, int model, double (&cam)[10]) {
  if (model == CT_CAMERA_EIGEN_QUATERNION) {
    cam[0] = src[3]; cam[1] = src[0]; cam[2] = src[1]; cam[3] = src[2];
  } else {
    cam[0] = src[0]; cam[1] = src[1]; cam[2] = src[2]; cam[3] = src[3];
  }
#pragma unroll
  for (int j = 4; j < 10; ++j) cam[j] = src[j];
}

__device__ __forceinline__ void store_quat_camera(const double (&cam)[10], int model, double* __restrict__ dst) {
  if (model == CT_CAMERA_EIGEN_QUATERNION) {
    dst[3] = cam[0]; dst[0] = cam[1]; dst[1] = cam[2]; dst[2] = cam[3];
  } else {
    dst[0] = cam[0]; dst[1] = cam[1]; dst[2] = cam[2]; dst[3] = cam[3];
  }
#pragma unroll
  for (int j = 4; j < 10; ++j) dst[j] = cam[j];
}

// R(q / |q|) X; returns 1 / |q| (|q| clamped at the smallest normal, as normalize_quaternion).
__device__ __forceinline__ double quat_rotate(const double* q, const double* X, double* out) {
  const double inv = 1.0 / fmax(sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]),
                                2.2250738585072014e-308);
  unit_quat_rotate(q[0] * inv, q[1] * inv, q[2] * inv, q[3] * inv, X, out);
  return inv;
}

__device__ __forceinline__ void quat_camera_residual(const double (&cam)[10], const double* X, const double* obs,
                                                     double* r) {
  double p[3];
  quat_rotate(cam, X, p);
  p[0] += cam[4]; p[1] += cam[5]; p[2] += cam[6];
  const double xp = -p[0] / p[2];
  const double yp = -p[1] / p[2];
  const double r2 = xp * xp + yp * yp;
  const double d = 1.0 + r2 * (cam[8] + cam[9] * r2);
  const double fd = cam[7] * d;
  r[0] = fd * xp - obs[0];
  r[1] = fd * yp - obs[1];
}

__device__ __forceinline__ void quat_camera_jacobian(const double (&cam)[10], const double* X, const double* obs,
                                                     double (&r)[2], double (&Jc)[2][9], double (&Jp)[2][3]) {
  double pr[3];
  const double inv = quat_rotate(cam, X, pr);
  const double p0 = pr[0] + cam[4], p1 = pr[1] + cam[5], p2 = pr[2] + cam[6];
  const double f = cam[7], k1 = cam[8], k2 = cam[9];
  const double iz = 1.0 / p2;
  const double xp = -p0 * iz;
  const double yp = -p1 * iz;
  const double r2 = xp * xp + yp * yp;
  const double d = 1.0 + r2 * (k1 + k2 * r2);
  const double fd = f * d;
  r[0] = fd * xp - obs[0];
  r[1] = fd * yp - obs[1];
  const double dd = k1 + 2.0 * k2 * r2;
  const double u_x = f * (d + 2.0 * xp * xp * dd);
  const double u_y = f * 2.0 * xp * yp * dd;
  const double v_y = f * (d + 2.0 * yp * yp * dd);
  double G[2][3];
  G[0][0] = -u_x * iz; G[0][1] = -u_y * iz; G[0][2] = -(u_x * xp + u_y * yp) * iz;
  G[1][0] = -u_y * iz; G[1][1] = -v_y * iz; G[1][2] = -(u_y * xp + v_y * yp) * iz;
  // rotation columns: -2 [R X]_x / |q|
  const double s = -2.0 * inv;
  const double D[3][3] = {{0.0, -s * pr[2], s * pr[1]}, {s * pr[2], 0.0, -s * pr[0]}, {-s * pr[1], s * pr[0], 0.0}};
  // d(R X)/dX = I + 2w [u]_x + 2 (u u^T - |u|^2 I), the derivative of the rotate formula at the unit quaternion
  const double w = cam[0] * inv, u0 = cam[1] * inv, u1 = cam[2] * inv, u2 = cam[3] * inv;
  const double uu = u0 * u0 + u1 * u1 + u2 * u2;
  const double M[3][3] = {{1.0 + 2.0 * (u0 * u0 - uu), 2.0 * (u0 * u1 - w * u2), 2.0 * (u0 * u2 + w * u1)},
                          {2.0 * (u1 * u0 + w * u2), 1.0 + 2.0 * (u1 * u1 - uu), 2.0 * (u1 * u2 - w * u0)},
                          {2.0 * (u2 * u0 - w * u1), 2.0 * (u2 * u1 + w * u0), 1.0 + 2.0 * (u2 * u2 - uu)}};
#pragma unroll
  for (int k = 0; k < 2; ++k) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      Jc[k][j] = G[k][0] * D[0][j] + G[k][1] * D[1][j] + G[k][2] * D[2][j];
      Jc[k][3 + j] = G[k][j];
      Jp[k][j] = G[k][0] * M[0][j] + G[k][1] * M[1][j] + G[k][2] * M[2][j];
    }
  }
  Jc[0][6] = d * xp; Jc[1][6] = d * yp;
  Jc[0][7] = f * xp * r2; Jc[1][7] = f * yp * r2;
  Jc[0][8] = f * xp * r2 * r2; Jc[1][8] = f * yp * r2 * r2;
}

// QuaternionManifold.plus (manifolds.py:130-137) on the rotation and Euclidean
// Plus on t, f, k1, k2: cam (Ceres order) <- Plus(cam, d[0..8]).  A constant
// rotation block (rotate false) keeps its stored values: the reference's
// apply_delta (problem.py:462-474) only moves blocks with a tangent space.
__device__ __forceinline__ void quat_camera_plus(double (&cam)[10], const double* d, bool rotate) {
#pragma unroll
  for (int j = 0; j < 6; ++j) cam[4 + j] += d[3 + j];
  if (!rotate) return;
  const double inv = 1.0 / fmax(sqrt(cam[0] * cam[0] + cam[1] * cam[1] + cam[2] * cam[2] + cam[3] * cam[3]),
                                2.2250738585072014e-308);
  const double x0 = cam[0] * inv, x1 = cam[1] * inv, x2 = cam[2] * inv, x3 = cam[3] * inv;
  const double nd = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  double sn, cs;
  sincos(nd, &sn, &cs);
  const double scale = nd > 2.220446049250313e-16 ? sn / nd : 1.0 - nd * nd / 6.0;
  const double z0 = cs, z1 = scale * d[0], z2 = scale * d[1], z3 = scale * d[2];
  // quaternion_product(q_delta, x) (rotation.py:72-83)
  const double y0 = z0 * x0 - z1 * x1 - z2 * x2 - z3 * x3;
  const double y1 = z0 * x1 + z1 * x0 + z2 * x3 - z3 * x2;
  const double y2 = z0 * x2 - z1 * x3 + z2 * x0 + z3 * x1;
  const double y3 = z0 * x3 + z1 * x2 - z2 * x1 + z3 * x0;
  const double iy = 1.0 / fmax(sqrt(y0 * y0 + y1 * y1 + y2 * y2 + y3 * y3), 2.2250738585072014e-308);
  cam[0] = y0 * iy; cam[1] = y1 * iy; cam[2] = y2 * iy; cam[3] = y3 * iy;
}

// The rotation of quaternion camera ci moves unless its block is constant.
__device__ __forceinline__ bool rotation_free(const BAView& v, int ci) {
  return v.camera_mask == nullptr || v.camera_mask[9 * (size_t)ci] != 0;
}

// The problem's camera model: residual (and Jacobian) of one observation of
// camera ci, whose values start at cams + cam_stride * ci.
__device__ __forceinline__ void camera_jacobian(const BAView& v, const double* __restrict__ cams, int ci,
                                                const double* X, const double* obs, double (&r)[2],
                                                double (&Jc)[2][9], double (&Jp)[2][3]) {
  if (v.cam_model == CT_CAMERA_ANGLE_AXIS) {
    double cam[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) cam[j] = cams[9 * (size_t)ci + j];
    snavely_jacobian(cam, X, obs, r, Jc, Jp);
  } else {
    double cam[10];
    load_quat_camera(cams + 10 * (size_t)ci, v.cam_model, cam);
    quat_camera_jacobian(cam, X, obs, r, Jc, Jp);
  }
}

__device__ __forceinline__ void camera_residual(const BAView& v, const double* __restrict__ cams, int ci,
                                                const double* X, const double* obs, double* r) {
  if (v.cam_model == CT_CAMERA_ANGLE_AXIS) {
    snavely_residual(cams + 9 * (size_t)ci, X, obs, r);
  } else {
    double cam[10];
    load_quat_camera(cams + 10 * (size_t)ci, v.cam_model, cam);
    quat_camera_residual(cam, X, obs, r);
  }
}

// Residual of camera ci after its step d[0..8] (Plus applied on the fly, as the
// camera step stores it).
__device__ __forceinline__ void camera_residual_after_step(const BAView& v, const double* __restrict__ cams, int ci,
                                                           const double* d, const double* X, const double* obs,
                                                           double* r) {
  if (v.cam_model == CT_CAMERA_ANGLE_AXIS) {
    double cam[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) cam[j] = cams[9 * (size_t)ci + j] + d[j];
    snavely_residual(cam, X, obs, r);
  } else {
    double cam[10];
    load_quat_camera(cams + 10 * (size_t)ci, v.cam_model, cam);
    quat_camera_plus(cam, d, rotation_free(v, ci));
    quat_camera_residual(cam, X, obs, r);
  }
}

// Step of quaternion camera ci: ncams <- Plus(cams, dxc); adds |dx|^2 and |x'|^2
// (all ten stored values) and flags a non-finite step.
__device__ __forceinline__ void step_quat_camera(const BAView& v, const double* __restrict__ cams,
                                                 double* __restrict__ ncams, const double* __restrict__ dxc, int ci,
                                                 double& dx2, double& x2, double& bad) {
  double d[9];
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    d[j] = dxc[9 * (size_t)ci + j];
    dx2 += d[j] * d[j];
    if (!isfinite(d[j])) bad = 1.0;
  }
  double cam[10];
  load_quat_camera(cams + 10 * (size_t)ci, v.cam_model, cam);
  quat_camera_plus(cam, d, rotation_free(v, ci));
  store_quat_camera(cam, v.cam_model, ncams + 10 * (size_t)ci);
#pragma unroll
  for (int j = 0; j < 10; ++j) x2 += cam[j] * cam[j];
}

// Ceres corrector on a 2-residual block (losses.py:158-193).  Returns the
// block cost 1/2 rho0 and rescales r, Jc, Jp in place.
__device__ __forceinline__ double robustify2(const ct_loss& L, double (&r)[2], double (&Jc)[2][9],
                                             double (&Jp)[2][3]) {
  const double s = r[0] * r[0] + r[1] * r[1];
  if (L.type == CT_LOSS_TRIVIAL && L.scale == 1.0) return 0.5 * s;
  double rho1, rho2;
  const double rho0 = loss_rho(L, s, &rho1, &rho2);
  const double sq = sqrt(fmax(rho1, 0.0));
  double scale = sq, alpha_s = 0.0;
  if (s > 0.0 && rho2 > 0.0) {
    const double alpha = 1.0 - sqrt(1.0 + 2.0 * s * rho2 / rho1);
    scale = sq / (1.0 - alpha);
    alpha_s = alpha / s;
  }
  // J <- sqrt(rho1) (J - alpha/s r r^T J); alpha = 0 gives the plain scaling
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    const double rtj = alpha_s * (r[0] * Jc[0][j] + r[1] * Jc[1][j]);
    Jc[0][j] = sq * (Jc[0][j] - r[0] * rtj);
    Jc[1][j] = sq * (Jc[1][j] - r[1] * rtj);
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const double rtj = alpha_s * (r[0] * Jp[0][j] + r[1] * Jp[1][j]);
    Jp[0][j] = sq * (Jp[0][j] - r[0] * rtj);
    Jp[1][j] = sq * (Jp[1][j] - r[1] * rtj);
  }
  r[0] *= scale;
  r[1] *= scale;
  return 0.5 * rho0;
}

__device__ __forceinline__ double robust_cost2(const ct_loss& L, const double* r) {
  const double s = r[0] * r[0] + r[1] * r[1];
  if (L.type == CT_LOSS_TRIVIAL && L.scale == 1.0) return 0.5 * s;
  double rho1, rho2;
  return 0.5 * loss_rho(L, s, &rho1, &rho2);
}

}  // namespace ct
