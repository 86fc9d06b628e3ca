"""Quaternion manifolds and quaternion cameras: the host API and the oracle, pinned on the CPU.

Pins (each would fail on a plausible mistake - a missing factor 2, a product in the
wrong order, a dropped 1/|q|, a transposed plus_jacobian, a wrong block split):

* the reference's known answer for Plus/Minus (tests/test_manifolds_rotation.py:6-13);
* Plus against scipy's rotation composition (q_delta rotates by 2|delta|, applied on the left);
* the quaternion projection against the angle-axis projection of the same rotation;
* the closed-form tangent Jacobian against the reference's own procedure, autograd of the
  residual times plus_jacobian (problem.py:505-514), at non-unit quaternions, and against
  finite differences of r(Plus(x, delta)) at unit quaternions;
* the retained preconditioner blocks (linear.py:1269-1289) by known answers, and the
  block-sparse oracle solve against the dense restatement that applies that rule.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch
from scipy.spatial.transform import Rotation

import ceres_mi355x as tc
from oracle import ba as oba
from oracle import generators as gen
from oracle.lm import OracleOptions, retained_block_sizes, trust_region_solve


def test_quaternion_plus_minus_roundtrip_reference_known_answer():
    """tests/test_manifolds_rotation.py:6-13 of the reference, for the host manifold and the oracle."""
    q = torch.tensor([1.0, 0.0, 0.0, 0.0], dtype=torch.float64)
    delta = torch.tensor([0.01, -0.02, 0.03], dtype=torch.float64)
    m = tc.QuaternionManifold()
    y = m.plus(q, delta)
    torch.testing.assert_close(m.minus(y, q), delta, rtol=1e-8, atol=1e-10)
    torch.testing.assert_close(torch.linalg.norm(y), torch.tensor(1.0, dtype=torch.float64))
    y_o = oba.quaternion_plus(q.numpy()[None], delta.numpy()[None])[0]
    np.testing.assert_allclose(y_o, y.numpy(), rtol=0, atol=1e-15)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_quaternion_plus_is_a_left_rotation_by_twice_delta(seed):
    rng = np.random.default_rng(seed)
    q = rng.normal(size=(6, 4)) * 1.7             # not unit: Plus normalises first
    d = rng.normal(size=(6, 3)) * 0.3
    d[0] = 0.0                                    # the small-angle branch
    d[1] = 1e-17
    y = oba.quaternion_plus(q, d)
    qn = q / np.linalg.norm(q, axis=1, keepdims=True)
    expect = (Rotation.from_rotvec(2.0 * d) * Rotation.from_quat(qn[:, [1, 2, 3, 0]])).as_quat()[:, [3, 0, 1, 2]]
    sign = np.sign(np.sum(expect * y, axis=1, keepdims=True))
    np.testing.assert_allclose(y, sign * expect, atol=1e-14)
    np.testing.assert_allclose(np.linalg.norm(y, axis=1), 1.0, atol=1e-15)
    # the host manifold agrees with the oracle
    for k in range(q.shape[0]):
        yk = tc.QuaternionManifold().plus(torch.tensor(q[k]), torch.tensor(d[k]))
        np.testing.assert_allclose(yk.numpy(), y[k], atol=1e-15)


def test_eigen_quaternion_manifold_is_the_ceres_one_reordered():
    rng = np.random.default_rng(3)
    q = rng.normal(size=4)
    d = rng.normal(size=3) * 0.2
    ceres = tc.QuaternionManifold()
    eigen = tc.EigenQuaternionManifold()
    qe = torch.tensor(q[[1, 2, 3, 0]])
    y_c = ceres.plus(torch.tensor(q), torch.tensor(d)).numpy()
    y_e = eigen.plus(qe, torch.tensor(d)).numpy()
    np.testing.assert_allclose(y_e, y_c[[1, 2, 3, 0]], atol=1e-15)
    np.testing.assert_allclose(eigen.plus_jacobian(qe).numpy(), ceres.plus_jacobian(torch.tensor(q)).numpy()[[1, 2, 3, 0]])
    # oracle: the same step in either storage order
    cams = np.concatenate([q, rng.normal(size=6)])[None]
    dc = rng.normal(size=(1, 9)) * 0.1
    out_c = oba.camera_plus(cams, dc, "quaternion")
    out_e = oba.camera_plus(oba.storage_order(cams, "eigen_quaternion"), dc, "eigen_quaternion")
    np.testing.assert_allclose(oba.ceres_order(out_e, "eigen_quaternion"), out_c, atol=1e-15)


@pytest.mark.parametrize("order", ["ceres", "eigen"])
def test_quaternion_projection_equals_angle_axis_projection(order):
    data = gen.ba_config("ba4")
    qdata = gen.with_quaternion_cameras(data, order, scale=np.linspace(0.5, 2.0, data.num_cameras))
    assert qdata.cameras.shape == (data.num_cameras, 10)
    st = oba.BAStructure.build(data.obs_camera, data.obs_point, data.num_cameras, data.num_points)
    a = oba.ba_evaluate(st, data.cameras, data.points, data.observations, data.loss, False)
    b = oba.ba_evaluate(st, qdata.cameras, qdata.points, qdata.observations, qdata.loss, False, qdata.camera_model)
    np.testing.assert_allclose(b.r, a.r, rtol=1e-12, atol=1e-9)
    # the host cost function gives the same residuals
    k = 5
    c, p = int(data.obs_camera[k]), int(data.obs_point[k])
    cost = tc.QuaternionReprojectionError(*data.observations[k], order=order)
    r = cost.residuals(torch.tensor(qdata.cameras[c, :4]), torch.tensor(qdata.cameras[c, 4:]),
                       torch.tensor(qdata.points[p]))
    np.testing.assert_allclose(r.numpy(), a.r[k], rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("order", ["ceres", "eigen"])
def test_quaternion_tangent_jacobian_is_ambient_jacobian_times_plus_jacobian(order):
    """problem.py:505-514: the reference multiplies the autograd Jacobian of the residual in the
    stored quaternion by plus_jacobian; the oracle's closed form must equal it at non-unit q."""
    data = gen.ba_config("ba4")
    scale = np.linspace(0.6, 1.9, data.num_cameras)
    qdata = gen.with_quaternion_cameras(data, order, scale=scale)
    st = oba.BAStructure.build(data.obs_camera, data.obs_point, data.num_cameras, data.num_points)
    ev = oba.ba_evaluate(st, qdata.cameras, qdata.points, qdata.observations, qdata.loss, True, qdata.camera_model)
    manifold = tc.QuaternionManifold() if order == "ceres" else tc.EigenQuaternionManifold()
    for k in range(0, data.num_observations, 17):
        c, p = int(data.obs_camera[k]), int(data.obs_point[k])
        q = torch.tensor(qdata.cameras[c, :4], requires_grad=True)
        rest = torch.tensor(qdata.cameras[c, 4:], requires_grad=True)
        X = torch.tensor(qdata.points[p], requires_grad=True)
        cost = tc.QuaternionReprojectionError(*data.observations[k], order=order)
        J_q, J_rest, J_X = torch.autograd.functional.jacobian(lambda a, b, x: cost.residuals(a, b, x), (q, rest, X))
        J_rot = (J_q @ manifold.plus_jacobian(q.detach())).numpy()
        np.testing.assert_allclose(ev.Jc[k, :, 0:3], J_rot, rtol=1e-10, atol=1e-8)
        np.testing.assert_allclose(ev.Jc[k, :, 3:9], J_rest.numpy(), rtol=1e-10, atol=1e-8)
        np.testing.assert_allclose(ev.Jp[k], J_X.numpy(), rtol=1e-10, atol=1e-8)


def test_quaternion_tangent_jacobian_is_the_derivative_along_plus_at_unit_quaternions():
    data = gen.ba_config("ba4")
    qdata = gen.with_quaternion_cameras(data, "ceres")
    st = oba.BAStructure.build(data.obs_camera, data.obs_point, data.num_cameras, data.num_points)
    ev = oba.ba_evaluate(st, qdata.cameras, qdata.points, qdata.observations, qdata.loss, True, "quaternion")
    h = 1e-6
    for k in (0, 31, 99):
        c, p = int(data.obs_camera[k]), int(data.obs_point[k])
        cam, X = qdata.cameras[c:c + 1], qdata.points[p:p + 1]
        for j in range(3):
            d = np.zeros((1, 9))
            d[0, j] = h
            rp = oba.quaternion_project(oba.camera_plus(cam, d, "quaternion"), X)[0]
            rm = oba.quaternion_project(oba.camera_plus(cam, -d, "quaternion"), X)[0]
            np.testing.assert_allclose(ev.Jc[k, :, j], (rp - rm) / (2 * h), rtol=1e-6, atol=1e-5)


def test_retained_block_sizes_known_answers():
    """linear.py:1269-1289: eliminated leading blocks drop out, a straddled block keeps its tail."""
    assert retained_block_sizes([3, 3, 9, 9], 6, 24) == [9, 9]
    assert retained_block_sizes([3, 3, 3, 6, 3, 6], 6, 24) == [3, 6, 3, 6]
    assert retained_block_sizes([3, 3, 9, 9], 4, 24) == [2, 9, 9]
    assert retained_block_sizes(None, 2, 5) == [1, 1, 1]
    with pytest.raises(ValueError):
        retained_block_sizes([3, 3], 6, 6)


@pytest.mark.parametrize("linear,precond", [("dense_schur", None), ("iterative_schur", "schur_jacobi"),
                                            ("iterative_schur", "jacobi")])
def test_quaternion_block_sparse_oracle_matches_dense_reference_algebra(linear, precond):
    """The block-sparse oracle (3x3 + 6x6 SCHUR_JACOBI blocks per quaternion camera) against the dense
    restatement of the reference's algebra on the full Jacobian, whose preconditioner blocks come from
    the parameter blocks' tangent sizes: same trajectory, PCG iteration counts included."""
    data = gen.with_quaternion_cameras(gen.ba_config("ba4"), "ceres", scale=1.2)
    # JACOBI's PCG runs grow past 20 iterations, where rounding differences between the two
    # restatements move the count; it is compared while the counts are still exact
    oo = OracleOptions(max_num_iterations=3 if precond == "jacobi" else 6, function_tolerance=0.0,
                       gradient_tolerance=0.0, parameter_tolerance=0.0)
    if precond:
        oo.preconditioner_type = precond
    a, (ac, ap) = trust_region_solve(oba.BAModel(data, linear_solver=linear), oo)
    b, (bc, bp) = trust_region_solve(oba.BAModel(data, linear_solver=linear, dense_faithful=True), oo)
    assert [it["linear_solver_iterations"] for it in a.iterations] == [it["linear_solver_iterations"]
                                                                      for it in b.iterations]
    assert [it["step_is_successful"] for it in a.iterations] == [it["step_is_successful"] for it in b.iterations]
    np.testing.assert_allclose([it["cost"] for it in a.iterations], [it["cost"] for it in b.iterations], rtol=1e-10)
    if precond != "jacobi":  # the last JACOBI step comes from an 18-iteration PCG run
        np.testing.assert_allclose(ac, bc, atol=1e-9)
    np.testing.assert_allclose(np.linalg.norm(ac[:, :4], axis=1), 1.0, atol=1e-14)


def test_quaternion_split_preconditioner_differs_from_one_nine_block():
    """The 3 + 6 split is not a no-op: on the same system a 9x9 block preconditioner takes a
    different number of PCG iterations (so the split is what the trajectories above pin)."""
    data = gen.with_quaternion_cameras(gen.ba_config("ba6_huber"), "ceres")
    model = oba.BAModel(data, linear_solver="iterative_schur")
    ev, _ = model.evaluate(model.x0(), True)
    st = model.struct
    Dp, Dc = oba.split_columns(st, 1e-4 * model.column_sq(ev))
    S, rhs, *_ = oba.schur_system(st, ev, Dp, Dc)
    _, it_split, _ = oba.block_jacobi_pcg(S, rhs, [3, 6] * st.num_cameras, tol=1e-12, max_iter=200)
    _, it_nine, _ = oba.block_jacobi_pcg(S, rhs, 9, tol=1e-12, max_iter=200)
    assert it_split != it_nine


def test_eigen_order_oracle_trajectory_equals_ceres_order():
    base = gen.ba_config("ba4")
    oo = OracleOptions(max_num_iterations=5, function_tolerance=0.0, gradient_tolerance=0.0, parameter_tolerance=0.0)
    a, (ac, _) = trust_region_solve(oba.BAModel(gen.with_quaternion_cameras(base, "ceres", 1.1)), oo)
    b, (bc, _) = trust_region_solve(oba.BAModel(gen.with_quaternion_cameras(base, "eigen", 1.1)), oo)
    np.testing.assert_allclose([it["cost"] for it in a.iterations], [it["cost"] for it in b.iterations], rtol=1e-12)
    np.testing.assert_allclose(oba.ceres_order(bc, "eigen_quaternion"), ac, atol=1e-12)


def test_quaternion_problem_extraction_pairs_blocks_and_refuses_wrong_manifolds():
    from ceres_mi355x.solver import _BAExtract

    data = gen.with_quaternion_cameras(gen.ba_config("ba4"), "ceres")
    problem = tc.Problem()
    rots = [torch.tensor(data.cameras[i, :4]) for i in range(data.num_cameras)]
    rests = [torch.tensor(data.cameras[i, 4:]) for i in range(data.num_cameras)]
    pts = [torch.tensor(data.points[i]) for i in range(data.num_points)]
    for r in rots:
        problem.AddParameterBlock(r, 4, tc.QuaternionManifold())
    for k in range(data.num_observations):
        c, p = int(data.obs_camera[k]), int(data.obs_point[k])
        problem.AddResidualBlock(tc.QuaternionReprojectionError(*data.observations[k]), None, [rots[c], rests[c], pts[p]])
    problem.SetParameterBlockConstant(rots[1])
    ex = _BAExtract(problem)
    assert ex.camera_model == "quaternion"
    cams, _ = ex.stacked()
    # cameras in order of first appearance, each the rotation block then the rest block
    index = [next(i for i, r in enumerate(rots) if r is c.rotation.tensor) for c in ex.cam_blocks]
    assert sorted(index) == list(range(data.num_cameras))
    assert all(ex.cam_blocks[k].rest.tensor is rests[i] for k, i in enumerate(index))
    np.testing.assert_array_equal(cams.numpy(), data.cameras[index])
    cm, _ = ex.masks()
    assert cm.shape == (data.num_cameras, 9)
    k1 = index.index(1)
    assert cm[k1, :3].sum() == 0 and cm[k1, 3:].all()
    assert all(cm[k].all() for k in range(data.num_cameras) if k != k1)
    # an Eigen-order cost on a Ceres-order manifold is refused
    p2 = tc.Problem()
    p2.AddParameterBlock(rots[0], 4, tc.QuaternionManifold())
    p2.AddResidualBlock(tc.QuaternionReprojectionError(1.0, 2.0, order="eigen"), None, [rots[0], rests[0], pts[0]])
    with pytest.raises(NotImplementedError, match="EigenQuaternionManifold"):
        _BAExtract(p2)
    # a rotation block without a quaternion manifold is refused
    p3 = tc.Problem()
    p3.AddResidualBlock(tc.QuaternionReprojectionError(1.0, 2.0), None, [rots[0].clone(), rests[0], pts[0]])
    with pytest.raises(NotImplementedError, match="QuaternionManifold"):
        _BAExtract(p3)


def test_constant_quaternion_camera_keeps_its_stored_values():
    """apply_delta (problem.py:462-474) moves only blocks with a tangent space: a constant camera's
    non-unit quaternion is not normalised by a step, while a free one is."""
    data = gen.with_quaternion_cameras(gen.ba_config("ba4"), "ceres", scale=1.5)
    active = np.array([True, False, True, True])
    model = oba.BAModel(data, camera_active=active)
    x = model.x0()
    step = np.full(model.num_active_columns(), 1e-3)
    cams, _ = model.apply(x, step)
    np.testing.assert_array_equal(cams[1], data.cameras[1])
    np.testing.assert_allclose(np.linalg.norm(cams[[0, 2, 3], :4], axis=1), 1.0, atol=1e-15)
