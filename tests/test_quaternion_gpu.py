"""Quaternion cameras on the MI355X against the float64 oracle.

A camera is two parameter blocks (rotation q with a QuaternionManifold or
EigenQuaternionManifold, then t, f, k1, k2) of QuaternionReprojectionError; the
device solves them as one 10-value camera with a 9-wide tangent.  The quaternions
start at non-unit norms (the Jacobian's 1/|q| and the first Plus's normalisation
are exercised).  Oracle: the block-sparse float64 restatement on the same inputs,
whose quaternion functions are pinned in tests/test_quaternion.py.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import ceres_mi355x as tc
from oracle import ba as oba
from oracle import generators as gen
from oracle.lm import OracleOptions, trust_region_solve

from test_oracle import assert_trajectory_matches

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _loop_mode_reset():
    from ceres_mi355x.ba import set_schur_operator

    yield
    tc.set_device_loop("device", 8)
    set_schur_operator("auto")


def _loss_obj(loss):
    if loss is None or loss[0] == "trivial":
        return None
    return {"huber": tc.HuberLoss, "cauchy": tc.CauchyLoss}[loss[0]](loss[1])


def quaternion_problem(data, order, constant_cameras=(), device="cuda"):
    """Per-block Problem: rotation blocks with their manifold, the 6-value camera blocks, points in
    ordering group 0 (as simple_bundle_adjuster.py:82-112 orders a BA problem)."""
    manifold = tc.QuaternionManifold if order == "ceres" else tc.EigenQuaternionManifold
    rots = [torch.tensor(data.cameras[i, :4], dtype=torch.float64, device=device) for i in range(data.num_cameras)]
    rests = [torch.tensor(data.cameras[i, 4:], dtype=torch.float64, device=device) for i in range(data.num_cameras)]
    pts = [torch.tensor(data.points[i], dtype=torch.float64, device=device) for i in range(data.num_points)]
    problem = tc.Problem()
    for r, c in zip(rots, rests):
        problem.AddParameterBlock(r, 4, manifold())
        problem.AddParameterBlock(c, 6)
    for p in pts:
        problem.AddParameterBlock(p, 3)
        problem.SetParameterBlockOrderingGroup(p, 0)
    loss = _loss_obj(data.loss)
    for k in range(data.num_observations):
        c, p = int(data.obs_camera[k]), int(data.obs_point[k])
        problem.AddResidualBlock(tc.QuaternionReprojectionError(*data.observations[k], order=order), loss,
                                 [rots[c], rests[c], pts[p]])
    for i in constant_cameras:
        problem.SetParameterBlockConstant(rots[i])
        problem.SetParameterBlockConstant(rests[i])
    return problem, rots, rests, pts


def _stacked(rots, rests):
    return np.stack([np.concatenate([r.cpu().numpy(), c.cpu().numpy()]) for r, c in zip(rots, rests)])


def _data(name="ba16", order="ceres", seed=None):
    base = gen.ba_config(name, seed)
    return gen.with_quaternion_cameras(base, order, scale=np.linspace(0.7, 1.6, base.num_cameras))


def _run(data, order, *, linear, loop, iters, precond=None, strategy=None, mixed=False, constant=()):
    tc.set_device_loop(loop)
    problem, rots, rests, pts = quaternion_problem(data, order, constant)
    o = tc.SolverOptions(max_num_iterations=iters, linear_solver_type=tc.LinearSolverType[linear.upper()])
    o.function_tolerance = o.gradient_tolerance = o.parameter_tolerance = 0.0
    oo = OracleOptions(max_num_iterations=iters, function_tolerance=0.0, gradient_tolerance=0.0,
                       parameter_tolerance=0.0)
    if precond:
        o.preconditioner_type = tc.PreconditionerType[precond.upper()]
        oo.preconditioner_type = precond
    if strategy:
        o.trust_region_strategy_type = tc.TrustRegionStrategyType[strategy.upper()]
        oo.trust_region_strategy_type = strategy
    if mixed:
        o.use_mixed_precision_solves = True
    summary = tc.solve(o, problem)
    print(f"termination {summary.termination_type} {summary.message!r}", flush=True)
    for it in summary.iterations:
        print(f"  it {it.iteration} valid {it.step_is_valid} ok {it.step_is_successful} cost {it.cost:.12e} "
              f"rho {it.relative_decrease:.4e} pcg {it.linear_solver_iterations} radius {it.trust_region_radius:.3e}",
              flush=True)
    active = np.ones(data.num_cameras, bool)
    active[list(constant)] = False
    ref, (rc, rp) = trust_region_solve(oba.BAModel(data, linear_solver=linear, camera_active=active), oo)
    return summary, ref, rc, rp, _stacked(rots, rests), torch.stack(pts).cpu().numpy()


@pytest.mark.parametrize("loop", ["host", "device"])
@pytest.mark.parametrize("order", ["ceres", "eigen"])
def test_quaternion_ba_dense_schur_matches_oracle_trajectory(order, loop):
    print(f"quaternion ba16 dense_schur order={order} loop={loop}", flush=True)
    data = _data(order=order)
    s, ref, rc, rp, cams, pts = _run(data, order, linear="dense_schur", loop=loop, iters=10)
    assert_trajectory_matches(s, {"iterations": ref.iterations, "final_cost": ref.final_cost,
                                  "initial_cost": ref.initial_cost})
    np.testing.assert_allclose(cams, rc, atol=1e-7)
    np.testing.assert_allclose(pts, rp, atol=1e-6)
    q = cams[:, :4]
    np.testing.assert_allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-14)


@pytest.mark.parametrize("schur", ["auto", "explicit"])
@pytest.mark.parametrize("loop", ["host", "device"])
def test_quaternion_ba_iterative_schur_split_blocks_match_oracle(loop, schur):
    """SCHUR_JACOBI with a 3x3 and a 6x6 block per camera (the reference's parameter blocks): the PCG
    iteration counts are compared exactly, so a 9x9 block preconditioner would fail."""
    from ceres_mi355x.ba import set_schur_operator

    from ceres_mi355x.ba import debug_fill_lds

    print(f"quaternion ba16 iterative_schur loop={loop} schur={schur}", flush=True)
    set_schur_operator(schur)
    data = _data()
    with debug_fill_lds():  # a solver may not depend on what the LDS held before it
        s, ref, rc, rp, cams, pts = _run(data, "ceres", linear="iterative_schur", loop=loop, iters=10,
                                         precond="schur_jacobi")
    assert_trajectory_matches(s, {"iterations": ref.iterations, "final_cost": ref.final_cost,
                                  "initial_cost": ref.initial_cost})
    np.testing.assert_allclose(pts, rp, atol=1e-6)


def test_quaternion_ba_huber_dogleg_with_constant_cameras_matches_oracle():
    print("quaternion ba6_huber dogleg", flush=True)
    data = _data("ba6_huber")
    s, ref, rc, rp, cams, pts = _run(data, "ceres", linear="dense_schur", loop="device", iters=12,
                                     strategy="dogleg", constant=(0, 3))
    assert_trajectory_matches(s, {"iterations": ref.iterations, "final_cost": ref.final_cost,
                                  "initial_cost": ref.initial_cost})
    np.testing.assert_allclose(cams, rc, atol=1e-7)
    np.testing.assert_array_equal(cams[[0, 3]], data.cameras[[0, 3]])


def test_quaternion_ba_mixed_precision_reaches_oracle_cost():
    print("quaternion ba16_huber mixed", flush=True)
    data = _data("ba16_huber")
    s, ref, *_ = _run(data, "ceres", linear="dense_schur", loop="device", iters=25, mixed=True)
    assert abs(s.final_cost - ref.final_cost) <= 1e-6 * ref.final_cost


def test_quaternion_evaluate_jacobian_matches_oracle():
    """Problem.evaluate(compute_jacobian=True): tangent columns of the rotation (3), the camera block
    (6) and the point (3) per residual block, with one rotation held constant."""
    print("quaternion ba4 evaluate", flush=True)
    data = _data("ba4")
    problem, rots, rests, pts = quaternion_problem(data, "ceres")
    problem.SetParameterBlockConstant(rots[2])
    out = problem.evaluate(compute_jacobian=True)
    st = oba.BAStructure.build(data.obs_camera, data.obs_point, data.num_cameras, data.num_points)
    ev = oba.ba_evaluate(st, data.cameras, data.points, data.observations, data.loss, True, "quaternion")
    assert float(out.cost) == pytest.approx(ev.cost, rel=1e-12)
    J = out.jacobian.cpu().numpy()
    # column layout: blocks in insertion order (rotation, camera per camera, then points), constant skipped
    col, offsets = 0, {}
    for i in range(data.num_cameras):
        if i != 2:
            offsets[("q", i)] = col
            col += 3
        offsets[("c", i)] = col
        col += 6
    for p in range(data.num_points):
        offsets[("p", p)] = col
        col += 3
    assert J.shape == (2 * data.num_observations, col)
    for k in range(0, data.num_observations, 7):
        c, p = int(data.obs_camera[k]), int(data.obs_point[k])
        rows = J[2 * k:2 * k + 2]
        if c != 2:
            np.testing.assert_allclose(rows[:, offsets[("q", c)]:offsets[("q", c)] + 3], ev.Jc[k, :, :3],
                                       rtol=1e-10, atol=1e-8)
        np.testing.assert_allclose(rows[:, offsets[("c", c)]:offsets[("c", c)] + 6], ev.Jc[k, :, 3:],
                                   rtol=1e-10, atol=1e-8)
        np.testing.assert_allclose(rows[:, offsets[("p", p)]:offsets[("p", p)] + 3], ev.Jp[k], rtol=1e-10, atol=1e-8)
