"""GPU parity of the bundle-adjustment LM solve against the reference's own
trajectories (tests/golden, written by running the reference) and the oracle.

Tolerances (written here, per the plain FP64 contract): the accept/reject
sequence and per-iteration costs match until the cost has converged (1e-9
relative per iteration), and the final cost matches to 1e-10 relative for
trivial-loss problems; 1e-7 for the Huber/outlier case whose 50-iteration cap
falls mid-descent (see tests/test_oracle.py).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import ceres_mi355x as tc
from oracle import ba as oba
from oracle import generators as gen
from oracle.lm import OracleOptions, trust_region_solve

from conftest import GOLDEN, load_golden
from test_oracle import assert_trajectory_matches

pytestmark = pytest.mark.gpu


def _loss_obj(loss):
    if loss is None or loss[0] == "trivial":
        return None
    return {"huber": tc.HuberLoss, "cauchy": tc.CauchyLoss}[loss[0]](loss[1])


def ba_problem(data, device="cuda"):
    cams = torch.tensor(data.cameras, dtype=torch.float64, device=device)
    pts = torch.tensor(data.points, dtype=torch.float64, device=device)
    problem = tc.Problem()
    problem.add_bundle_adjustment_observations(cams, pts, data.obs_camera, data.obs_point, data.observations,
                                               loss=_loss_obj(data.loss))
    return problem, cams, pts


def options_for(golden, **kw):
    o = tc.SolverOptions(max_num_iterations=golden["max_num_iterations"],
                         linear_solver_type=tc.LinearSolverType[golden["linear_solver"].upper()], **kw)
    if golden.get("tolerances") == "zero":
        o.function_tolerance = o.gradient_tolerance = o.parameter_tolerance = 0.0
    if golden["linear_solver"] == "iterative_schur":
        o.preconditioner_type = tc.PreconditionerType.SCHUR_JACOBI
    return o


@pytest.fixture(autouse=True)
def _loop_mode_reset():
    yield
    tc.set_device_loop("device", 8)


@pytest.mark.parametrize("loop", ["host", "device"])
@pytest.mark.parametrize("case", ["ba4_dense_schur_zero_tol", "ba4_dense_schur_default_tol", "ba6_huber_dense_schur",
                                  "ba4_iterative_schur_zero_tol"])
def test_ba_matches_reference_trajectory(case, loop):
    print(f"case={case} loop={loop}", flush=True)
    tc.set_device_loop(loop)
    g = load_golden(case)
    data = gen.ba_config(g["config"])
    problem, cams, pts = ba_problem(data)
    summary = tc.solve(options_for(g), problem)
    final_rtol = 1e-7 if "huber" in case else 1e-10
    assert_trajectory_matches(summary, g["summary"], final_rtol=final_rtol)
    if g["tolerances"] == "default":
        assert summary.termination_type.value == g["summary"]["termination_type"]
        assert summary.message == g["summary"]["message"]
        assert len(summary.iterations) == len(g["summary"]["iterations"])
        assert summary.num_successful_steps == g["summary"]["num_successful_steps"]
        assert summary.num_jacobian_evaluations == g["summary"]["num_jacobian_evaluations"]
        assert summary.num_residual_evaluations == g["summary"]["num_residual_evaluations"]
    if "huber" not in case:
        np.testing.assert_allclose(pts.cpu().numpy(), np.array(g["final_points"]), atol=1e-6)


@pytest.mark.parametrize("loop", ["host", "device"])
def test_ba16_matches_reference_trajectory(loop):
    print(f"ba16 loop={loop}", flush=True)
    tc.set_device_loop(loop)
    g = load_golden("ba16_dense_schur_zero_tol")
    data = gen.ba_config("ba16")
    problem, cams, pts = ba_problem(data)
    summary = tc.solve(options_for(g), problem)
    assert_trajectory_matches(summary, g["summary"])
    ref = np.load(GOLDEN / "ba16_dense_schur_zero_tol_final.npz")
    np.testing.assert_allclose(pts.cpu().numpy(), ref["points"], atol=1e-6)


def test_per_block_api_matches_bulk_api():
    g = load_golden("ba4_dense_schur_zero_tol")
    data = gen.ba_config("ba4")
    cams = torch.tensor(data.cameras, dtype=torch.float64)
    pts = torch.tensor(data.points, dtype=torch.float64)
    problem = tc.Problem()
    cam_blocks = [cams[i] for i in range(cams.shape[0])]
    pt_blocks = [pts[i] for i in range(pts.shape[0])]
    for c in cam_blocks:
        problem.AddParameterBlock(c)
    for p in pt_blocks:
        problem.AddParameterBlock(p)
        problem.SetParameterBlockOrderingGroup(p, 0)
    for k in range(data.num_observations):
        problem.AddResidualBlock(tc.SnavelyReprojectionError(*data.observations[k]), None,
                                 [cam_blocks[int(data.obs_camera[k])], pt_blocks[int(data.obs_point[k])]])
    summary = tc.solve(options_for(g), problem)
    assert_trajectory_matches(summary, g["summary"])
    np.testing.assert_allclose(pts.numpy(), np.array(g["final_points"]), atol=1e-6)


@pytest.mark.parametrize("loop", ["host", "device"])
@pytest.mark.parametrize("case", ["ba4_dogleg_zero_tol", "ba4_subspace_dogleg_zero_tol", "ba4_dogleg_default_tol"])
def test_ba_dogleg_matches_reference_trajectory(case, loop):
    """DOGLEG (traditional and subspace), gauge fixed by two constant cameras.

    The small initial radius (0.05) makes the first steps the scaled Cauchy
    point, then points on the dogleg path / the subspace boundary, then full
    Gauss-Newton steps.  Same bar as the LM trajectories.
    """
    print(f"case={case} loop={loop}", flush=True)
    tc.set_device_loop(loop)
    g = load_golden(case)
    data = gen.ba_config(g["config"])
    problem, cams, pts = ba_problem(data)
    for i in g["constant_cameras"]:
        problem.SetParameterBlockConstant(problem.parameter_blocks[i])
    o = options_for(g, trust_region_strategy_type=tc.TrustRegionStrategyType.DOGLEG,
                    dogleg_type=(tc.DoglegType.SUBSPACE_DOGLEG if g["dogleg_type"] == "subspace"
                                 else tc.DoglegType.TRADITIONAL_DOGLEG),
                    initial_trust_region_radius=g["initial_trust_region_radius"])
    cams0 = cams.clone()
    summary = tc.solve(o, problem)
    assert_trajectory_matches(summary, g["summary"])
    assert summary.num_linear_solves == 0
    if g["tolerances"] == "default":
        assert summary.termination_type.value == g["summary"]["termination_type"]
        assert summary.message == g["summary"]["message"]
        assert len(summary.iterations) == len(g["summary"]["iterations"])
    np.testing.assert_allclose(pts.cpu().numpy(), np.array(g["final_points"]), atol=1e-6)
    torch.testing.assert_close(cams[:2], cams0[:2], rtol=0, atol=0)


def test_tiny_bal_with_constant_cameras():
    # examples/simple_bundle_adjuster.py:142-176; tests/test_examples.py:151-165
    true_cams = np.array([[0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 800.0, 0.0, 0.0],
                          [0.0, 0.02, 0.0, -0.4, 0.0, 0.05, 800.0, 0.0, 0.0]])
    true_pts = np.array([[0.2, -0.1, 4.0], [-0.3, 0.15, 4.5], [0.1, 0.2, 5.0]])
    init_pts = true_pts + np.array([[0.04, -0.03, 0.15], [-0.05, 0.02, -0.12], [0.03, 0.04, 0.10]])
    oc = np.array([0, 0, 0, 1, 1, 1])
    op = np.array([0, 1, 2, 0, 1, 2])
    obs = oba.project(true_cams[oc], true_pts[op])
    cams = torch.tensor(true_cams)
    pts = torch.tensor(init_pts)
    problem = tc.Problem()
    problem.add_bundle_adjustment_observations(cams, pts, oc, op, obs)
    for i in range(2):
        problem.SetParameterBlockConstant(problem.parameter_blocks[i])
    summary = tc.solve(tc.SolverOptions(max_num_iterations=40, linear_solver_type=tc.LinearSolverType.DENSE_SCHUR,
                                        gradient_tolerance=1e-12, function_tolerance=1e-12,
                                        parameter_tolerance=1e-12), problem)
    assert summary.IsSolutionUsable()
    assert summary.final_cost < 1e-12
    np.testing.assert_allclose(pts.numpy(), true_pts, atol=1e-5, rtol=1e-5)
    np.testing.assert_array_equal(cams.numpy(), true_cams)


def test_tiny_bal_file_solve_recovers_points(tmp_path):
    # tests/test_examples.py:151-165: the example's BAL file, cameras fixed
    from ceres_mi355x.bal import BALProblem, build_problem
    from test_bal import tiny_bal_text

    text, true_cams, true_pts, _ = tiny_bal_text()
    path = tmp_path / "tiny.bal"
    path.write_text(text)
    bal = BALProblem.from_file(path)
    problem, cams, pts = build_problem(bal, fix_cameras=True, device="cuda")
    summary = tc.solve(tc.SolverOptions(max_num_iterations=40, linear_solver_type=tc.LinearSolverType.DENSE_SCHUR,
                                        gradient_tolerance=1e-12, function_tolerance=1e-12,
                                        parameter_tolerance=1e-12), problem)
    assert summary.IsSolutionUsable()
    assert summary.final_cost < 1e-12
    np.testing.assert_allclose(pts.cpu().numpy(), true_pts, atol=1e-5, rtol=1e-5)
    np.testing.assert_array_equal(cams.cpu().numpy(), true_cams)


def test_evaluate_matches_oracle_cost_and_residuals():
    data = gen.ba_config("ba6_huber")
    problem, cams, pts = ba_problem(data)
    res = problem.evaluate()
    m = oba.BAModel(data)
    ev, _ = m.evaluate(m.x0(), False)
    assert float(res.cost) == pytest.approx(ev.cost, rel=1e-13)
    np.testing.assert_allclose(res.residuals.cpu().numpy(), ev.r.reshape(-1), rtol=1e-12, atol=1e-9)


def test_mixed_precision_reaches_fp64_final_cost():
    """Mixed form (FP32 J / J^T J / Schur, FP64 Cholesky, cost, gain ratio, 2 refinement steps) vs plain FP64.

    Both are run to convergence (function tolerance 0: the solve ends at the
    minimum trust-region radius); the Huber problem with outliers descends
    slowly, so a looser tolerance stops the two runs at different points of
    the same descent.  Bar: final cost within 1e-6 relative.
    """
    data = gen.ba_config("ba16_huber")
    o = tc.SolverOptions(max_num_iterations=2000, linear_solver_type=tc.LinearSolverType.DENSE_SCHUR,
                         function_tolerance=0.0, gradient_tolerance=0.0, parameter_tolerance=0.0)
    problem, _, _ = ba_problem(data)
    s64 = tc.solve(o, problem)
    o2 = tc.SolverOptions(**{**o.__dict__, "use_mixed_precision_solves": True, "max_num_refinement_iterations": 2})
    problem2, _, _ = ba_problem(data)
    s32 = tc.solve(o2, problem2)
    assert s64.termination_type == tc.TerminationType.CONVERGENCE
    assert s32.termination_type == tc.TerminationType.CONVERGENCE
    assert s32.final_cost == pytest.approx(s64.final_cost, rel=1e-6)
    assert s32.precision == "mixed"
