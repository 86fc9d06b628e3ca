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
from test_oracle import assert_trajectory_matches, long_pcg_bound

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
        o.preconditioner_type = tc.PreconditionerType[(golden.get("preconditioner") or "schur_jacobi").upper()]
    return o


@pytest.fixture(autouse=True)
def _loop_mode_reset():
    from ceres_mi355x.ba import set_schur_operator

    yield
    tc.set_device_loop("device", 8)
    set_schur_operator("auto")


@pytest.mark.parametrize("loop", ["host", "device"])
@pytest.mark.parametrize("case", ["ba4_dense_schur_zero_tol", "ba4_dense_schur_default_tol", "ba6_huber_dense_schur",
                                  "ba4_iterative_schur_zero_tol", "ba4_iterative_jacobi_zero_tol",
                                  "ba4_iterative_identity_zero_tol"])
def test_ba_matches_reference_trajectory(case, loop):
    print(f"case={case} loop={loop}", flush=True)
    tc.set_device_loop(loop)
    _check_reference_trajectory(case)


@pytest.mark.parametrize("case", ["ba4_iterative_schur_zero_tol", "ba4_iterative_jacobi_zero_tol",
                                  "ba4_iterative_identity_zero_tol"])
def test_ba_explicit_schur_operator_matches_reference_trajectory(case):
    """The explicitly assembled S (CT_SCHUR_EXPLICIT) on the ITERATIVE_SCHUR goldens; the default applies S
    from the records."""
    from ceres_mi355x.ba import set_schur_operator

    print(f"case={case} explicit S", flush=True)
    set_schur_operator("explicit")
    _check_reference_trajectory(case)


@pytest.mark.parametrize("loop", ["host", "device"])
@pytest.mark.parametrize("case,schur", [("ba4_iterative_schur_zero_tol", "auto"),
                                        ("ba4_iterative_schur_zero_tol", "explicit"),
                                        ("ba4_iterative_jacobi_zero_tol", "auto"),
                                        ("ba4_dense_schur_zero_tol", "auto"),
                                        ("ba6_huber_dense_schur", "auto")])
def test_solvers_read_no_lds_they_did_not_write(case, schur, loop):
    """With every compute unit's LDS filled with NaN right before each reduced-camera solve (the
    debug_fill_lds hook), the reference trajectories still hold: no PCG or Cholesky kernel reads
    LDS it has not written in its own launch.  (The implicit PCG once formed its first direction as
    z + 0 * p from an unwritten p, which left NaN behind another kernel's LDS as NaN.)"""
    from ceres_mi355x.ba import debug_fill_lds, set_schur_operator

    print(f"case={case} schur={schur} loop={loop} LDS filled with NaN", flush=True)
    tc.set_device_loop(loop)
    set_schur_operator(schur)
    with debug_fill_lds():
        _check_reference_trajectory(case)


def _check_reference_trajectory(case):
    g = load_golden(case)
    data = gen.ba_config(g["config"])
    problem, cams, pts = ba_problem(data)
    summary = tc.solve(options_for(g), problem)
    final_rtol = 1e-7 if "huber" in case else 1e-10
    assert_trajectory_matches(summary, g["summary"], final_rtol=final_rtol, max_exact_pcg=long_pcg_bound(g))
    if g["tolerances"] == "default":
        assert summary.termination_type.value == g["summary"]["termination_type"]
        assert summary.message == g["summary"]["message"]
        assert len(summary.iterations) == len(g["summary"]["iterations"])
        assert summary.num_successful_steps == g["summary"]["num_successful_steps"]
        assert summary.num_jacobian_evaluations == g["summary"]["num_jacobian_evaluations"]
        assert summary.num_residual_evaluations == g["summary"]["num_residual_evaluations"]
    if "huber" not in case and long_pcg_bound(g) is None:
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


@pytest.mark.parametrize("loop", ["host", "device"])
def test_max_solver_time_ends_the_solve_like_the_reference(loop):
    # solver.py:342-345: the time test follows the radius test of every iteration with a step; with
    # a zero budget the solve ends after the first step, NO_CONVERGENCE (same record count as the oracle)
    tc.set_device_loop(loop)
    data = gen.ba_config("ba4")
    problem, _, _ = ba_problem(data)
    o = tc.SolverOptions(max_num_iterations=25, linear_solver_type=tc.LinearSolverType.DENSE_SCHUR,
                         max_solver_time_in_seconds=0.0)
    summary = tc.solve(o, problem)
    ref, _ = trust_region_solve(oba.BAModel(data), OracleOptions(max_num_iterations=25, max_solver_time_in_seconds=0.0))
    assert summary.termination_type == tc.TerminationType.NO_CONVERGENCE
    assert summary.message == ref.message == "Maximum solver time reached."
    assert len(summary.iterations) == len(ref.iterations) == 2
    assert summary.final_cost == pytest.approx(ref.final_cost, rel=1e-10)


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


def test_evaluate_with_jacobian_matches_oracle():
    """Problem.evaluate(compute_jacobian=True) (problem.py:377-413): columns of the active tangent
    coordinates in problem order (cameras, then points, as registered), a constant camera and a
    subset manifold dropping columns; gradient J^T r; the CRS copy; apply_loss_function=False;
    an explicit parameter-block selection.  Oracle: the float64 evaluation of the same inputs."""
    data = gen.ba_config("ba6_huber")
    problem, cams, pts = ba_problem(data)
    C, P, N = data.num_cameras, data.num_points, data.num_observations
    problem.SetParameterBlockConstant(problem.parameter_blocks[0])
    problem.set_manifold(problem.parameter_blocks[1], tc.SubsetManifold(9, [6, 7, 8]))
    m = oba.BAModel(data)
    ev, _ = m.evaluate(m.x0(), True)
    cam_cols = [None] + [list(range(6))] + [list(range(9))] * (C - 2)
    cols = []  # (camera or point, index, coordinate) per tangent column
    for c in range(1, C):
        cols += [("c", c, j) for j in cam_cols[c]]
    cols += [("p", p, j) for p in range(P) for j in range(3)]
    J = np.zeros((2 * N, len(cols)))
    for k, (kind, idx, j) in enumerate(cols):
        sel = (data.obs_camera == idx) if kind == "c" else (data.obs_point == idx)
        blk = ev.Jc if kind == "c" else ev.Jp
        for i in range(2):
            J[2 * np.nonzero(sel)[0] + i, k] = blk[sel, i, j]
    r = ev.r.reshape(-1)
    res = problem.evaluate()
    assert float(res.cost) == pytest.approx(ev.cost, rel=1e-13)
    np.testing.assert_allclose(res.residuals.cpu().numpy(), r, rtol=1e-12, atol=1e-9)
    scale = np.abs(J).max()
    np.testing.assert_allclose(res.jacobian.cpu().numpy(), J, rtol=1e-9, atol=1e-11 * scale)
    np.testing.assert_allclose(res.gradient.cpu().numpy(), J.T @ r, rtol=1e-8, atol=1e-9 * np.abs(J.T @ r).max())
    torch.testing.assert_close(res.crs_jacobian.to_dense(), res.jacobian.cpu(), rtol=0, atol=0)
    assert res.crs_jacobian.rows[-1] == int((res.jacobian != 0).sum())
    # no loss: raw residuals, cost 1/2 |r|^2
    raw = oba.project(data.cameras[data.obs_camera], data.points[data.obs_point]) - data.observations
    res2 = problem.evaluate(tc.EvaluateOptions(apply_loss_function=False), compute_jacobian=False)
    np.testing.assert_allclose(res2.residuals.cpu().numpy(), raw.reshape(-1), rtol=1e-12, atol=1e-9)
    assert float(res2.cost) == pytest.approx(0.5 * float(np.sum(raw * raw)), rel=1e-13)
    # the points alone, in reverse order
    sel = [problem.parameter_blocks[C + p] for p in reversed(range(P))]
    res3 = problem.evaluate(tc.EvaluateOptions(parameter_blocks=sel))
    perm = [len(cols) - 3 * P + 3 * p + j for p in reversed(range(P)) for j in range(3)]
    np.testing.assert_allclose(res3.jacobian.cpu().numpy(), J[:, perm], rtol=1e-9, atol=1e-11 * scale)


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
    # against the FP64 oracle's own converged cost (tests/golden/make_oracle_fixtures.py)
    g = load_golden("oracle_ba16_huber_converged")
    assert g["input_sha256"] == _digest(data)
    assert s32.final_cost == pytest.approx(g["summary"]["final_cost"], rel=1e-6)
    assert s64.final_cost == pytest.approx(g["summary"]["final_cost"], rel=1e-9)


def _digest(data):
    from conftest import array_digest

    return array_digest(data.cameras, data.points, data.observations, data.obs_camera, data.obs_point)


@pytest.fixture(scope="module")
def ba200_data():
    return gen.ba_config("ba200")


@pytest.mark.parametrize("loop,linear,schur", [("host", "iterative_schur", "auto"),
                                               ("device", "iterative_schur", "auto"),
                                               ("device", "iterative_schur", "explicit"),
                                               ("host", "dense_schur", "auto"),
                                               ("device", "dense_schur", "auto")])
def test_ba200_matches_oracle_trajectory(ba200_data, linear, loop, schur):
    """The 200-camera workload (100,000 points, ~650k observations) at its full size: the first 5 LM
    iterations against the float64 oracle's (tests/golden/make_oracle_fixtures.py), with the same
    bar as the reference-recorded trajectories: accept/reject sequence, costs to 1e-9, PCG counts
    exactly, final cost to 1e-10; the final cameras to 1e-7."""
    from ceres_mi355x.ba import set_schur_operator

    print(f"ba200 linear={linear} loop={loop} schur={schur}", flush=True)
    tc.set_device_loop(loop)
    set_schur_operator(schur)
    g = load_golden(f"oracle_ba200_{linear}_5")
    data = ba200_data
    assert g["input_sha256"] == _digest(data)
    problem, cams, pts = ba_problem(data)
    summary = tc.solve(options_for(g), problem)
    assert_trajectory_matches(summary, g["summary"])
    assert summary.num_linear_solves == g["summary"]["num_linear_solves"]
    np.testing.assert_allclose(cams.cpu().numpy(), np.array(g["final_cameras"]), rtol=1e-7, atol=1e-7)
    np.testing.assert_allclose(pts[:200].cpu().numpy(), np.array(g["final_points_head"]), atol=1e-7)


def test_ba200_mixed_precision_reaches_oracle_cost(ba200_data):
    """The mixed form on the 200-camera workload, DENSE_SCHUR: after 8 iterations its cost is within
    1e-6 of the FP64 oracle's (which has converged to 1e-13 within its first 5)."""
    g = load_golden("oracle_ba200_dense_schur_5")
    problem, _, _ = ba_problem(ba200_data)
    o = tc.SolverOptions(max_num_iterations=8, linear_solver_type=tc.LinearSolverType.DENSE_SCHUR,
                         use_mixed_precision_solves=True, max_num_refinement_iterations=2,
                         function_tolerance=0.0, gradient_tolerance=0.0, parameter_tolerance=0.0)
    s = tc.solve(o, problem)
    assert s.precision == "mixed"
    assert s.final_cost == pytest.approx(g["summary"]["final_cost"], rel=1e-6)


@pytest.mark.parametrize("loop", ["host", "device"])
def test_largest_served_ba_request_matches_oracle(loop):
    """The largest bundle adjustment of the served mix (20 cameras, 5,000 points, 2-6 observations per
    point) against the oracle's trajectory."""
    print(f"ba20x5000 loop={loop}", flush=True)
    tc.set_device_loop(loop)
    data = gen.make_bundle_adjustment(num_cameras=20, num_points=5000, obs_per_point=(2, 6), seed=41)
    problem, cams, pts = ba_problem(data)
    o = tc.SolverOptions(max_num_iterations=12, linear_solver_type=tc.LinearSolverType.DENSE_SCHUR)
    o.function_tolerance = o.gradient_tolerance = o.parameter_tolerance = 0.0
    summary = tc.solve(o, problem)
    oo = OracleOptions(max_num_iterations=12, function_tolerance=0.0, gradient_tolerance=0.0, parameter_tolerance=0.0)
    ref, (rc, rp) = trust_region_solve(oba.BAModel(data), oo)
    assert_trajectory_matches(summary, {"iterations": ref.iterations, "final_cost": ref.final_cost,
                                        "initial_cost": ref.initial_cost})
    np.testing.assert_allclose(pts.cpu().numpy(), rp, atol=1e-6)


def test_plan_rejects_camera_pair_blocks_out_of_row_major_order():
    """The assembly indexes S's blocks by camera pair, so the C-ABI refuses a block list that is not
    every pair i <= j in row-major order (here the same pairs reversed, and one pair missing)."""
    from ceres_mi355x import ba as nba
    from ceres_mi355x._native import NativeLibraryError
    data = gen.ba_config("ba4")
    cams = torch.tensor(data.cameras, dtype=torch.float64, device="cuda")
    pts = torch.tensor(data.points, dtype=torch.float64, device="cuda")
    opts = nba.lm_options(tc.SolverOptions(max_num_iterations=2))
    dev = nba.BADevice(cams, pts, data.obs_camera, data.obs_point, data.observations)
    good = dev.block_cameras
    dev.block_cameras = good.flip(0).contiguous()
    with pytest.raises(NativeLibraryError, match="row-major order"):
        dev.run(opts)
    dev.block_cameras = good[:-1].contiguous()
    dev.block_offsets = dev.block_offsets[:-1].contiguous()
    with pytest.raises(NativeLibraryError, match=r"C \(C \+ 1\) / 2"):
        dev.run(opts)
    dev.close()


@pytest.mark.parametrize("loop", ["host", "device"])
@pytest.mark.parametrize("linear", ["dense_schur", "iterative_schur"])
def test_points_with_more_observations_than_a_wavefront_match_oracle(linear, loop):
    """A point observed by more than 64 cameras cannot be one lane per observation of a wavefront, so
    the plan evaluates and steps per observation and per point (k_eval_obs + k_eval_points,
    k_step_points + k_candidate_obs) instead of by point groups.  Oracle: the float64 block-sparse
    restatement on the same inputs, same tolerances as the golden cases."""
    print(f"wide points linear={linear} loop={loop}", flush=True)
    tc.set_device_loop(loop)
    data = gen.make_bundle_adjustment(num_cameras=72, num_points=80, obs_per_point=(3, 72), seed=21)
    assert np.bincount(data.obs_point).max() > 64
    problem, cams, pts = ba_problem(data)
    o = tc.SolverOptions(max_num_iterations=10, linear_solver_type=tc.LinearSolverType[linear.upper()])
    oo = OracleOptions(max_num_iterations=10)
    if linear == "iterative_schur":
        o.preconditioner_type = tc.PreconditionerType.SCHUR_JACOBI
        oo.preconditioner_type = "schur_jacobi"
    summary = tc.solve(o, problem)
    ref, (rc, rp) = trust_region_solve(oba.BAModel(data, linear_solver=linear), oo)
    assert_trajectory_matches(summary, {"iterations": ref.iterations, "final_cost": ref.final_cost,
                                        "initial_cost": ref.initial_cost})
    np.testing.assert_allclose(pts.cpu().numpy(), rp, atol=1e-6)


def test_points_with_more_observations_than_a_wavefront_mixed_precision():
    """The per-observation kernels in the mixed form: final cost within 1e-6 of the FP64 oracle's."""
    data = gen.make_bundle_adjustment(num_cameras=72, num_points=80, obs_per_point=(3, 72), seed=21)
    problem, cams, pts = ba_problem(data)
    o = tc.SolverOptions(max_num_iterations=20, linear_solver_type=tc.LinearSolverType.DENSE_SCHUR,
                         use_mixed_precision_solves=True, max_num_refinement_iterations=2)
    summary = tc.solve(o, problem)
    ref, _ = trust_region_solve(oba.BAModel(data), OracleOptions(max_num_iterations=20))
    assert summary.final_cost == pytest.approx(ref.final_cost, rel=1e-6)


def _sparse_visibility_problem():
    """Points seen once (rank-2 V_p without damping), points seen by no camera and a camera that sees
    nothing: their columns are zero, so only the damping keeps the system nonsingular."""
    data = gen.make_bundle_adjustment(num_cameras=6, num_points=90, obs_per_point=(1, 4), seed=23)
    keep = data.obs_camera != 5
    data.obs_camera, data.obs_point, data.observations = (data.obs_camera[keep], data.obs_point[keep],
                                                          data.observations[keep])
    counts = np.bincount(data.obs_point, minlength=data.points.shape[0])
    assert (counts == 1).any() and (counts == 0).any()
    return data


@pytest.mark.parametrize("loop", ["host", "device"])
@pytest.mark.parametrize("linear", ["dense_schur", "iterative_schur"])
def test_unobserved_blocks_and_single_observation_points_match_oracle(linear, loop):
    print(f"sparse visibility linear={linear} loop={loop}", flush=True)
    tc.set_device_loop(loop)
    data = _sparse_visibility_problem()
    problem, cams, pts = ba_problem(data)
    o = tc.SolverOptions(max_num_iterations=10, linear_solver_type=tc.LinearSolverType[linear.upper()])
    oo = OracleOptions(max_num_iterations=10)
    if linear == "iterative_schur":
        o.preconditioner_type = tc.PreconditionerType.SCHUR_JACOBI
        oo.preconditioner_type = "schur_jacobi"
    summary = tc.solve(o, problem)
    ref, (rc, rp) = trust_region_solve(oba.BAModel(data, linear_solver=linear), oo)
    assert_trajectory_matches(summary, {"iterations": ref.iterations, "final_cost": ref.final_cost,
                                        "initial_cost": ref.initial_cost})
    np.testing.assert_allclose(pts.cpu().numpy(), rp, atol=1e-6)
    np.testing.assert_allclose(cams.cpu().numpy(), rc, atol=1e-6)
    unseen = np.bincount(data.obs_point, minlength=data.points.shape[0]) == 0
    np.testing.assert_array_equal(pts.cpu().numpy()[unseen], data.points[unseen])
    np.testing.assert_array_equal(cams.cpu().numpy()[5], data.cameras[5])


@pytest.mark.parametrize("loop", ["host", "device"])
def test_invalid_steps_end_in_failure_like_the_reference(loop):
    """The reference's invalid-step test (tests/test_problem_solver.py:637-664)
    on a bundle-adjustment problem: every step is overwritten with NaN on the
    device; two in a row end the solve with FAILURE, the failing iteration is
    not recorded, the radius goes 4 -> 2 and the parameters stay unchanged."""
    print(f"invalid steps loop={loop}", flush=True)
    from ceres_mi355x.ba import debug_invalid_steps

    tc.set_device_loop(loop)
    data = gen.ba_config("ba4")
    problem, cams, pts = ba_problem(data)
    cams0, pts0 = cams.clone(), pts.clone()
    o = tc.SolverOptions(max_num_iterations=3, initial_trust_region_radius=4.0, max_num_consecutive_invalid_steps=2,
                         gradient_tolerance=0.0, linear_solver_type=tc.LinearSolverType.DENSE_SCHUR)
    with debug_invalid_steps(100):
        summary = tc.solve(o, problem)
    assert summary.termination_type is tc.TerminationType.FAILURE
    assert summary.message == "Too many invalid trust-region steps."
    assert summary.num_unsuccessful_steps == 2
    assert len(summary.iterations) == 2
    assert summary.iterations[1].step_is_valid is False
    assert summary.iterations[1].trust_region_radius == pytest.approx(2.0)
    assert summary.final_cost == summary.initial_cost
    assert torch.equal(cams, cams0) and torch.equal(pts, pts0)


@pytest.mark.parametrize("loop", ["host", "device"])
@pytest.mark.parametrize("linear", ["dense_schur", "iterative_schur"])
def test_one_invalid_step_then_the_oracle_trajectory(loop, linear):
    """Two NaN steps, then the solve continues: the whole trajectory, counters
    included, matches the oracle whose first two linear solves return NaN."""
    print(f"two invalid steps loop={loop} linear={linear}", flush=True)
    from ceres_mi355x.ba import debug_invalid_steps
    from test_oracle import PoisonedSolves

    tc.set_device_loop(loop)
    data = gen.ba_config("ba4")
    problem, cams, pts = ba_problem(data)
    o = tc.SolverOptions(max_num_iterations=12, linear_solver_type=tc.LinearSolverType[linear.upper()],
                         preconditioner_type=tc.PreconditionerType.SCHUR_JACOBI)
    o.function_tolerance = o.gradient_tolerance = o.parameter_tolerance = 0.0
    with debug_invalid_steps(2):
        summary = tc.solve(o, problem)
    oo = OracleOptions(max_num_iterations=12, function_tolerance=0.0, gradient_tolerance=0.0, parameter_tolerance=0.0)
    if linear == "iterative_schur":
        oo.preconditioner_type = "schur_jacobi"
    ref, _ = trust_region_solve(PoisonedSolves(oba.BAModel(data, linear_solver=linear), 2), oo)
    assert [it.step_is_valid for it in summary.iterations[:4]] == [True, False, False, True]
    # accept/reject decisions once the cost has converged are rounding noise
    # (the trajectory comparison stops there); the counters that do not depend
    # on them must agree
    assert summary.num_residual_evaluations == ref.num_residual_evaluations
    assert summary.num_linear_solves == ref.num_linear_solves
    assert_trajectory_matches(summary, {"iterations": ref.iterations, "final_cost": ref.final_cost,
                                        "initial_cost": ref.initial_cost})
