"""Pins the CPU oracle before it is trusted as the checker of the HIP path.

Every oracle function is checked against something independent of its own
formula:

| oracle function                         | pinned by |
|-----------------------------------------|-----------|
| ba.angle_axis_rotate_point              | reference KAT (tests/test_manifolds_rotation.py:114-118), scipy Rotation (Rodrigues), small-angle branch |
| ba.snavely_residual_jacobian            | central finite differences; torch autograd of the quaternion formula |
| losses.loss_rho                         | reference KATs (tests/test_losses.py:6-19), finite differences of rho0 |
| losses.robustify                        | Triggs identity J^T r = rho1 J0^T r0 (tests/test_losses.py:31-46) |
| ba.schur_system / back-substitution     | damped least squares [J; sqrt D] (textbook identity), reference 6x4 KAT (tests/test_linear_schur.py:109-134) |
| ba.block_jacobi_pcg                     | reference KAT (tests/test_linear_schur.py:137-174), exact CG termination; JACOBI on a diagonal system in one step; IDENTITY in at most (distinct eigenvalues) steps |
| lm.LMRadius / StepEvaluator             | reference KATs (tests/test_problem_solver.py:551-577) |
| lm.updated_trust_region_radius (DOGLEG) | known answers of each branch of solver.py:1051-1061 |
| lm.traditional/subspace_dogleg_step     | brute force: the boundary crossing of the dogleg path on a fine grid; the model minimum over a polar grid of the 2-D subspace disc; normal equations for the interior step |
| ba.BAModel.gauss_newton_step / jmul     | dense J^T J solve and dense J v on the same evaluation |
| small_models.*.gauss_newton_step / jmul | least squares min |J dx + r| (numpy lstsq); J v vs central differences of the residual |
| covariance.dense_covariance             | reference KATs (tests/test_covariance_interpolation.py:7-223); (J^T J)^-1 for full-rank J |
| lm.refine_solution                      | reference KAT (tests/test_linear_schur.py:455-483); damped fixed point = (J^T J + D)^-1 J^T b |
| lm.trust_region_solve (BA, curve, SE2)  | golden trajectories written by running the reference (tests/golden/) |
| lm.trust_region_solve, invalid steps    | reference KAT (tests/test_problem_solver.py:637-664): NaN steps -> FAILURE, radius 4 -> 2 |
| generators                              | sha256 of the inputs recorded with each golden trajectory |
| ba.quaternion_plus / camera_plus        | reference KAT (tests/test_manifolds_rotation.py:6-13); scipy Rotation composition (rotvec 2 delta on the left); Eigen order = Ceres order permuted (tests/test_quaternion.py) |
| ba.quaternion_project                   | equals the angle-axis projection of the same rotation at non-unit q (tests/test_quaternion.py) |
| ba.quaternion_residual_jacobian         | torch autograd of the residual times plus_jacobian, the reference's procedure (problem.py:505-514), at non-unit q; central differences along Plus at unit q (tests/test_quaternion.py) |
| line_search.* (the LINE_SEARCH loop)    | interpolation steps exact on quadratics and cubics; LBFGS two-loop = dense BFGS updates; NCG betas = linear CG's under exact line searches; reference KATs: Rosenbrock to (1, 1) for LBFGS, BFGS, PR-CG and the three interpolations (tests/test_line_search.py) |
| lm.retained_block_sizes, split blocks   | known answers of linear.py:1269-1289; block-sparse 3 + 6 SCHUR_JACOBI trajectory = dense restatement with the reference's block rule (tests/test_quaternion.py) |
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import ba as oba
from oracle import generators as gen
from oracle.lm import (LMRadius, OracleOptions, StepEvaluator, dense_linear_solve, orthonormal_basis,
                       refine_solution, schur_solve_dense, subspace_dogleg_step, traditional_dogleg_step,
                       trust_region_solve, updated_trust_region_radius)
from oracle.losses import loss_rho, robustify
from oracle.small_models import CurveModel, PoseGraphModel

from conftest import load_golden


# ---------------------------------------------------------------------------
# rotation and Snavely residual
# ---------------------------------------------------------------------------


def test_angle_axis_rotation_known_answer():
    out = oba.angle_axis_rotate_point(np.array([0.0, 0.0, np.pi / 2]), np.array([1.0, 0.0, 0.0]))
    np.testing.assert_allclose(out, [0.0, 1.0, 0.0], atol=1e-12)


def test_angle_axis_rotation_matches_rodrigues_library():
    from scipy.spatial.transform import Rotation

    rng = np.random.default_rng(0)
    aa = rng.normal(size=(200, 3)) * np.array([[1.0], [1e-3], [1e-9], [3.0]]).repeat(50, 0)
    x = rng.normal(size=(200, 3))
    ref = Rotation.from_rotvec(aa).apply(x)
    np.testing.assert_allclose(oba.angle_axis_rotate_point(aa, x), ref, rtol=0, atol=1e-13)
    # theta == 0 takes the series branch: identity
    np.testing.assert_allclose(oba.angle_axis_rotate_point(np.zeros(3), x[0]), x[0], atol=0)


def _torch_project(cam, pt):
    aa = cam[:3]
    theta = torch.linalg.norm(aa)
    half = 0.5 * theta
    k = torch.sin(half) / theta
    q = torch.cat([torch.cos(half).reshape(1), k * aa])
    q = q / torch.linalg.norm(q)
    u = q[1:]
    uv = 2.0 * torch.linalg.cross(u, pt)
    p = pt + q[0] * uv + torch.linalg.cross(u, uv) + cam[3:6]
    xp, yp = -p[0] / p[2], -p[1] / p[2]
    r2 = xp * xp + yp * yp
    d = 1.0 + r2 * (cam[7] + cam[8] * r2)
    return cam[6] * d * torch.stack([xp, yp])


def test_snavely_jacobian_matches_autograd_and_finite_differences():
    rng = np.random.default_rng(1)
    d = gen.ba_config("ba4")
    cams = d.cameras[d.obs_camera][:40].copy()
    cams[:, 7] = rng.normal(0, 0.05, 40)   # exercise the distortion terms
    cams[:, 8] = rng.normal(0, 0.01, 40)
    pts = d.points[d.obs_point][:40]
    r, Jc, Jp = oba.snavely_residual_jacobian(cams, pts, d.observations[:40])
    for k in range(40):
        c = torch.tensor(cams[k], requires_grad=True)
        p = torch.tensor(pts[k], requires_grad=True)
        jac = torch.autograd.functional.jacobian(_torch_project, (c, p))
        np.testing.assert_allclose(Jc[k], jac[0].numpy(), rtol=1e-10, atol=1e-9)
        np.testing.assert_allclose(Jp[k], jac[1].numpy(), rtol=1e-10, atol=1e-9)
    # central differences (brute force) on a few entries
    h = 1e-6
    for j in range(9):
        cp, cm = cams.copy(), cams.copy()
        cp[:, j] += h
        cm[:, j] -= h
        fd = (oba.project(cp, pts) - oba.project(cm, pts)) / (2 * h)
        np.testing.assert_allclose(Jc[:, :, j], fd, rtol=1e-5, atol=1e-4)


def test_small_angle_jacobian_series_matches_finite_differences():
    cams = np.array([[1e-6, -2e-6, 5e-7, 0.1, 0.2, -5.0, 500.0, 0.01, 0.001]])
    pts = np.array([[0.3, -0.2, 1.0]])
    _, Jc, _ = oba.snavely_residual_jacobian(cams, pts, np.zeros((1, 2)))
    h = 1e-7
    for j in range(3):
        cp, cm = cams.copy(), cams.copy()
        cp[0, j] += h
        cm[0, j] -= h
        fd = (oba.project(cp, pts) - oba.project(cm, pts)) / (2 * h)
        np.testing.assert_allclose(Jc[0, :, j], fd[0], rtol=1e-6, atol=1e-5)


# ---------------------------------------------------------------------------
# losses and corrector
# ---------------------------------------------------------------------------


def test_loss_known_answers():
    r0, r1, r2 = loss_rho(("trivial",), np.array(2.0))
    assert (r0, r1, r2) == (2.0, 1.0, 0.0)
    r0, r1, r2 = loss_rho(("huber", 1.0), np.array(4.0))
    assert r0 == pytest.approx(3.0) and r1 == pytest.approx(0.5) and r2 < 0


@pytest.mark.parametrize("loss", [("huber", 1.0), ("softlone", 0.7), ("cauchy", 0.5), ("arctan", 2.0),
                                  ("tolerant", 1.0, 0.5), ("tukey", 3.0), ("scaled", ("cauchy", 0.5), 2.5)])
def test_loss_derivatives_match_finite_differences(loss):
    s = np.array([0.05, 0.3, 1.7, 4.0, 7.5])
    h = 1e-6
    r0, r1, r2 = loss_rho(loss, s)
    f = lambda v: loss_rho(loss, v)[0]
    g = lambda v: loss_rho(loss, v)[1]
    np.testing.assert_allclose(r1, (f(s + h) - f(s - h)) / (2 * h), rtol=1e-6, atol=1e-8)
    np.testing.assert_allclose(r2, (g(s + h) - g(s - h)) / (2 * h), rtol=1e-5, atol=1e-7)


def test_triggs_correction_identity():
    # tests/test_losses.py:31-46: J^T r after correction equals rho1 J0^T r0
    r = np.array([[2.0, 1.0]])
    J = np.array([[[1.0], [0.0]]])
    _, rc, Jc = robustify(("tolerant", 1.0, 1.0), r, J)
    _, rho1, _ = loss_rho(("tolerant", 1.0, 1.0), np.array([5.0]))
    assert not np.allclose(Jc, np.sqrt(rho1) * J)
    np.testing.assert_allclose(Jc[0].T @ rc[0], rho1 * (J[0].T @ r[0]), rtol=1e-12)


# ---------------------------------------------------------------------------
# linear algebra
# ---------------------------------------------------------------------------


def test_reference_dense_schur_known_answer():
    A = np.array([[2.0, 0.0, 1.0, 0.0], [0.0, 3.0, 0.0, 1.0], [1.0, -1.0, 2.0, 1.0], [0.0, 1.0, -1.0, 2.0],
                  [1.0, 2.0, 0.0, -1.0], [3.0, 0.0, 1.0, 1.0]])
    b = np.array([1.0, -2.0, 0.5, 3.0, -1.0, 2.0])
    D = np.array([0.1, 0.2, 0.3, 0.4])
    qr = np.linalg.lstsq(np.vstack([A, np.diag(np.sqrt(D))]), np.concatenate([b, np.zeros(4)]), rcond=None)[0]
    np.testing.assert_allclose(schur_solve_dense(A, b, 2, D), qr, atol=1e-10, rtol=1e-10)


def test_block_sparse_schur_step_equals_damped_least_squares():
    d = gen.ba_config("ba4")
    model = oba.BAModel(d)
    ev, _ = model.evaluate(model.x0(), True)
    J = model.dense_jacobian(ev)
    D = np.random.default_rng(3).uniform(0.1, 2.0, J.shape[1]) * model.column_sq(ev).clip(1e-6)
    step, _ = model.solve(ev, D, OracleOptions())
    ref = np.linalg.lstsq(np.vstack([J, np.diag(np.sqrt(D))]), np.concatenate([-ev.r.reshape(-1), np.zeros(J.shape[1])]),
                          rcond=None)[0]
    np.testing.assert_allclose(step, ref, rtol=1e-8, atol=1e-10 * np.abs(ref).max())


def test_pcg_known_answer_and_exact_termination():
    A = np.array([[2.0, 0.0, 1.0, 0.0], [0.0, 3.0, 0.0, 1.0], [1.0, -1.0, 2.0, 1.0], [0.0, 1.0, -1.0, 2.0],
                  [1.0, 2.0, 0.0, -1.0], [3.0, 0.0, 1.0, 1.0]])
    b = np.array([1.0, -2.0, 0.5, 3.0, -1.0, 2.0])
    D = np.array([0.1, 0.2, 0.3, 0.4])
    dense = schur_solve_dense(A, b, 2, D)
    H = A.T @ A + np.diag(D)
    rhs = A.T @ b
    S = H[2:, 2:] - H[2:, :2] @ np.linalg.solve(H[:2, :2], H[:2, 2:])
    rs = rhs[2:] - H[2:, :2] @ np.linalg.solve(H[:2, :2], rhs[:2])
    xb, iters, ok = oba.block_jacobi_pcg(S, rs, 1, tol=1e-14, max_iter=20)
    assert ok and 0 < iters <= 2
    np.testing.assert_allclose(xb, dense[2:], atol=1e-10, rtol=1e-10)
    # a 9x9-block SPD system converges in at most n iterations
    rng = np.random.default_rng(5)
    M = rng.normal(size=(27, 27))
    S2 = M @ M.T + 27 * np.eye(27)
    b2 = rng.normal(size=27)
    # (the reference also stops, unconverged, once |r.z| <= eps; tol=1e-6 stays clear of that)
    x2, it2, ok2 = oba.block_jacobi_pcg(S2, b2, 9, tol=1e-6, max_iter=100)
    assert ok2 and it2 <= 27
    np.testing.assert_allclose(x2, np.linalg.solve(S2, b2), rtol=1e-5, atol=1e-6)


# ---------------------------------------------------------------------------
# LM bookkeeping known answers
# ---------------------------------------------------------------------------


def test_lm_radius_known_answers():
    st = LMRadius(2.0, 20.0)
    assert st.step_rejected() == 1.0
    assert st.step_rejected() == 0.25
    assert st.step_accepted(1.0) == pytest.approx(0.75)
    assert st.step_accepted(1.0) == pytest.approx(2.25)
    assert st.step_accepted(0.25) == pytest.approx(2.0)
    assert st.step_accepted(1.0) == pytest.approx(6.0)
    assert st.step_accepted(1.0) == pytest.approx(18.0)
    assert st.step_accepted(1.0) == pytest.approx(20.0)


def test_step_evaluator_known_answers():
    ev = StepEvaluator(10.0, 2)
    assert ev.step_quality(9.0, 2.0) == pytest.approx(0.5)
    ev.step_accepted(9.0, 2.0)
    assert ev.step_quality(9.5, 1.0) == pytest.approx(1.0 / 6.0)
    ev.step_accepted(9.5, 1.0)
    assert ev.step_quality(9.4, 1.0) == pytest.approx(0.15)
    ev.step_accepted(9.4, 1.0)
    assert ev.step_quality(9.6, 1.0) == pytest.approx(-0.05)


class _HelloWorld:
    """residual 10 - x (tests/test_golden_validation.py:6-20)."""

    def x0(self):
        return np.array([0.5])

    def num_active_columns(self):
        return 1

    def state_norm(self, x):
        return float(np.abs(x[0]))

    def apply(self, x, step):
        return x + step

    def evaluate(self, x, jacobian):
        class E:
            pass

        e = E()
        e.r = np.array([10.0 - x[0]])
        e.J = np.array([[-1.0]])
        e.cost = 0.5 * float(e.r @ e.r)
        return e, (e.J.T @ e.r if jacobian else None)

    def column_sq(self, ev):
        return np.sum(ev.J * ev.J, axis=0)

    def solve(self, ev, damping, opts):
        return dense_linear_solve(ev.J, -ev.r, damping, 0, "dense_qr", opts)

    def model_decrease(self, ev, step):
        after = ev.r + ev.J @ step
        return max(0.0, 0.5 * float(ev.r @ ev.r) - 0.5 * float(after @ after))


def test_hello_world_known_answer():
    s, x = trust_region_solve(_HelloWorld(), OracleOptions(max_num_iterations=25, gradient_tolerance=1e-12))
    assert s.initial_cost == pytest.approx(45.125)
    assert abs(x[0] - 10.0) < 1e-6
    assert s.termination_type == "CONVERGENCE"


class PoisonedSolves:
    """Wraps an oracle model so that its first ``count`` linear solves return
    NaN, as the reference's test patches ``solve_linear_system``
    (tests/test_problem_solver.py:637-664)."""

    def __init__(self, model, count: int):
        self._model = model
        self.left = count

    def __getattr__(self, name):
        return getattr(self._model, name)

    def solve(self, ev, damping, opts):
        step, iters = self._model.solve(ev, damping, opts)
        if self.left > 0:
            self.left -= 1
            return np.full_like(step, np.nan), iters
        return step, iters


class _OneMinusX(_HelloWorld):
    """residual 1 - x from x = 0 (tests/test_problem_solver.py:640-642)."""

    def x0(self):
        return np.array([0.0])

    def evaluate(self, x, jacobian):
        e, g = super().evaluate(x, jacobian)
        e.r = np.array([1.0 - x[0]])
        e.cost = 0.5 * float(e.r @ e.r)
        return e, (e.J.T @ e.r if jacobian else None)


def test_invalid_steps_known_answer():
    # solver.py:168-190: every step NaN; two in a row end the solve with
    # FAILURE; the failing iteration is not recorded; LM halves the radius
    model = PoisonedSolves(_OneMinusX(), 100)
    s, x = trust_region_solve(model, OracleOptions(max_num_iterations=3, initial_trust_region_radius=4.0,
                                                   max_num_consecutive_invalid_steps=2, gradient_tolerance=0.0))
    assert s.termination_type == "FAILURE"
    assert s.message == "Too many invalid trust-region steps."
    assert s.num_unsuccessful_steps == 2
    assert len(s.iterations) == 2
    assert s.iterations[1]["step_is_valid"] is False
    assert s.iterations[1]["trust_region_radius"] == pytest.approx(2.0)
    assert x[0] == 0.0 and s.final_cost == s.initial_cost == 0.5
    # one NaN step, then the solve goes on from radius 2 (decrease factor 4 after the rejection)
    s1, x1 = trust_region_solve(PoisonedSolves(_OneMinusX(), 1),
                                OracleOptions(max_num_iterations=3, initial_trust_region_radius=4.0,
                                              gradient_tolerance=0.0))
    assert [it["step_is_valid"] for it in s1.iterations[:3]] == [True, False, True]
    assert s1.iterations[1]["trust_region_radius"] == pytest.approx(2.0)
    assert s1.iterations[2]["step_is_successful"] and s1.num_unsuccessful_steps == 1


def test_tiny_bal_known_answer():
    # examples/simple_bundle_adjuster.py:142-176 with fixed cameras;
    # tests/test_examples.py:151-165 expects final_cost < 1e-12, points to 1e-5
    true_cams = np.array([[0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 800.0, 0.0, 0.0],
                          [0.0, 0.02, 0.0, -0.4, 0.0, 0.05, 800.0, 0.0, 0.0]])
    true_pts = np.array([[0.2, -0.1, 4.0], [-0.3, 0.15, 4.5], [0.1, 0.2, 5.0]])
    init_pts = true_pts + np.array([[0.04, -0.03, 0.15], [-0.05, 0.02, -0.12], [0.03, 0.04, 0.10]])
    oc = np.array([0, 0, 0, 1, 1, 1])
    op = np.array([0, 1, 2, 0, 1, 2])
    obs = oba.project(true_cams[oc], true_pts[op])
    order = np.lexsort((oc, op))
    data = gen.BAProblemData(cameras=true_cams, points=init_pts, obs_camera=oc[order], obs_point=op[order],
                             observations=obs[order])
    model = oba.BAModel(data, camera_active=np.zeros(2, bool))
    s, (cams, pts) = trust_region_solve(model, OracleOptions(max_num_iterations=40, gradient_tolerance=1e-12,
                                                             function_tolerance=1e-12, parameter_tolerance=1e-12))
    assert s.final_cost < 1e-12
    np.testing.assert_allclose(pts, true_pts, atol=1e-5, rtol=1e-5)
    np.testing.assert_array_equal(cams, true_cams)


# ---------------------------------------------------------------------------
# trajectories written by the reference
# ---------------------------------------------------------------------------


def assert_trajectory_matches(summary, golden_summary, *, final_rtol=1e-10, cost_rtol=1e-9, max_exact_pcg=None):
    """Accept/reject sequence and costs agree until the cost has converged.

    Once the reference's cost is within 1e-11 (relative) of its final value,
    every remaining gain ratio is rounding noise in both implementations, so
    the comparison stops there; the final cost must agree to ``final_rtol``.

    ``max_exact_pcg``: preconditioned CG that runs many iterations (the
    JACOBI and IDENTITY preconditioners) amplifies rounding differences in the
    Krylov recurrence - the numpy restatement of the reference's own dense
    algebra lands 2.5e-6 away in the cost after a 19-iteration solve - so the
    comparison also stops after the first PCG solve longer than this.
    """
    gits = golden_summary["iterations"]
    final = golden_summary["final_cost"]
    conv = next((k for k, it in enumerate(gits) if abs(it["cost"] - final) <= 1e-11 * abs(final)), len(gits))
    if max_exact_pcg is not None:
        long_solve = next((k for k, it in enumerate(gits) if it["linear_solver_iterations"] > max_exact_pcg), len(gits))
        conv = min(conv, long_solve + 1)
    mine = summary.iterations if isinstance(summary.iterations[0], dict) else [vars(i) for i in summary.iterations]
    assert len(mine) >= min(conv, len(gits)), (len(mine), conv)
    for k in range(min(conv, len(gits))):
        a, b = mine[k], gits[k]
        assert a["iteration"] == b["iteration"]
        assert bool(a["step_is_successful"]) == b["step_is_successful"], k
        assert a["cost"] == pytest.approx(b["cost"], rel=cost_rtol), k
        assert a["trust_region_radius"] == pytest.approx(b["trust_region_radius"], rel=1e-3), k
        if "linear_solver_iterations" in b and "linear_solver_iterations" in a and (
                max_exact_pcg is None or b["linear_solver_iterations"] <= max_exact_pcg):
            # PCG iteration counts (1 for direct solves) are part of the trajectory
            assert int(a["linear_solver_iterations"]) == int(b["linear_solver_iterations"]), k
    assert summary.final_cost == pytest.approx(final, rel=final_rtol)
    assert summary.initial_cost == pytest.approx(golden_summary["initial_cost"], rel=1e-12)


def long_pcg_bound(g):
    """PCG length up to which trajectories are compared exactly (None: all) - see assert_trajectory_matches."""
    return 10 if g.get("preconditioner") in ("jacobi", "identity") else None


def _ba_opts(g):
    o = OracleOptions(max_num_iterations=g["max_num_iterations"])
    if g["linear_solver"] == "iterative_schur":
        o.preconditioner_type = g.get("preconditioner") or "schur_jacobi"
    if g["tolerances"] == "zero":
        o.function_tolerance = o.gradient_tolerance = o.parameter_tolerance = 0.0
    return o


@pytest.mark.parametrize("case", ["ba4_dense_schur_zero_tol", "ba4_dense_schur_default_tol", "ba6_huber_dense_schur",
                                  "ba4_iterative_schur_zero_tol", "ba4_iterative_jacobi_zero_tol",
                                  "ba4_iterative_identity_zero_tol"])
@pytest.mark.parametrize("faithful", [False, True])
def test_ba_trajectory_matches_reference(case, faithful):
    g = load_golden(case)
    d = gen.ba_config(g["config"])
    model = oba.BAModel(d, linear_solver=g["linear_solver"], dense_faithful=faithful)
    s, (cams, pts) = trust_region_solve(model, _ba_opts(g))
    # The Huber/outlier case is still moving at its 50-iteration cap and its
    # last steps are ill-conditioned: even the dense restatement (same algebra
    # as the reference, numpy LAPACK instead of torch) lands 2e-9 away, so the
    # final-cost bound there is 1e-7; trivial-loss cases are held to 1e-10.
    final_rtol = 1e-7 if "huber" in case else 1e-10
    assert_trajectory_matches(s, g["summary"], final_rtol=final_rtol, max_exact_pcg=long_pcg_bound(g))
    if g["tolerances"] == "default":
        assert s.termination_type == g["summary"]["termination_type"]
        assert s.message == g["summary"]["message"]
        assert len(s.iterations) == len(g["summary"]["iterations"])
    if g.get("final_points") is not None and "huber" not in case and long_pcg_bound(g) is None:
        # (after a trajectory that differs by rounding, the 7-dof gauge of the final state drifts)
        # BA has a 7-dof gauge freedom; compare reprojection-level agreement
        np.testing.assert_allclose(pts, np.array(g["final_points"]), atol=1e-6)


def test_ba16_trajectory_matches_reference():
    g = load_golden("ba16_dense_schur_zero_tol")
    d = gen.ba_config("ba16")
    s, (cams, pts) = trust_region_solve(oba.BAModel(d), _ba_opts(g))
    assert_trajectory_matches(s, g["summary"])
    ref = np.load(__import__("conftest").GOLDEN / "ba16_dense_schur_zero_tol_final.npz")
    np.testing.assert_allclose(pts, ref["points"], atol=1e-6)


def test_curve_trajectories_match_reference():
    g = load_golden("curve_cauchy")
    batch = gen.make_curve_batch(num_problems=g["num_problems"], seed=g["seed"])
    for case in g["problems"]:
        i = case["index"]
        lo, hi = batch.offsets[i], batch.offsets[i + 1]
        model = CurveModel(batch.x[lo:hi], batch.y[lo:hi], loss=tuple(g["loss"]))
        s, x = trust_region_solve(model, OracleOptions(max_num_iterations=g["max_num_iterations"]))
        assert_trajectory_matches(s, case["summary"])
        assert s.termination_type == case["summary"]["termination_type"]
        np.testing.assert_allclose(x, case["final"], rtol=1e-7, atol=1e-9)


def test_pose_graph_trajectory_matches_reference():
    g = load_golden("pose2d_40")
    pg = gen.make_pose_graph(num_nodes=g["num_nodes"], seed=g["seed"])
    model = PoseGraphModel(pg.poses, pg.edges, pg.measurements)
    s, x = trust_region_solve(model, OracleOptions(max_num_iterations=g["max_num_iterations"]))
    assert_trajectory_matches(s, g["summary"])
    assert s.termination_type == g["summary"]["termination_type"]
    np.testing.assert_allclose(x, np.array(g["final_poses"]), atol=1e-8)


@pytest.mark.parametrize("case", ["ba4_dense_schur_zero_tol", "ba6_huber_dense_schur", "ba16_dense_schur_zero_tol"])
def test_generator_inputs_match_golden_digest(case):
    import hashlib

    g = load_golden(case)
    d = gen.ba_config(g["config"])
    h = hashlib.sha256()
    for a in (d.cameras, d.points, d.observations, d.obs_camera, d.obs_point):
        h.update(np.ascontiguousarray(a).tobytes())
    assert h.hexdigest() == g["input_sha256"]


# ---------------------------------------------------------------------------
# dogleg (linear.py:554-664, solver.py:864-883, 1051-1061)
# ---------------------------------------------------------------------------


def test_dogleg_radius_known_answers():
    assert updated_trust_region_radius(1.0, 0.8, 0.1, 1e16) == 2.0      # expand to 2 radius
    assert updated_trust_region_radius(1.0, 0.8, 1.0, 1e16) == 3.0      # ... or 3 |step|
    assert updated_trust_region_radius(1.0, 0.8, 1.0, 2.5) == 2.5       # capped
    assert updated_trust_region_radius(1.0, 0.5, 1.0, 1e16) == 1.0      # unchanged
    assert updated_trust_region_radius(1.0, 0.1, 1.0, 1e16) == 0.5      # shrink
    assert updated_trust_region_radius(1e-320, 0.1, 1.0, 1e16) == np.finfo(np.float64).tiny


def _dogleg_case(seed, m=14, n=5):
    rng = np.random.default_rng(seed)
    J = rng.normal(size=(m, n)) * np.linspace(0.3, 3.0, n)
    r = rng.normal(size=m)
    g = J.T @ r
    gn = np.linalg.solve(J.T @ J, -g)
    return J, r, g, gn


def test_traditional_dogleg_branches_by_brute_force():
    for seed in range(6):
        J, r, g, gn = _dogleg_case(seed)
        jmul = lambda v: J @ v  # noqa: E731
        # interior: the Gauss-Newton step solves the normal equations
        s = traditional_dogleg_step(jmul, g, gn, 2.0 * np.linalg.norm(gn))
        np.testing.assert_allclose(J.T @ (J @ s), -g, rtol=1e-10, atol=1e-12)
        # the Cauchy point, by minimising the model along -g on a fine grid
        ts = np.linspace(0.0, 10.0, 200001) * np.linalg.norm(g) ** 2 / max(np.linalg.norm(J @ g) ** 2, 1e-300)
        model = [0.5 * np.sum((r - t * (J @ g)) ** 2) for t in ts[::100]]
        t_star = ts[::100][int(np.argmin(model))]
        sd = -t_star * g
        # small radius: the step is -g scaled to the boundary
        rad = 0.25 * np.linalg.norm(sd)
        s = traditional_dogleg_step(jmul, g, gn, rad)
        assert np.linalg.norm(s) == pytest.approx(rad, rel=1e-12)
        assert float(s @ g) < 0 and abs(float(s @ g)) == pytest.approx(np.linalg.norm(s) * np.linalg.norm(g), rel=1e-12)
        # between: the boundary crossing of the path sd -> gn, found on a grid of the segment
        sd_exact = -(g @ g) / np.sum((J @ g) ** 2) * g
        np.testing.assert_allclose(sd_exact, sd, rtol=2e-3)
        lo, hi = np.linalg.norm(sd_exact), np.linalg.norm(gn)
        if hi <= lo * 1.01:
            continue
        rad = 0.5 * (lo + hi)
        s = traditional_dogleg_step(jmul, g, gn, rad)
        taus = np.linspace(0.0, 1.0, 100001)
        pts = sd_exact[None, :] + taus[:, None] * (gn - sd_exact)[None, :]
        k = int(np.argmax(np.linalg.norm(pts, axis=1) >= rad))
        np.testing.assert_allclose(s, pts[k], atol=2e-4 * np.linalg.norm(gn))
        assert np.linalg.norm(s) == pytest.approx(rad, rel=1e-12)


def test_subspace_dogleg_minimises_the_model_on_the_subspace_disc():
    for seed in range(4):
        J, r, g, gn = _dogleg_case(10 + seed)
        jmul = lambda v: J @ v  # noqa: E731
        sd = -(g @ g) / np.sum((J @ g) ** 2) * g
        rad = 0.5 * (np.linalg.norm(sd) + np.linalg.norm(gn))
        if np.linalg.norm(gn) <= np.linalg.norm(sd) * 1.01:
            continue
        s = subspace_dogleg_step(jmul, r, g, gn, rad)
        B = orthonormal_basis((-g, gn))
        assert B.shape[1] == 2
        np.testing.assert_allclose(B.T @ B, np.eye(2), atol=1e-12)
        # brute force over the disc |y| <= rad in span(B) (the minimum is on the boundary)
        ang = np.linspace(0.0, 2.0 * np.pi, 20001)
        ys = rad * np.stack([np.cos(ang), np.sin(ang)], axis=1)
        vals = 0.5 * np.sum((r[None, :] + (J @ B @ ys.T).T) ** 2, axis=1)
        best = float(vals.min())
        mine = 0.5 * float(np.sum((r + J @ s) ** 2))
        assert np.linalg.norm(s) <= rad * (1.0 + 1e-9)
        assert mine <= best + 1e-9 * abs(best)
        assert mine >= best - 1e-6 * abs(best)  # and it is that minimum, not a better point off the subspace
        np.testing.assert_allclose(B @ (B.T @ s), s, atol=1e-12 * max(1.0, np.linalg.norm(s)))


def test_ba_gauss_newton_step_and_jmul_match_dense_algebra():
    d = gen.ba_config("ba4")
    active = np.ones(d.num_cameras, bool)
    active[:2] = False
    model = oba.BAModel(d, camera_active=active)
    ev, g = model.evaluate(model.x0(), True)
    J = model.dense_jacobian(ev)
    np.testing.assert_allclose(model.gauss_newton_step(ev), np.linalg.solve(J.T @ J, -g), rtol=1e-8, atol=1e-12)
    v = np.random.default_rng(3).normal(size=J.shape[1])
    np.testing.assert_allclose(model.jmul(ev, v), J @ v, rtol=1e-12, atol=1e-12)


def _small_model_cases():
    batch = gen.make_curve_batch(num_problems=2, seed=5)
    lo, hi = batch.offsets[0], batch.offsets[1]
    curve = CurveModel(batch.x[lo:hi], batch.y[lo:hi], loss=None, m0=0.1, c0=0.2)
    pg = gen.make_pose_graph(num_nodes=12, seed=6)
    pose = PoseGraphModel(pg.poses, pg.edges, pg.measurements, loss=None)
    return [curve, pose]


@pytest.mark.parametrize("which", [0, 1])
def test_small_model_gauss_newton_step_and_jmul(which):
    # gauss_newton_step is the minimiser of |J dx + r| (linear.py:570-571 solves
    # the normal equations of it); numpy's lstsq reaches it by an SVD instead.
    # jmul is J v for the Jacobian of the residual: with the trivial loss, the
    # directional derivative of r along v, by central differences.
    model = _small_model_cases()[which]
    x0 = model.x0()
    ev, _ = model.evaluate(x0, True)
    gn = model.gauss_newton_step(ev)
    ref = np.linalg.lstsq(ev.J, -ev.r, rcond=None)[0]
    np.testing.assert_allclose(gn, ref, rtol=1e-7, atol=1e-10)
    v = np.random.default_rng(11 + which).normal(size=ev.J.shape[1])
    h = 1e-6
    rp, _ = model.evaluate(model.apply(x0, h * v), False)
    rm, _ = model.evaluate(model.apply(x0, -h * v), False)
    fd = (rp.r - rm.r) / (2 * h)
    np.testing.assert_allclose(model.jmul(ev, v), fd, rtol=1e-6, atol=1e-6 * np.abs(fd).max())


def test_refine_solution_reference_known_answer():
    # tests/test_linear_schur.py:455-483: a float32 least-squares solution of a
    # nearly rank-deficient system, refined in float64 (linear.py:1331-1363),
    # recovers the double solution and lowers the residual
    A = np.array([[1.0, 1.0], [1.0, 1.0 + 1e-7], [1.0, 1.0 - 1e-7]])
    true_x = np.array([2.0, -3.0])
    b = A @ true_x
    low = np.linalg.lstsq(A.astype(np.float32), b.astype(np.float32), rcond=None)[0].astype(np.float64)
    refined = refine_solution(A, b, low, None, 5)
    assert np.linalg.norm(b - A @ refined) < np.linalg.norm(b - A @ low)
    np.testing.assert_allclose(refined, true_x, atol=1e-6, rtol=1e-6)
    # no refinement steps: unchanged
    assert refine_solution(A, b, low, None, 0) is low


def test_refine_solution_damped_fixed_point():
    # with damping the refinement works on [A; sqrt(D)] x = [b; 0]: from any
    # start it lands on the damped normal-equation solution; a solution that
    # already meets the 10 eps |b| test is returned as it is
    rng = np.random.default_rng(12)
    A = rng.normal(size=(30, 6))
    b = rng.normal(size=30)
    d = rng.uniform(0.5, 2.0, size=6)
    exact = np.linalg.solve(A.T @ A + np.diag(d), A.T @ b)
    out = refine_solution(A, b, np.zeros(6), d, 2)
    np.testing.assert_allclose(out, exact, rtol=1e-12, atol=1e-13)
    undamped = np.linalg.lstsq(A, b, rcond=None)[0]
    assert np.linalg.norm(out - undamped) > 1e-3  # the damping rows were not dropped
    A2 = rng.normal(size=(6, 6))
    x2 = rng.normal(size=6)
    b2 = A2 @ x2
    assert not np.any(b2 - A2 @ x2)  # the same product: the residual is exactly zero
    np.testing.assert_array_equal(refine_solution(A2, b2, x2.copy(), None, 3), x2)
    # dense_linear_solve refines the DENSE_SCHUR step only in the mixed option (linear.py:181-187)
    o = OracleOptions(use_mixed_precision_solves=True, max_num_refinement_iterations=2)
    x_mixed, _ = dense_linear_solve(A, b, d, 2, "dense_schur", o)
    np.testing.assert_allclose(x_mixed, exact, rtol=1e-11, atol=1e-12)


def _dogleg_opts(g):
    o = _ba_opts(g)
    o.trust_region_strategy_type = g["trust_region_strategy"]
    o.dogleg_type = g["dogleg_type"]
    if g.get("initial_trust_region_radius") is not None:
        o.initial_trust_region_radius = g["initial_trust_region_radius"]
    return o


@pytest.mark.parametrize("case", ["ba4_dogleg_zero_tol", "ba4_subspace_dogleg_zero_tol", "ba4_dogleg_default_tol"])
@pytest.mark.parametrize("faithful", [False, True])
def test_ba_dogleg_trajectory_matches_reference(case, faithful):
    g = load_golden(case)
    d = gen.ba_config(g["config"])
    active = np.ones(d.num_cameras, bool)
    active[g["constant_cameras"]] = False
    model = oba.BAModel(d, linear_solver=g["linear_solver"], dense_faithful=faithful, camera_active=active)
    s, _ = trust_region_solve(model, _dogleg_opts(g))
    assert_trajectory_matches(s, g["summary"])
    assert s.num_linear_solves == g["summary"]["num_linear_solves"] == 0
    if g["tolerances"] == "default":
        assert s.termination_type == g["summary"]["termination_type"]
        assert len(s.iterations) == len(g["summary"]["iterations"])


def test_pcg_jacobi_and_identity_preconditioners():
    # JACOBI (linear.py:531-537) is exact on a diagonal system: one iteration
    rng = np.random.default_rng(8)
    d = rng.uniform(0.5, 4.0, size=18)
    b = rng.normal(size=18)
    x, it, ok = oba.block_jacobi_pcg(np.diag(d), b, 9, tol=1e-14, max_iter=50, preconditioner="jacobi")
    assert ok and it == 1
    np.testing.assert_allclose(x, b / d, rtol=1e-14)
    # IDENTITY (plain CG) ends in at most as many steps as the matrix has distinct eigenvalues
    Q, _ = np.linalg.qr(rng.normal(size=(18, 18)))
    S = (Q * np.repeat([1.0, 3.0, 7.0], 6)) @ Q.T
    x, it, ok = oba.block_jacobi_pcg(S, b, 9, tol=1e-12, max_iter=50, preconditioner="identity")
    assert ok and it <= 3
    np.testing.assert_allclose(S @ x, b, rtol=1e-10, atol=1e-10)
    with pytest.raises(ValueError):
        oba.block_jacobi_pcg(S, b, 9, tol=1e-12, max_iter=5, preconditioner="cluster_jacobi")


def test_segment_sum_handles_empty_segments():
    from oracle.ba import _segment_sum
    rng = np.random.default_rng(3)
    counts = np.array([0, 2, 0, 0, 3, 1, 0])
    offsets = np.r_[0, np.cumsum(counts)]
    values = rng.normal(size=(offsets[-1], 2, 2))
    want = np.stack([values[offsets[k]: offsets[k + 1]].sum(axis=0) for k in range(len(counts))])
    got = _segment_sum(values, offsets)
    np.testing.assert_array_equal(got[counts == 0], 0.0)
    np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-15)


def test_unobserved_blocks_stay_fixed_and_match_the_dense_algebra():
    """Points seen by no camera and a camera that sees nothing keep zero Jacobian columns in the
    reference (problem.py:522-530 keeps every block with a tangent space in the column map), so the
    LM damping alone fixes their step at zero.  The block-sparse oracle must agree with the
    reference's dense algebra on the full Jacobian (dense_faithful) and leave those blocks unchanged."""
    d = gen.make_bundle_adjustment(num_cameras=6, num_points=90, obs_per_point=(1, 4), seed=23)
    keep = d.obs_camera != 5
    d.obs_camera, d.obs_point, d.observations = d.obs_camera[keep], d.obs_point[keep], d.observations[keep]
    unseen = np.bincount(d.obs_point, minlength=d.points.shape[0]) == 0
    assert unseen.sum() == 3
    opts = OracleOptions(max_num_iterations=10)
    s, (cams, pts) = trust_region_solve(oba.BAModel(d), opts)
    sd, (cams_d, pts_d) = trust_region_solve(oba.BAModel(d, dense_faithful=True), opts)
    assert_trajectory_matches(s, {"iterations": sd.iterations, "final_cost": sd.final_cost,
                                  "initial_cost": sd.initial_cost})
    np.testing.assert_allclose(pts, pts_d, atol=1e-9)
    np.testing.assert_array_equal(pts[unseen], d.points[unseen])
    np.testing.assert_array_equal(cams[5], d.cameras[5])


# ---------------------------------------------------------------------------
# covariance (src/ceres_torch/covariance.py), restated on dense Jacobians
# ---------------------------------------------------------------------------


def test_covariance_reference_known_answers():
    from oracle.covariance import dense_covariance

    # tests/test_covariance_interpolation.py:7-24: residual 2x - 1
    cov, s = dense_covariance(np.array([[2.0]]))
    np.testing.assert_allclose(cov, [[0.25]], atol=1e-9)
    assert s["success"] and s["rank"] == 1 and s["nullity"] == 0 and s["reciprocal_condition_number"] == 1.0
    # :27-44: NormalPrior diag(2, 4)
    cov, _ = dense_covariance(np.diag([2.0, 4.0]))
    np.testing.assert_allclose(cov, np.diag([0.25, 0.0625]), atol=1e-12)
    # :62-81: SubsetManifold(3, [1]) leaves the tangent Jacobian diag(2, 3)
    cov, _ = dense_covariance(np.array([[2.0, 0.0], [0.0, 3.0]]))
    np.testing.assert_allclose(cov, np.diag([0.25, 1.0 / 9.0]), atol=1e-12)
    # :84-105: x + y, rank 1 of 2
    cov, s = dense_covariance(np.array([[1.0, 1.0]]))
    assert not s["success"] and s["rank"] == 1 and s["nullity"] == 1 and "rank deficient" in s["message"]
    cov, s = dense_covariance(np.array([[1.0, 1.0]]), null_space_rank=-1)
    assert s["success"] and s["rank"] == 1
    np.testing.assert_allclose(cov, np.full((2, 2), 0.25), atol=1e-9)
    # :108-130: b - a - 1, the translation gauge
    cov, s = dense_covariance(np.array([[-1.0, 1.0]]))
    assert not s["success"] and s["rank"] == 1 and s["nullity"] == 1
    cov, s = dense_covariance(np.array([[-1.0, 1.0]]), null_space_rank=-1)
    np.testing.assert_allclose(cov @ np.ones(2), np.zeros(2), atol=1e-10)
    assert np.array([-1.0, 1.0]) @ cov @ np.array([-1.0, 1.0]) == pytest.approx(1.0)
    # :140-168: near singular, eigenvalue-ratio policy
    A = np.diag([1.0, 1e-4])
    cov, s = dense_covariance(A, min_reciprocal_condition_number=1e-6)
    assert not s["success"] and s["rank"] == 2 and s["reciprocal_condition_number"] < 1e-6
    cov, s = dense_covariance(A, min_reciprocal_condition_number=1e-6, null_space_rank=-1)
    assert s["success"] and s["rank"] == 1 and s["nullity"] == 1
    np.testing.assert_allclose(cov, np.diag([1.0, 0.0]), atol=1e-12)
    # :171-183: Huber(1) on x - 10 at x = 0: the corrected Jacobian is sqrt(rho1) = sqrt(0.1)
    _, _, Jc = robustify(("huber", 1.0), np.array([[-10.0]]), np.array([[[1.0]]]))
    cov, _ = dense_covariance(Jc.reshape(1, 1))
    np.testing.assert_allclose(cov, [[10.0]], rtol=1e-12)
    # :186-198: SPARSE_QR equals DENSE_SVD for a full-rank problem
    A = np.array([[2.0, 0.0], [1.0, 3.0], [0.5, -1.0]])
    np.testing.assert_allclose(dense_covariance(A, algorithm="sparse_qr")[0], dense_covariance(A)[0], atol=1e-10)
    # :201-223: SPARSE_QR ignores null_space_rank and fails on a rank-deficient J
    A = np.array([[1.0, 1.0], [2.0, 2.0]])
    assert dense_covariance(A, null_space_rank=-1)[1]["success"]
    cov, s = dense_covariance(A, algorithm="sparse_qr", null_space_rank=-1)
    assert cov is None and not s["success"] and "Sparse QR" in s["message"]


def test_covariance_full_rank_is_the_inverse_normal_matrix():
    # for a full-rank J both algorithms give (J^T J)^-1 (textbook); the QR
    # form reaches it through R^-1 R^-T, the SVD form through V diag(s^-2) V^T
    from oracle.covariance import dense_covariance

    rng = np.random.default_rng(21)
    J = rng.normal(size=(40, 7)) @ np.diag(np.logspace(0, 3, 7))
    inv = np.linalg.inv(J.T @ J)
    for alg in ("dense_svd", "sparse_qr"):
        cov, s = dense_covariance(J, algorithm=alg)
        np.testing.assert_allclose(cov, inv, rtol=1e-9, atol=1e-12 * np.abs(inv).max())
        assert s["success"] and s["rank"] == 7
        sv = np.linalg.svd(J, compute_uv=False)
        assert s["reciprocal_condition_number"] == pytest.approx((sv.min() / sv.max()) ** 2)
