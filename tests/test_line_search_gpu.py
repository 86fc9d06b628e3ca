"""The LINE_SEARCH minimizer on the MI355X (ct_ba_line_search_solve) against the float64 oracle.

The oracle (oracle/line_search.py, pinned in tests/test_line_search.py) restates
_line_search_solve (solver.py:360-607) on the block-sparse BA model.  The interpolated
step sizes are built from candidate_cost - cost - derivative * t, a difference of nearly
equal costs, so the summation-order rounding of any two FP64 implementations (1e-16 of
the cost) becomes 1e-8 to 1e-5 of the step, and a few iterations later a different trial
count.  So the comparison is exact over the prefix on which the oracle itself is stable -
the iterations on which oracle runs with rounding-sized perturbations (costs by 1e-14,
gradient components by 1e-15 and 1e-14) agree with the plain run: there every accept flag,
trial count and restart agrees, costs to 1e-7 and step sizes to 1e-4.  Past it, the device run must still be a descent trajectory (every
accepted step lowers the cost) with the oracle's termination.  With the
reference's default 20 trials a unit first step on bundle adjustment cannot shrink enough
(each trial contracts by at most 0.5), so the default-option cases end in the reference's
"Line search failed" at the same iteration; 60 trials let the solve make progress.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import ceres_mi355x as tc
from oracle import ba as oba
from oracle import generators as gen
from oracle.line_search import LineSearchOptions, line_search_solve

from test_ba_gpu import ba_problem
from test_quaternion_gpu import quaternion_problem

pytestmark = pytest.mark.gpu

_MESSAGES = {"Gradient tolerance reached.", "Parameter tolerance reached.", "Function tolerance reached.",
             "Maximum iterations reached.", "Line search failed to find a decreasing step.",
             "Projected gradient tolerance reached."}


def _options(**kw):
    o = tc.SolverOptions(minimizer_type=tc.MinimizerType.LINE_SEARCH)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _oracle_options(o: tc.SolverOptions) -> LineSearchOptions:
    return LineSearchOptions(
        max_num_iterations=o.max_num_iterations, direction=o.line_search_direction_type.name.lower(),
        line_search_type=o.line_search_type.name.lower(), ncg_type=o.nonlinear_conjugate_gradient_type.name.lower(),
        max_lbfgs_rank=o.max_lbfgs_rank, use_approximate_eigenvalue_bfgs_scaling=o.use_approximate_eigenvalue_bfgs_scaling,
        interpolation=o.line_search_interpolation_type.name.lower(), min_line_search_step_size=o.min_line_search_step_size,
        sufficient_function_decrease=o.line_search_sufficient_function_decrease,
        max_step_contraction=o.max_line_search_step_contraction, min_step_contraction=o.min_line_search_step_contraction,
        max_step_size_iterations=o.max_num_line_search_step_size_iterations,
        max_direction_restarts=o.max_num_line_search_direction_restarts,
        sufficient_curvature_decrease=o.line_search_sufficient_curvature_decrease,
        function_tolerance=o.function_tolerance, gradient_tolerance=o.gradient_tolerance,
        parameter_tolerance=o.parameter_tolerance)


def _agree(a, b, cost_rtol, step_rtol) -> bool:
    return (a["step_is_successful"] == b["step_is_successful"]
            and a["line_search_iterations"] == b["line_search_iterations"]
            and a["line_search_function_evaluations"] == b["line_search_function_evaluations"]
            and a["line_search_direction_restarts"] == b["line_search_direction_restarts"]
            and abs(a["cost"] - b["cost"]) <= cost_rtol * abs(b["cost"])
            and abs(a["step_size"] - b["step_size"]) <= step_rtol * abs(b["step_size"]))


class _Rounded:
    """The model with rounding of the size two FP64 implementations differ by: every cost scaled by
    1 + 1e-14 and, with a seed, every gradient component by 1 +- rel (signs drawn from the seed).  A
    gradient component is a sum with cancellation, so its rounding relative to itself exceeds eps."""

    def __init__(self, model, seed=None, rel=1e-15):
        self.model = model
        self.rel = rel
        self.rng = None if seed is None else np.random.default_rng(seed)

    def x0(self):
        return self.model.x0()

    def evaluate(self, x, jacobian=True):
        ev, g = self.model.evaluate(x, jacobian)
        ev.cost = ev.cost * (1.0 + 1e-14)
        if g is not None and self.rng is not None:
            g = g * (1.0 + self.rel * self.rng.choice([-1.0, 1.0], size=g.shape))
        return ev, g

    def apply(self, x, step):
        return self.model.apply(x, step)

    def state_norm(self, x):
        return self.model.state_norm(x)


def stable_prefix(model, opts, ref, *, cost_rtol=1e-7, step_rtol=1e-4) -> int:
    """Leading iterations on which the oracle agrees with itself under rounding-sized perturbations
    (costs alone; costs and gradients at 1e-15 and 1e-14, two seeds each): the smallest such prefix."""
    prefix = len(ref.iterations)
    for seed, rel in ((None, 0.0), (0, 1e-15), (1, 1e-15), (0, 1e-14), (1, 1e-14)):
        other, _ = line_search_solve(_Rounded(model, seed, rel), opts)
        k = 0
        for a, b in zip(ref.iterations, other.iterations):
            if not _agree(a, b, cost_rtol, step_rtol):
                break
            k += 1
        prefix = min(prefix, k)
    return prefix


def assert_line_search_matches(summary, ref, prefix=None, *, cost_rtol=1e-7, step_rtol=1e-4):
    """Exact agreement over the first ``prefix`` iterations (all when None), a descent trajectory after."""
    assert summary.minimizer_type is tc.MinimizerType.LINE_SEARCH
    full = prefix is None or prefix >= len(ref.iterations)
    if full:
        assert len(summary.iterations) == len(ref.iterations), (len(summary.iterations), len(ref.iterations))
    else:
        assert prefix >= 5, prefix  # the comparison pins at least five iterations
        mine = [vars(it) for it in summary.iterations]
        assert len(mine) >= prefix
        costs = [it["cost"] for it in mine]
        for k in range(1, len(mine)):
            if mine[k - 1]["step_is_successful"]:
                assert costs[k] < costs[k - 1], k
        assert summary.termination_type.name == ref.termination
        assert summary.final_cost <= costs[prefix - 1]
    for a, b in list(zip(summary.iterations, ref.iterations))[: None if full else prefix]:
        k = b["iteration"]
        assert a.iteration == k
        assert a.cost == pytest.approx(b["cost"], rel=cost_rtol), k
        assert a.step_is_successful == b["step_is_successful"], k
        assert a.line_search_iterations == b["line_search_iterations"], k
        assert a.line_search_function_evaluations == b["line_search_function_evaluations"], k
        assert a.line_search_direction_restarts == b["line_search_direction_restarts"], k
        assert a.step_size == pytest.approx(b["step_size"], rel=step_rtol, abs=1e-300), k
        assert a.gradient_max_norm == pytest.approx(b["gradient_max_norm"], rel=1e-6), k
        if b["step_is_successful"]:
            assert a.step_norm == pytest.approx(b["step_norm"], rel=step_rtol), k
    if not full:
        return
    assert summary.message == ref.message
    assert summary.termination_type.name == ref.termination
    assert summary.final_cost == pytest.approx(ref.final_cost, rel=cost_rtol)
    assert summary.initial_cost == pytest.approx(ref.initial_cost, rel=1e-12)
    assert summary.num_successful_steps == ref.num_successful_steps
    assert summary.num_unsuccessful_steps == ref.num_unsuccessful_steps
    assert summary.num_jacobian_evaluations == ref.num_jacobian_evaluations
    assert summary.num_line_search_steps == ref.num_line_search_steps
    assert summary.num_line_search_function_evaluations == ref.num_line_search_function_evaluations
    assert summary.num_line_search_direction_restarts == ref.num_line_search_direction_restarts


def _run_bulk(data, o, camera_active=None):
    """Device and oracle runs and the oracle's stable prefix (None: the whole run is stable)."""
    problem, cams, pts = ba_problem(data)
    if camera_active is not None:
        for i in np.nonzero(~camera_active)[0]:
            problem.SetParameterBlockConstant(problem.parameter_blocks[int(i)])  # cameras come first
    summary = tc.solve(o, problem)
    model = oba.BAModel(data, camera_active=camera_active)
    oo = _oracle_options(o)
    ref, (rc, rp) = line_search_solve(model, oo)
    prefix = stable_prefix(model, oo, ref)
    return summary, ref, prefix, cams.cpu().numpy(), pts.cpu().numpy(), rc, rp


@pytest.mark.parametrize("name", ["ba4", "ba16"])
def test_line_search_default_options_fail_like_the_reference(name):
    """Default options: the unit first step shrinks by at most 0.5 per trial, 20 trials do not reach a
    decrease on bundle adjustment, and the solve ends with the reference's message at the same iteration."""
    print(f"line search {name} defaults", flush=True)
    data = gen.ba_config(name)
    s, ref, prefix, cams, pts, rc, rp = _run_bulk(data, _options(max_num_iterations=10))
    assert ref.message == "Line search failed to find a decreasing step."
    assert prefix == len(ref.iterations)
    assert_line_search_matches(s, ref)
    np.testing.assert_allclose(pts, rp, atol=1e-9)


@pytest.mark.parametrize("direction,ncg", [("LBFGS", "FLETCHER_REEVES"), ("STEEPEST_DESCENT", "FLETCHER_REEVES"),
                                           ("NONLINEAR_CONJUGATE_GRADIENT", "FLETCHER_REEVES"),
                                           ("NONLINEAR_CONJUGATE_GRADIENT", "POLAK_RIBIERE"),
                                           ("NONLINEAR_CONJUGATE_GRADIENT", "HESTENES_STIEFEL"),
                                           ("BFGS", "FLETCHER_REEVES")])
def test_line_search_directions_match_oracle(direction, ncg):
    print(f"line search ba4 {direction} {ncg}", flush=True)
    data = gen.ba_config("ba4")
    o = _options(max_num_iterations=12, max_num_line_search_step_size_iterations=60,
                 line_search_direction_type=tc.LineSearchDirectionType[direction],
                 nonlinear_conjugate_gradient_type=tc.NonlinearConjugateGradientType[ncg])
    s, ref, prefix, cams, pts, rc, rp = _run_bulk(data, o)
    assert ref.num_successful_steps >= 5
    assert_line_search_matches(s, ref, prefix)
    if prefix == len(ref.iterations):
        np.testing.assert_allclose(pts, rp, rtol=1e-6, atol=1e-8)


@pytest.mark.parametrize("interpolation,ls_type", [("BISECTION", "WOLFE"), ("QUADRATIC", "ARMIJO"),
                                                   ("CUBIC", "ARMIJO")])
def test_line_search_interpolations_match_oracle(interpolation, ls_type):
    """ARMIJO with nonlinear CG (the reference's validation takes WOLFE only for BFGS and LBFGS)."""
    print(f"line search ba4 {interpolation} {ls_type}", flush=True)
    data = gen.ba_config("ba4")
    o = _options(max_num_iterations=12, max_num_line_search_step_size_iterations=60,
                 line_search_interpolation_type=tc.LineSearchInterpolationType[interpolation],
                 line_search_type=tc.LineSearchType[ls_type],
                 line_search_direction_type=tc.LineSearchDirectionType.NONLINEAR_CONJUGATE_GRADIENT)
    s, ref, prefix, *_ = _run_bulk(data, o)
    assert_line_search_matches(s, ref, prefix)


def test_line_search_lbfgs_ba16_with_constant_cameras_and_huber():
    print("line search ba16_huber lbfgs", flush=True)
    data = gen.ba_config("ba16_huber")
    active = np.ones(data.num_cameras, bool)
    active[[0, 5]] = False
    o = _options(max_num_iterations=15, max_num_line_search_step_size_iterations=60, max_lbfgs_rank=5)
    s, ref, prefix, cams, pts, rc, rp = _run_bulk(data, o, camera_active=active)
    assert_line_search_matches(s, ref, prefix)
    np.testing.assert_array_equal(cams[[0, 5]], data.cameras[[0, 5]])


@pytest.mark.parametrize("order", ["ceres", "eigen"])
def test_line_search_quaternion_cameras_match_oracle(order):
    print(f"line search quaternion ba4 {order}", flush=True)
    base = gen.ba_config("ba4")
    data = gen.with_quaternion_cameras(base, order, scale=np.linspace(0.8, 1.3, base.num_cameras))
    problem, rots, rests, pts = quaternion_problem(data, order)
    o = _options(max_num_iterations=12, max_num_line_search_step_size_iterations=60)
    s = tc.solve(o, problem)
    model = oba.BAModel(data)
    ref, (rc, rp) = line_search_solve(model, _oracle_options(o))
    prefix = stable_prefix(model, _oracle_options(o), ref)
    assert_line_search_matches(s, ref, prefix)
    if prefix == len(ref.iterations):
        np.testing.assert_allclose(torch.stack(pts).cpu().numpy(), rp, rtol=1e-6, atol=1e-8)


def test_line_search_refuses_what_it_does_not_cover():
    data = gen.make_curve_batch(num_problems=1, seed=3)
    m = torch.zeros(1, dtype=torch.float64)
    c = torch.zeros(1, dtype=torch.float64)
    problem = tc.Problem()
    problem.AddResidualBlock(tc.ExpCurveResidual(float(data.x[0]), float(data.y[0])), None, [m, c])
    with pytest.raises(NotImplementedError, match="bundle adjustment only"):
        tc.solve(_options(), problem)
    big = gen.ba_config("ba16")  # n = 9 * 16 + 3 * 2000 > CT_LS_MAX_BFGS_N
    bp, *_ = ba_problem(big)
    with pytest.raises(tc.NativeLibraryError, match="dense inverse Hessian"):
        tc.solve(_options(line_search_direction_type=tc.LineSearchDirectionType.BFGS), bp)
