"""The LINE_SEARCH minimizer's oracle (oracle/line_search.py), pinned on the CPU.

Pins, each chosen so that a plausible mistake fails it:

* interpolation: the quadratic step is the exact minimiser of a quadratic through f(0),
  f'(0) and f(t); the cubic step is the exact local minimiser of a cubic sampled at two
  trial points (line_search.py:37-81); the clamp to [0.5, 0.6] x step with the default
  contractions (line_search.py:31-35);
* directions: the LBFGS two-loop recursion equals -H g for the inverse Hessian built by
  the same pairs through the dense BFGS update from gamma I or I (Nocedal & Wright,
  Algorithm 7.4 against (6.17)); the three nonlinear CG betas coincide with linear CG's
  on a quadratic under exact line searches;
* the loop: the reference's own known answers, Rosenbrock from (-1.2, 1) to (1, 1)
  within 1e-4 for LBFGS, BFGS and Polak-Ribiere CG, and for the three interpolations
  (tests/test_problem_solver.py:307-380).
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import line_search as ols


def test_quadratic_interpolation_is_the_minimiser_of_the_parabola():
    a, b, c = 3.0, -4.0, 2.0  # f(t) = a t^2 + b t + c, minimum at 2/3
    f = lambda t: a * t * t + b * t + c  # noqa: E731
    t = ols.quadratic_interpolation_step(1.5, f(0.0), f(1.5), b)
    assert t == pytest.approx(-b / (2 * a), rel=1e-14)
    # a parabola opening downwards, or no decrease information: bisection
    assert ols.quadratic_interpolation_step(1.0, 0.0, -5.0, -1.0) == 0.5


def test_cubic_interpolation_is_the_local_minimiser_of_the_cubic():
    c3, c2, d, f0 = 0.7, -0.4, -1.3, 2.0  # f'(0) = d < 0
    f = lambda t: c3 * t ** 3 + c2 * t ** 2 + d * t + f0  # noqa: E731
    t1, t2 = 2.0, 1.4  # previous and current trial
    t = ols.cubic_interpolation_step(t2, f0, f(t2), d, t1, f(t1))
    roots = np.roots([3 * c3, 2 * c2, d])
    local_min = min(r.real for r in roots if abs(r.imag) < 1e-12 and 6 * c3 * r.real + 2 * c2 > 0)
    assert 0 < local_min < t2
    assert t == pytest.approx(local_min, rel=1e-12)
    # without a previous trial it is the quadratic step
    assert ols.cubic_interpolation_step(t2, f0, f(t2), d, None, None) == \
        ols.quadratic_interpolation_step(t2, f0, f(t2), d)


def test_next_step_size_clamps_to_the_contraction_interval():
    o = ols.LineSearchOptions()  # max contraction 1e-3, min contraction 0.6
    # the quadratic step here is tiny: clamped up to 0.5 x step
    assert ols.next_step_size(o, 1.0, 0.0, 1e6, -1.0, None, None) == 0.5
    # and a step above 0.6 x step is clamped down
    assert ols.next_step_size(o, 1.0, 0.0, -0.9, -1.0, None, None) == pytest.approx(0.6)
    b = ols.LineSearchOptions(interpolation="bisection")
    assert ols.next_step_size(b, 2.0, 0.0, 1.0, -1.0, None, None) == 1.0


@pytest.mark.parametrize("scaling", [False, True])
def test_lbfgs_two_loop_equals_dense_bfgs_updates(scaling):
    rng = np.random.default_rng(0)
    n = 7
    A = rng.normal(size=(n, n))
    A = A @ A.T + n * np.eye(n)  # s, y pairs of a convex quadratic: y = A s, positive curvature
    s_hist = [rng.normal(size=n) for _ in range(4)]
    y_hist = [A @ s for s in s_hist]
    g = rng.normal(size=n)
    s_last, y_last = s_hist[-1], y_hist[-1]
    H = (s_last @ y_last) / (y_last @ y_last) * np.eye(n) if scaling else np.eye(n)
    for s, y in zip(s_hist, y_hist):
        H = ols.bfgs_update(H, s, y)
    d = ols.lbfgs_two_loop(g, s_hist, y_hist, scaling)
    np.testing.assert_allclose(d, -H @ g, rtol=1e-11, atol=1e-12)


def test_nonlinear_cg_betas_agree_with_linear_cg_on_a_quadratic():
    rng = np.random.default_rng(1)
    n = 6
    A = rng.normal(size=(n, n))
    A = A @ A.T + n * np.eye(n)
    b = rng.normal(size=n)
    x = np.zeros(n)
    g = A @ x - b
    d = -g
    for _ in range(3):
        alpha = -(g @ d) / (d @ A @ d)  # exact line search
        x = x + alpha * d
        g_new = A @ x - b
        cg_beta = (g_new @ g_new) / (g @ g)  # linear CG (Fletcher-Reeves form)
        for kind in ("fletcher_reeves", "polak_ribiere", "hestenes_stiefel"):
            beta = ols.ncg_beta(ols.LineSearchOptions(ncg_type=kind), g_new, g, d)
            assert beta == pytest.approx(cg_beta, rel=1e-9), kind
        d = -g_new + cg_beta * d
        g = g_new


class _Rosenbrock:
    """tests/test_problem_solver.py:319-352: residuals (1 - x0, 10 (x1 - x0^2))."""

    class _Ev:
        def __init__(self, cost):
            self.cost = cost

    def x0(self):
        return np.array([-1.2, 1.0])

    def evaluate(self, x, jacobian=True):
        r = np.array([1.0 - x[0], 10.0 * (x[1] - x[0] ** 2)])
        J = np.array([[-1.0, 0.0], [-20.0 * x[0], 10.0]])
        return self._Ev(0.5 * float(r @ r)), J.T @ r

    def apply(self, x, step):
        return x + step

    def state_norm(self, x):
        return float(np.linalg.norm(x))


_ROSENBROCK_TOLS = dict(max_num_iterations=250, function_tolerance=1e-12, gradient_tolerance=1e-8,
                        parameter_tolerance=1e-12)


@pytest.mark.parametrize("direction,ncg", [("lbfgs", "fletcher_reeves"), ("bfgs", "fletcher_reeves"),
                                           ("nonlinear_conjugate_gradient", "polak_ribiere")])
def test_rosenbrock_reference_known_answer_directions(direction, ncg):
    s, x = ols.line_search_solve(_Rosenbrock(), ols.LineSearchOptions(direction=direction, ncg_type=ncg,
                                                                      **_ROSENBROCK_TOLS))
    assert s.termination == "CONVERGENCE"
    np.testing.assert_allclose(x, [1.0, 1.0], atol=1e-4, rtol=1e-4)


@pytest.mark.parametrize("interpolation", ["bisection", "quadratic", "cubic"])
def test_rosenbrock_reference_known_answer_interpolations(interpolation):
    s, x = ols.line_search_solve(_Rosenbrock(), ols.LineSearchOptions(interpolation=interpolation,
                                                                      **_ROSENBROCK_TOLS))
    assert s.termination == "CONVERGENCE"
    np.testing.assert_allclose(x, [1.0, 1.0], atol=1e-4, rtol=1e-4)
    successful = [it for it in s.iterations if it["step_is_successful"]]
    assert successful and all(it["line_search_function_evaluations"] >= it["line_search_iterations"]
                              for it in successful)


def test_counters_add_up():
    s, _ = ols.line_search_solve(_Rosenbrock(), ols.LineSearchOptions(**_ROSENBROCK_TOLS))
    assert s.num_line_search_steps == sum(it["line_search_iterations"] for it in s.iterations)
    assert s.num_jacobian_evaluations == len(s.iterations) + s.num_line_search_function_evaluations
    assert s.num_successful_steps == sum(it["step_is_successful"] for it in s.iterations)
