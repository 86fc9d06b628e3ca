"""The LINE_SEARCH minimizer in float64 numpy: restatement of the reference's
``_line_search_solve`` for a model.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows ``src/ceres_torch/``:

* ``solver.py:360-607``  ``_line_search_solve`` (the loop, restarts, counters, tests)
* ``solver.py:1078-1120`` ``_line_search_direction``, ``_nonlinear_conjugate_gradient_beta``
* ``solver.py:1123-1146`` ``_lbfgs_two_loop``
* ``solver.py:1149-1162`` ``_bfgs_update``, ``_has_positive_curvature``
* ``solver.py:1165-1175`` ``_gradient_converged``, ``_parameter_converged``
* ``line_search.py:9-81``  ``next_line_search_step_size`` and the interpolations

A model provides ``x0()``, ``evaluate(x, jacobian=True) -> (ev, g)`` (``ev.cost``
and the gradient in tangent coordinates), ``apply(x, step)`` (Plus from x) and
``state_norm(x)`` (the norm of every parameter value, constant blocks included).
No bounds, callbacks or time limit (the device path refuses them).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

_EPS = float(np.finfo(np.float64).eps)


@dataclass
class LineSearchOptions:
    """The SolverOptions fields the loop reads (types.py:164-198 defaults)."""

    max_num_iterations: int = 50
    direction: str = "lbfgs"                  # steepest_descent | nonlinear_conjugate_gradient | lbfgs | bfgs
    line_search_type: str = "wolfe"           # | armijo
    ncg_type: str = "fletcher_reeves"         # | polak_ribiere | hestenes_stiefel
    max_lbfgs_rank: int = 20
    use_approximate_eigenvalue_bfgs_scaling: bool = False
    interpolation: str = "cubic"              # | quadratic | bisection
    min_line_search_step_size: float = 1e-9
    sufficient_function_decrease: float = 1e-4
    max_step_contraction: float = 1e-3
    min_step_contraction: float = 0.6
    max_step_size_iterations: int = 20
    max_direction_restarts: int = 5
    sufficient_curvature_decrease: float = 0.9
    function_tolerance: float = 1e-6
    gradient_tolerance: float = 1e-10
    parameter_tolerance: float = 1e-8


# ----------------------------------------------------------------------------
# step-size interpolation (line_search.py:9-81)
# ----------------------------------------------------------------------------


def quadratic_interpolation_step(step, cost, candidate_cost, derivative):
    """Minimiser of the parabola through f(0) = cost, f'(0) = derivative, f(step) = candidate_cost."""
    denom = 2.0 * (candidate_cost - cost - derivative * step)
    if denom <= 0.0 or not math.isfinite(denom):
        return 0.5 * step
    t = -(derivative * step * step) / denom
    return t if math.isfinite(t) and t > 0.0 else 0.5 * step


def cubic_interpolation_step(step, cost, candidate_cost, derivative, previous_step, previous_candidate_cost):
    """Smallest positive root inside (0, step) of the derivative of the cubic through f(0), f'(0) and
    the last two trial points; the quadratic step where that cubic is not usable."""
    if previous_step is None or previous_candidate_cost is None or previous_step == step:
        return quadratic_interpolation_step(step, cost, candidate_cost, derivative)
    t1, t2 = previous_step, step
    y1 = previous_candidate_cost - cost - derivative * t1
    y2 = candidate_cost - cost - derivative * t2
    det = t1 * t1 * t2 * t2 * (t2 - t1)
    if det == 0.0 or not math.isfinite(det):
        return quadratic_interpolation_step(step, cost, candidate_cost, derivative)
    c2 = (y1 * t2 ** 3 - y2 * t1 ** 3) / det
    c3 = (y2 * t1 * t1 - y1 * t2 * t2) / det
    if not math.isfinite(c2) or not math.isfinite(c3) or abs(c3) < 1e-30:
        return quadratic_interpolation_step(step, cost, candidate_cost, derivative)
    disc = c2 * c2 - 3.0 * c3 * derivative
    if disc < 0.0:
        return quadratic_interpolation_step(step, cost, candidate_cost, derivative)
    roots = [(-c2 + math.sqrt(disc)) / (3.0 * c3), (-c2 - math.sqrt(disc)) / (3.0 * c3)]
    inside = [r for r in roots if math.isfinite(r) and 0.0 < r < step]
    return min(inside) if inside else quadratic_interpolation_step(step, cost, candidate_cost, derivative)


def next_step_size(opts: LineSearchOptions, step, cost, candidate_cost, derivative, previous_step,
                   previous_candidate_cost):
    """line_search.py:9-35: the interpolated step clamped to [lower, upper] times the current one."""
    if opts.interpolation == "bisection":
        t = 0.5 * step
    elif opts.interpolation == "quadratic":
        t = quadratic_interpolation_step(step, cost, candidate_cost, derivative)
    else:
        t = cubic_interpolation_step(step, cost, candidate_cost, derivative, previous_step, previous_candidate_cost)
    lower = step * max(0.5, min(opts.max_step_contraction, opts.min_step_contraction))
    upper = step * max(opts.max_step_contraction, opts.min_step_contraction)
    return min(max(t, lower), upper)


# ----------------------------------------------------------------------------
# search directions (solver.py:1078-1162)
# ----------------------------------------------------------------------------


def ncg_beta(opts: LineSearchOptions, g, g_prev, d_prev) -> float:
    if opts.ncg_type == "polak_ribiere":
        return float(g @ (g - g_prev)) / max(float(g_prev @ g_prev), _EPS)
    if opts.ncg_type == "hestenes_stiefel":
        y = g - g_prev
        return float(g @ y) / max(float(d_prev @ y), _EPS)
    return float(g @ g) / max(float(g_prev @ g_prev), _EPS)


def lbfgs_two_loop(g, s_hist, y_hist, scaling: bool):
    q = g.copy()
    alphas, rhos = [], []
    for s, y in reversed(list(zip(s_hist, y_hist))):
        rho = 1.0 / max(float(y @ s), _EPS)
        alpha = rho * float(s @ q)
        q = q - alpha * y
        alphas.append(alpha)
        rhos.append(rho)
    s, y = s_hist[-1], y_hist[-1]
    gamma = float(s @ y) / max(float(y @ y), _EPS)
    r = gamma * q if scaling else q
    for s, y, alpha, rho in zip(s_hist, y_hist, reversed(alphas), reversed(rhos)):
        beta = rho * float(y @ r)
        r = r + s * (alpha - beta)
    return -r


def has_positive_curvature(s, y) -> bool:
    return float(s @ y) > 10.0 * _EPS * float(np.linalg.norm(s)) * float(np.linalg.norm(y))


def bfgs_update(H, s, y):
    if not has_positive_curvature(s, y):
        return np.eye(H.shape[0])
    rho = 1.0 / float(y @ s)
    eye = np.eye(H.shape[0])
    return (eye - rho * np.outer(s, y)) @ H @ (eye - rho * np.outer(y, s)) + rho * np.outer(s, s)


def search_direction(opts: LineSearchOptions, g, g_prev, d_prev, s_hist, y_hist, H):
    if opts.direction == "steepest_descent":
        return -g
    if opts.direction == "nonlinear_conjugate_gradient" and g_prev is not None and d_prev is not None:
        return -g + max(ncg_beta(opts, g, g_prev, d_prev), 0.0) * d_prev
    if opts.direction == "lbfgs" and s_hist:
        return lbfgs_two_loop(g, s_hist, y_hist, opts.use_approximate_eigenvalue_bfgs_scaling)
    if opts.direction == "bfgs" and H is not None:
        return -(H @ g)
    return -g


# ----------------------------------------------------------------------------
# the loop (solver.py:360-607)
# ----------------------------------------------------------------------------


@dataclass
class LineSearchSummary:
    iterations: list = field(default_factory=list)
    initial_cost: float = 0.0
    final_cost: float = 0.0
    termination: str = "NO_CONVERGENCE"
    message: str = ""
    num_successful_steps: int = 0
    num_unsuccessful_steps: int = 0
    num_jacobian_evaluations: int = 0
    num_line_search_steps: int = 0
    num_line_search_function_evaluations: int = 0
    num_line_search_gradient_evaluations: int = 0
    num_line_search_direction_restarts: int = 0


def _iteration(k, cost, g):
    return dict(iteration=k, cost=cost, gradient_norm=float(np.linalg.norm(g)),
                gradient_max_norm=float(np.max(np.abs(g))) if g.size else 0.0, step_size=0.0,
                line_search_iterations=0, line_search_function_evaluations=0, line_search_gradient_evaluations=0,
                line_search_direction_restarts=0, step_norm=0.0, cost_change=0.0, step_is_successful=False)


def line_search_solve(model, opts: LineSearchOptions):
    """Returns (summary, x): x is the state after the last accepted step."""
    x = model.x0()
    s = LineSearchSummary()
    ev, g = model.evaluate(x, True)
    s.initial_cost = s.final_cost = ev.cost
    g_prev = d_prev = None
    s_hist, y_hist = [], []
    H = None
    for k in range(opts.max_num_iterations + 1):
        ev, g = model.evaluate(x, True)
        s.num_jacobian_evaluations += 1
        cost = ev.cost
        it = _iteration(k, cost, g)
        s.iterations.append(it)
        if it["gradient_max_norm"] <= opts.gradient_tolerance:
            s.termination, s.message, s.final_cost = "CONVERGENCE", "Gradient tolerance reached.", cost
            break
        if k == opts.max_num_iterations:
            s.termination, s.message = "NO_CONVERGENCE", "Maximum iterations reached."
            break
        if H is None or H.shape[0] != g.size:
            H = np.eye(g.size) if opts.direction == "bfgs" else np.zeros((0, 0))
        d = search_direction(opts, g, g_prev, d_prev, s_hist, y_hist, H if opts.direction == "bfgs" else None)
        restarts = 0
        dd = float(g @ d)
        if np.linalg.norm(d) <= _EPS or dd >= 0:
            restarts += 1
            d = -g
            dd = float(g @ d)
            s_hist.clear()
            y_hist.clear()
            H = np.eye(g.size) if opts.direction == "bfgs" else H
        if np.linalg.norm(d) <= _EPS or dd >= 0:
            s.termination, s.message, s.final_cost = "CONVERGENCE", "Projected gradient tolerance reached.", cost
            it["line_search_direction_restarts"] = restarts
            break
        directions = [d]
        fallback = -g
        if np.linalg.norm(fallback) > _EPS and not np.allclose(d, fallback):
            directions.append(fallback)
        accepted = False
        accepted_d = d
        g_acc = None
        trials = 0
        for ti, dt in enumerate(directions):
            if ti > 0:
                restarts += 1
                if restarts > opts.max_direction_restarts:
                    break
            t = 1.0
            der = float(g @ dt)
            t_prev = c_prev = None
            for ls in range(opts.max_step_size_iterations):
                xc = model.apply(x, t * dt)
                evc, gc = model.evaluate(xc, True)
                s.num_jacobian_evaluations += 1
                trials += 1
                cc = evc.cost
                armijo = cc <= cost + opts.sufficient_function_decrease * t * der
                if opts.line_search_type == "wolfe":
                    curvature = abs(float(gc @ dt)) <= opts.sufficient_curvature_decrease * abs(der)
                else:
                    curvature = True
                if armijo and curvature:
                    accepted, accepted_d, g_acc = True, dt, gc
                    x = xc
                    s.final_cost = cc
                    it.update(step_size=t, line_search_iterations=ls + 1, line_search_function_evaluations=trials,
                              line_search_gradient_evaluations=trials, line_search_direction_restarts=restarts,
                              step_norm=float(np.linalg.norm(t * dt)), cost_change=cost - cc,
                              step_is_successful=True)
                    s.num_line_search_steps += ls + 1
                    s.num_line_search_function_evaluations += trials
                    s.num_line_search_gradient_evaluations += trials
                    s.num_line_search_direction_restarts += restarts
                    s.num_successful_steps += 1
                    break
                nt = next_step_size(opts, t, cost, cc, der, t_prev, c_prev)
                t_prev, c_prev, t = t, cc, nt
                if t < opts.min_line_search_step_size:
                    break
            if accepted:
                break
        if not accepted:
            s.num_unsuccessful_steps += 1
            s.num_line_search_steps += trials
            s.num_line_search_function_evaluations += trials
            s.num_line_search_gradient_evaluations += trials
            s.num_line_search_direction_restarts += restarts
            it["line_search_direction_restarts"] = restarts
            s.termination, s.message = "NO_CONVERGENCE", "Line search failed to find a decreasing step."
            break
        step = it["step_size"] * accepted_d
        y = g_acc - g
        if has_positive_curvature(step, y):
            s_hist.append(step)
            y_hist.append(y)
            if len(s_hist) > opts.max_lbfgs_rank:
                s_hist.pop(0)
                y_hist.pop(0)
            if opts.direction == "bfgs":
                H = bfgs_update(H, step, y)
        else:
            s_hist.clear()
            y_hist.clear()
            H = np.eye(g.size) if opts.direction == "bfgs" else H
        g_prev, d_prev = g, accepted_d
        gmax = float(np.max(np.abs(g_acc))) if g_acc.size else 0.0
        if gmax <= opts.gradient_tolerance:
            s.termination, s.message = "CONVERGENCE", "Gradient tolerance reached."
            break
        if float(np.linalg.norm(step)) <= opts.parameter_tolerance * (model.state_norm(x) + opts.parameter_tolerance):
            s.termination, s.message = "CONVERGENCE", "Parameter tolerance reached."
            break
        if abs(cost - s.final_cost) <= opts.function_tolerance * max(cost, 1.0):
            s.termination, s.message = "CONVERGENCE", "Function tolerance reached."
            break
    return s, x
