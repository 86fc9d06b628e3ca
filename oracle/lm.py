"""Line-faithful restatement of the reference's trust-region (LM) minimiser.

Test infrastructure only.  ``trust_region_solve`` follows
``src/ceres_torch/solver.py:35-357`` statement by statement (line search,
inner iterations and callbacks, which the measured workloads do not use, are
left out).  Helpers restate:

* ``_LevenbergMarquardtRadiusState``      solver.py:805-820
* ``_accepted/_rejected_trust_region_radius``, ``_updated_trust_region_radius``
  (DOGLEG)                                 solver.py:864-883, 1051-1061
* ``dogleg_step`` (traditional and subspace) linear.py:554-664
* ``_TrustRegionStepEvaluator``           solver.py:823-861
* ``_model_decrease``                     solver.py:798-802 (in each model)
* ``jacobi_damping_from_jacobian``        linear.py:539-551
* ``_parameter_converged``/``_gradient_converged``  solver.py:1165-1175
* dense solves for the dense-faithful mode: ``schur_solve_dense``
  (linear.py:667-699), damped ``lstsq`` (linear.py:332-371 / 674-679),
  ``_refine_solution`` (linear.py:1331-1363).

A *model* supplies ``x0``, ``evaluate(x, jacobian) -> (ev, g)``,
``column_sq(ev)``, ``solve(ev, damping, opts) -> (step, iters)``,
``apply(x, step)``, ``model_decrease(ev, step)``, ``state_norm(x)`` and
``num_active_columns()``; ``ev.cost`` is the loss-applied cost.
"""
from __future__ import annotations

import copy
from dataclasses import dataclass, field

import numpy as np

_EPS = np.finfo(np.float64).eps


@dataclass
class OracleOptions:
    """The subset of ``SolverOptions`` (types.py:164-229) the LM path reads."""

    max_num_iterations: int = 50
    initial_trust_region_radius: float = 1e4
    max_trust_region_radius: float = 1e16
    min_trust_region_radius: float = 1e-32
    min_relative_decrease: float = 1e-3
    min_lm_diagonal: float = 1e-6
    max_lm_diagonal: float = 1e32
    max_num_consecutive_invalid_steps: int = 5
    function_tolerance: float = 1e-6
    gradient_tolerance: float = 1e-10
    parameter_tolerance: float = 1e-8
    eta: float = 1e-1
    min_linear_solver_iterations: int = 0
    max_linear_solver_iterations: int = 500
    jacobi_scaling: bool = True
    use_nonmonotonic_steps: bool = False
    max_consecutive_nonmonotonic_steps: int = 5
    use_mixed_precision_solves: bool = False
    max_num_refinement_iterations: int = 0
    max_solver_time_in_seconds: float = 1e9
    preconditioner_type: str = "jacobi"                      # ITERATIVE_SCHUR: "jacobi", "schur_jacobi", "identity"
    trust_region_strategy_type: str = "levenberg_marquardt"  # or "dogleg"
    dogleg_type: str = "traditional"                          # or "subspace"


class LMRadius:
    """solver.py:805-820."""

    def __init__(self, radius: float, max_radius: float):
        self.radius = radius
        self.max_radius = max_radius
        self.decrease_factor = 2.0

    def step_accepted(self, rho: float) -> float:
        denominator = max(1.0 / 3.0, 1.0 - (2.0 * rho - 1.0) ** 3)
        self.radius = min(self.max_radius, self.radius / denominator)
        self.decrease_factor = 2.0
        return self.radius

    def step_rejected(self) -> float:
        self.radius = self.radius / self.decrease_factor
        self.decrease_factor *= 2.0
        return self.radius


class StepEvaluator:
    """solver.py:823-861."""

    def __init__(self, initial_cost: float, max_nonmonotonic: int):
        self.max_nonmonotonic = max(0, int(max_nonmonotonic))
        self.minimum_cost = initial_cost
        self.current_cost = initial_cost
        self.reference_cost = initial_cost
        self.candidate_cost = initial_cost
        self.acc_reference = 0.0
        self.acc_candidate = 0.0
        self.num_nonmonotonic = 0

    def step_quality(self, cost: float, model_change: float) -> float:
        if model_change <= 0.0:
            return float("-inf")
        relative = (self.current_cost - cost) / model_change
        historical = (self.reference_cost - cost) / (self.acc_reference + model_change)
        return max(relative, historical)

    def step_accepted(self, cost: float, model_change: float) -> None:
        self.current_cost = cost
        self.acc_candidate += model_change
        self.acc_reference += model_change
        if self.current_cost < self.minimum_cost:
            self.minimum_cost = self.current_cost
            self.num_nonmonotonic = 0
            self.candidate_cost = self.current_cost
            self.acc_candidate = 0.0
        else:
            self.num_nonmonotonic += 1
            if self.current_cost > self.candidate_cost:
                self.candidate_cost = self.current_cost
                self.acc_candidate = 0.0
        if self.num_nonmonotonic == self.max_nonmonotonic:
            self.reference_cost = self.candidate_cost
            self.acc_reference = self.acc_candidate


@dataclass
class OracleSummary:
    initial_cost: float = -1.0
    final_cost: float = -1.0
    termination_type: str = "FAILURE"
    message: str = ""
    iterations: list = field(default_factory=list)
    num_successful_steps: int = 0
    num_unsuccessful_steps: int = 0
    num_residual_evaluations: int = 0
    num_jacobian_evaluations: int = 0
    num_linear_solves: int = 0


def updated_trust_region_radius(radius: float, rho: float, step_norm: float, max_radius: float) -> float:
    """solver.py:1051-1061 (the DOGLEG radius after an accepted step)."""
    if rho > 0.75:
        return min(max_radius, max(2.0 * radius, 3.0 * max(step_norm, _EPS)))
    if rho < 0.25:
        return max(radius * 0.5, np.finfo(np.float64).tiny)
    return radius


def _dogleg(opts: OracleOptions) -> bool:
    return opts.trust_region_strategy_type == "dogleg"


def _rejected_radius(opts: OracleOptions, radius: float, lm: "LMRadius") -> float:
    """solver.py:876-883."""
    return radius * 0.25 if _dogleg(opts) else lm.step_rejected()


def steepest_descent_step(jmul, g: np.ndarray) -> np.ndarray:
    """linear.py:611-616."""
    Jg = jmul(g)
    denom = max(float(Jg @ Jg), _EPS)
    alpha = float(g @ g) / denom
    return -alpha * g


def traditional_dogleg_step(jmul, g: np.ndarray, gn: np.ndarray, radius: float) -> np.ndarray:
    """linear.py:568-583 with the Gauss-Newton step gn = solve(J^T J, -g) given."""
    if np.linalg.norm(gn) <= radius:
        return gn
    sd = steepest_descent_step(jmul, g)
    sd_norm = float(np.linalg.norm(sd))
    if sd_norm >= radius:
        return (radius / max(sd_norm, _EPS)) * sd
    diff = gn - sd
    a = float(diff @ diff)
    b = 2.0 * float(sd @ diff)
    c = float(sd @ sd) - radius * radius
    tau = (-b + np.sqrt(max(b * b - 4.0 * a * c, 0.0))) / max(2.0 * a, _EPS)
    return sd + tau * diff


def orthonormal_basis(vectors) -> np.ndarray:
    """linear.py:619-633 (Gram-Schmidt, dropping nearly dependent vectors)."""
    basis = []
    for v in vectors:
        cand = np.array(v, dtype=np.float64).reshape(-1)
        cand_norm = float(np.linalg.norm(cand))
        for q in basis:
            cand = cand - float(q @ cand) * q
        nrm = float(np.linalg.norm(cand))
        if nrm > 100.0 * _EPS * max(cand_norm, 1.0):
            basis.append(cand / nrm)
    if not basis:
        return np.zeros((np.asarray(vectors[0]).size, 0))
    return np.stack(basis, axis=1)


def trust_region_boundary_step(H: np.ndarray, g: np.ndarray, radius: float) -> np.ndarray:
    """linear.py:636-664: minimiser of the quadratic model inside the ball, by bisection on the shift."""
    unconstrained = _solve_square(H, -g)
    if np.linalg.norm(unconstrained) <= radius:
        return unconstrained
    eye = np.eye(H.shape[0])
    radius_value = max(float(radius), np.finfo(np.float64).tiny)

    def solve_shifted(shift):
        return _solve_square(H + shift * eye, -g)

    low, high = 0.0, 1.0
    y = solve_shifted(high)
    for _ in range(80):
        if np.linalg.norm(y) <= radius_value:
            break
        high *= 2.0
        y = solve_shifted(high)
    for _ in range(80):
        mid = 0.5 * (low + high)
        y = solve_shifted(mid)
        if np.linalg.norm(y) > radius_value:
            low = mid
        else:
            high = mid
    y = solve_shifted(high)
    yn = float(np.linalg.norm(y))
    if yn > radius_value:
        return (radius / max(yn, _EPS)) * y
    return y


def subspace_dogleg_step(jmul, r: np.ndarray, g: np.ndarray, gn: np.ndarray, radius: float) -> np.ndarray:
    """linear.py:586-608."""
    if np.linalg.norm(gn) <= radius:
        return gn
    sd = steepest_descent_step(jmul, g)
    sd_norm = float(np.linalg.norm(sd))
    if sd_norm >= radius:
        return (radius / max(sd_norm, _EPS)) * sd
    basis = orthonormal_basis((-g, gn))
    if basis.shape[1] == 0:
        return np.zeros_like(g)
    if basis.shape[1] < 2:
        return traditional_dogleg_step(jmul, g, gn, radius)
    RJ = np.stack([jmul(basis[:, 0]), jmul(basis[:, 1])], axis=1)
    RH = RJ.T @ RJ
    rg = RJ.T @ r.reshape(-1)
    return basis @ trust_region_boundary_step(RH, rg, radius)


def dogleg_step(model, ev, g: np.ndarray, radius: float, opts: OracleOptions) -> np.ndarray:
    """linear.py:554-565: dogleg_step(J, r, radius) for a model that supplies
    ``gauss_newton_step(ev)`` (solve(J^T J, -g)), ``jmul(ev, v)`` (J v) and the
    residual vector ``ev.r``."""
    if radius <= 0.0 or g.size == 0:
        return np.zeros(g.size)
    gn = model.gauss_newton_step(ev)

    def jmul(v):
        return model.jmul(ev, v)

    if opts.dogleg_type == "subspace":
        return subspace_dogleg_step(jmul, np.asarray(ev.r).reshape(-1), g, gn, radius)
    return traditional_dogleg_step(jmul, g, gn, radius)


def damping_from_column_sq(col_sq: np.ndarray, radius: float, opts: OracleOptions) -> np.ndarray:
    """linear.py:539-551."""
    diag = col_sq if opts.jacobi_scaling else np.ones_like(col_sq)
    diag = np.clip(diag, opts.min_lm_diagonal, opts.max_lm_diagonal)
    return diag / max(radius, _EPS)


def trust_region_solve(model, opts: OracleOptions):
    """solver.py:35-357 for a model without bounds or callbacks.

    Returns (summary, x_final) where x_final is the best state (solver.py:672).
    """
    import time

    start = time.perf_counter()
    s = OracleSummary()
    x = model.x0()
    lm = LMRadius(opts.initial_trust_region_radius, opts.max_trust_region_radius)
    radius = lm.radius
    consecutive_invalid = 0
    previous_cost = None

    ev0, g0 = model.evaluate(x, True)
    s.initial_cost = ev0.cost
    s.final_cost = ev0.cost
    best_cost = ev0.cost
    best_x = copy.deepcopy(x)
    evaluator = StepEvaluator(ev0.cost, opts.max_consecutive_nonmonotonic_steps if opts.use_nonmonotonic_steps else 0)
    if g0 is None or g0.size == 0 and model.num_active_columns() == 0:
        s.termination_type = "CONVERGENCE"
        s.message = "No active parameters."
        s.final_cost = best_cost
        return s, best_x

    finished = False
    for iteration in range(opts.max_num_iterations + 1):
        ev, g = model.evaluate(x, True)
        s.num_jacobian_evaluations += 1
        cost = ev.cost
        grad_max = float(np.max(np.abs(g))) if g.size else 0.0
        rec = dict(
            iteration=iteration,
            step_is_valid=True,
            step_is_successful=True,
            cost=cost,
            gradient_norm=float(np.linalg.norm(g)),
            gradient_max_norm=grad_max,
            trust_region_radius=radius,
            step_norm=0.0,
            cost_change=0.0,
            relative_decrease=0.0,
            linear_solver_iterations=0,
        )
        if iteration == 0:
            s.iterations.append(rec)
            if grad_max <= opts.gradient_tolerance:
                s.termination_type, s.message = "CONVERGENCE", "Gradient tolerance reached."
                s.final_cost = cost
                finished = True
                break
            if opts.max_num_iterations == 0:
                s.termination_type, s.message = "NO_CONVERGENCE", "Maximum iterations reached."
                s.final_cost = cost
                finished = True
                break
            continue
        if model.num_active_columns() == 0:
            s.termination_type, s.message = "CONVERGENCE", "No active parameters."
            s.final_cost = cost
            finished = True
            break
        if _dogleg(opts):
            # solver.py:129-133: no damping, one "linear iteration", not a linear solve
            step = dogleg_step(model, ev, g, radius, opts)
            lin_iters = 1
        else:
            damping = damping_from_column_sq(model.column_sq(ev), radius, opts)
            step, lin_iters = model.solve(ev, damping, opts)
            s.num_linear_solves += 1
        if not np.all(np.isfinite(step)):
            consecutive_invalid += 1
            radius = _rejected_radius(opts, radius, lm)
            rec.update(step_is_valid=False, step_is_successful=False, linear_solver_iterations=lin_iters,
                       trust_region_radius=radius)
            s.num_unsuccessful_steps += 1
            if consecutive_invalid >= opts.max_num_consecutive_invalid_steps:
                s.termination_type, s.message = "FAILURE", "Too many invalid trust-region steps."
                finished = True
                break
            s.iterations.append(rec)
            continue
        x_cand = model.apply(x, step)
        cand_ev, _ = model.evaluate(x_cand, False)
        s.num_residual_evaluations += 1
        cand_cost = cand_ev.cost
        md = model.model_decrease(ev, step)
        valid = md > 0.0
        rho = evaluator.step_quality(cand_cost, md) if valid else 0.0
        accepted = valid and rho > opts.min_relative_decrease
        rec.update(step_norm=float(np.linalg.norm(step)), cost_change=cost - cand_cost, relative_decrease=float(rho),
                   step_is_valid=valid, linear_solver_iterations=lin_iters)
        if accepted:
            consecutive_invalid = 0
            s.num_successful_steps += 1
            evaluator.step_accepted(cand_cost, md)
            x = x_cand
            if cand_cost < best_cost:
                best_cost = cand_cost
                best_x = copy.deepcopy(x)
            s.final_cost = best_cost
            if _dogleg(opts):
                radius = updated_trust_region_radius(radius, rho, rec["step_norm"], opts.max_trust_region_radius)
            else:
                radius = lm.step_accepted(rho)
            rec.update(trust_region_radius=radius, step_is_successful=True)
        else:
            s.num_unsuccessful_steps += 1
            radius = _rejected_radius(opts, radius, lm)
            rec.update(trust_region_radius=radius, step_is_successful=False)
        s.iterations.append(rec)
        if accepted:
            state_norm = model.state_norm(x)
            if rec["step_norm"] <= opts.parameter_tolerance * (state_norm + opts.parameter_tolerance):
                s.termination_type, s.message = "CONVERGENCE", "Parameter tolerance reached."
                finished = True
                break
            if previous_cost is not None and abs(previous_cost - cand_cost) <= opts.function_tolerance * max(previous_cost, 1.0):
                s.termination_type, s.message = "CONVERGENCE", "Function tolerance reached."
                finished = True
                break
            previous_cost = cand_cost
            nxt, g_next = model.evaluate(x, True)
            s.num_jacobian_evaluations += 1
            if g_next is not None and (float(np.max(np.abs(g_next))) if g_next.size else 0.0) <= opts.gradient_tolerance:
                s.termination_type, s.message = "CONVERGENCE", "Gradient tolerance reached."
                s.final_cost = nxt.cost
                finished = True
                break
        if radius <= opts.min_trust_region_radius:
            s.termination_type, s.message = "CONVERGENCE", "Minimum trust-region radius reached."
            finished = True
            break
        if time.perf_counter() - start >= opts.max_solver_time_in_seconds:  # solver.py:342-345
            s.termination_type, s.message = "NO_CONVERGENCE", "Maximum solver time reached."
            finished = True
            break
    if not finished:
        s.termination_type, s.message = "NO_CONVERGENCE", "Maximum iterations reached."
    s.final_cost = best_cost
    return s, best_x


# ----------------------------------------------------------------------------
# dense algebra of the reference (for the dense-faithful mode)
# ----------------------------------------------------------------------------


def _solve_square(A: np.ndarray, b: np.ndarray) -> np.ndarray:
    """linear.py:1366-1370: LU solve, least squares if singular."""
    try:
        return np.linalg.solve(A, b)
    except np.linalg.LinAlgError:
        return np.linalg.lstsq(A, b, rcond=None)[0]


def damped_lstsq(A: np.ndarray, b: np.ndarray, damping: np.ndarray | None) -> np.ndarray:
    if damping is not None:
        D = np.diag(np.sqrt(np.maximum(damping, 0.0)))
        A = np.vstack([A, D])
        b = np.concatenate([b, np.zeros(D.shape[0])])
    return np.linalg.lstsq(A, b, rcond=None)[0]


def schur_solve_dense(A: np.ndarray, b: np.ndarray, num_eliminate: int, damping: np.ndarray | None) -> np.ndarray:
    """linear.py:667-699."""
    if num_eliminate <= 0 or num_eliminate >= A.shape[1]:
        return damped_lstsq(A, b, damping)
    H = A.T @ A
    rhs = A.T @ b
    if damping is not None:
        H = H + np.diag(damping)
    e = num_eliminate
    Haa, Hab, Hba, Hbb = H[:e, :e], H[:e, e:], H[e:, :e], H[e:, e:]
    ga, gb = rhs[:e], rhs[e:]
    Haa_inv_Hab = _solve_square(Haa, Hab)
    Haa_inv_ga = _solve_square(Haa, ga)
    S = Hbb - Hba @ Haa_inv_Hab
    xb = _solve_square(S, gb - Hba @ Haa_inv_ga)
    xa = Haa_inv_ga - Haa_inv_Hab @ xb
    return np.concatenate([xa, xb])


def refine_solution(A, b, x, damping, iterations: int) -> np.ndarray:
    """linear.py:1331-1363."""
    if iterations <= 0:
        return x
    if damping is not None:
        A = np.vstack([A, np.diag(np.sqrt(np.maximum(damping, 0.0)))])
        b = np.concatenate([b, np.zeros(A.shape[1])])
    for _ in range(iterations):
        res = b - A @ x
        if np.linalg.norm(res) <= 10.0 * _EPS * max(np.linalg.norm(b), 1.0):
            break
        x = x + np.linalg.lstsq(A, res, rcond=None)[0]
    return x


def retained_block_sizes(block_sizes, num_eliminate: int, num_columns: int) -> list:
    """linear.py:1269-1289: the parameter blocks left after eliminating the first columns."""
    sizes = list(block_sizes) if block_sizes else [1] * num_columns
    out, offset = [], 0
    for size in sizes:
        nxt = offset + size
        if nxt > num_eliminate:
            out.append(nxt - num_eliminate if offset < num_eliminate else size)
        offset = nxt
    if not out:
        raise ValueError("num_eliminate leaves no retained Schur columns")
    return out


def dense_linear_solve(J, b, damping, num_eliminate, solver, opts, block_sizes=None):
    """The reference's pure-torch branches of ``solve_linear_system``."""
    if J.shape[1] == 0:
        return np.zeros(0), 0
    if solver == "dense_schur":
        x = schur_solve_dense(J, b, num_eliminate, damping)
        x = refine_solution(J, b, x, damping, opts.max_num_refinement_iterations if opts.use_mixed_precision_solves else 0)
        return x, 1
    if solver == "iterative_schur" and num_eliminate > 0:
        from .ba import block_jacobi_pcg

        H = J.T @ J + np.diag(damping)
        rhs = J.T @ b
        e = num_eliminate
        Haa_inv_Hab = _solve_square(H[:e, :e], H[:e, e:])
        Haa_inv_ga = _solve_square(H[:e, :e], rhs[:e])
        S = H[e:, e:] - H[e:, :e] @ Haa_inv_Hab
        rhs_s = rhs[e:] - H[e:, :e] @ Haa_inv_ga
        blocks = retained_block_sizes(block_sizes, e, J.shape[1]) if block_sizes is not None else 9
        xb, iters, _ = block_jacobi_pcg(S, rhs_s, blocks, tol=opts.eta, max_iter=opts.max_linear_solver_iterations,
                                        preconditioner=opts.preconditioner_type,
                                        min_iter=opts.min_linear_solver_iterations)
        return np.concatenate([Haa_inv_ga - Haa_inv_Hab @ xb, xb]), iters
    # DENSE_QR: damped least squares (linear.py:359-371)
    return damped_lstsq(J, b, damping), 1
