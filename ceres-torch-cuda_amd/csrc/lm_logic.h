// Levenberg-Marquardt trust-region arithmetic on a device-resident LMState
// (solver.py:35-357), shared by the bundle-adjustment kernels and the batched
// small-problem kernels.  Every function runs in one thread.
//
// An iteration t is: evaluation at x -> lm_begin_update -> (t > 0) step and
// candidate -> lm_end_update (or lm_skip_update for the evaluation-only
// iteration 0).  lm_begin_update also runs, deferred by one evaluation, the
// gradient and radius checks that follow an accepted step (solver.py:315-341):
// the evaluation after an accepted step is the evaluation at the top of the
// next iteration, so it is computed once.
#pragma once

#include "ba_device.h"

namespace ct {

__device__ __forceinline__ double lm_now(const LMState* st) {
  // wall_clock64 ticks at a constant 100 MHz on CDNA parts
  return (double)(wall_clock64() - st->t0) * 1e-8;
}

__device__ __forceinline__ double damping_of(double colsq, double min_diag, double max_diag, double radius,
                                             int jacobi_scaling) {
  const double dg = jacobi_scaling ? colsq : 1.0;
  return fmin(fmax(dg, min_diag), max_diag) / fmax(radius, 2.220446049250313e-16);
}

// Diagonal added to the normal equations of one column: the LM damping
// (linear.py:539-551), or for DOGLEG none (the Gauss-Newton system J^T J),
// except a unit diagonal on the columns of constant parameters, which keeps
// their (zero) rows of the system nonsingular without touching any other.
__device__ __forceinline__ double column_damping(const LMState* st, double colsq, bool free_column) {
  if (st->strategy == CT_DOGLEG) return free_column ? 0.0 : 1.0;
  return damping_of(colsq, st->min_diag, st->max_diag, st->radius, st->jacobi_scaling);
}

__device__ __forceinline__ void lm_terminate(LMState* st, int type, int reason) {
  st->done = 1;
  st->termination = type;
  st->reason = reason;
  st->t_end = wall_clock64();
}

__device__ __forceinline__ double lm_accept_radius(LMState* st, double rho) {
  // solver.py:810-814
  const double t = 2.0 * rho - 1.0;
  const double denom = fmax(1.0 / 3.0, 1.0 - t * t * t);
  st->radius = fmin(st->max_radius, st->radius / denom);
  st->decrease_factor = 2.0;
  return st->radius;
}

__device__ __forceinline__ double lm_reject_radius(LMState* st) {
  if (st->strategy == CT_DOGLEG) {  // solver.py:876-883
    st->radius *= 0.25;
    return st->radius;
  }
  // solver.py:816-819
  st->radius = st->radius / st->decrease_factor;
  st->decrease_factor *= 2.0;
  return st->radius;
}

// Radius after an accepted step: solver.py:864-873 (LM) and :1051-1061 (DOGLEG).
__device__ __forceinline__ double lm_accepted_radius(LMState* st, double rho, double step_norm) {
  if (st->strategy != CT_DOGLEG) return lm_accept_radius(st, rho);
  if (rho > 0.75) {
    st->radius = fmin(st->max_radius, fmax(2.0 * st->radius, 3.0 * fmax(step_norm, 2.220446049250313e-16)));
  } else if (rho < 0.25) {
    st->radius = fmax(st->radius * 0.5, 2.2250738585072014e-308);
  }
  return st->radius;
}

// x = A^-1 b for a 2x2 system by Gaussian elimination with partial pivoting
// (the LU solve of linear.py:1366-1370); false if a pivot is zero.
__device__ __forceinline__ bool solve2(double a00, double a01, double a10, double a11, double b0, double b1,
                                       double* x0, double* x1) {
  if (fabs(a10) > fabs(a00)) {
    double t = a00; a00 = a10; a10 = t;
    t = a01; a01 = a11; a11 = t;
    t = b0; b0 = b1; b1 = t;
  }
  if (a00 == 0.0) return false;
  const double l = a10 / a00;
  const double u11 = a11 - l * a01;
  if (u11 == 0.0) return false;
  *x1 = (b1 - l * b0) / u11;
  *x0 = (b0 - a01 * *x1) / a00;
  return true;
}

// Least-squares fallback of a singular 2x2 solve: the minimum-norm solution
// through the pseudo-inverse of the symmetric matrix [[a, b], [b, c]].
__device__ __forceinline__ void lstsq2_sym(double a, double b, double c, double g0, double g1, double* x0,
                                           double* x1) {
  const double tr = a + c;
  const double disc = sqrt(fmax(0.25 * (a - c) * (a - c) + b * b, 0.0));
  const double l1 = 0.5 * tr + disc, l2 = 0.5 * tr - disc;
  // eigenvector of l1
  double v0 = b, v1 = l1 - a;
  if (fabs(v0) + fabs(v1) == 0.0) { v0 = 1.0; v1 = 0.0; }
  const double nv = sqrt(v0 * v0 + v1 * v1);
  v0 /= nv; v1 /= nv;
  const double tol = 2.220446049250313e-16 * 2.0 * fmax(fabs(l1), fabs(l2));
  double y0 = 0.0, y1 = 0.0;
  const double p1 = v0 * g0 + v1 * g1, p2 = -v1 * g0 + v0 * g1;
  if (fabs(l1) > tol) { y0 += v0 * p1 / l1; y1 += v1 * p1 / l1; }
  if (fabs(l2) > tol) { y0 += -v1 * p2 / l2; y1 += v0 * p2 / l2; }
  *x0 = y0;
  *x1 = y1;
}

__device__ __forceinline__ void solve2_sym(double a, double b, double c, double g0, double g1, double* x0,
                                           double* x1) {
  if (!solve2(a, b, b, c, g0, g1, x0, x1)) lstsq2_sym(a, b, c, g0, g1, x0, x1);
}

// Minimiser of 1/2 y^T H y + g^T y over |y| <= radius for the 2x2 H = [[a, b],
// [b, c]]: linear.py:636-664 (bisection on the shift of H + shift I).
__device__ __forceinline__ void boundary_step2(double a, double b, double c, double g0, double g1, double radius,
                                               double* y0, double* y1) {
  double u0, u1;
  solve2_sym(a, b, c, -g0, -g1, &u0, &u1);
  if (sqrt(u0 * u0 + u1 * u1) <= radius) {
    *y0 = u0;
    *y1 = u1;
    return;
  }
  const double rv = fmax(radius, 2.2250738585072014e-308);
  double lo = 0.0, hi = 1.0;
  solve2_sym(a + hi, b, c + hi, -g0, -g1, &u0, &u1);
  for (int k = 0; k < 80; ++k) {
    if (sqrt(u0 * u0 + u1 * u1) <= rv) break;
    hi *= 2.0;
    solve2_sym(a + hi, b, c + hi, -g0, -g1, &u0, &u1);
  }
  for (int k = 0; k < 80; ++k) {
    const double mid = 0.5 * (lo + hi);
    solve2_sym(a + mid, b, c + mid, -g0, -g1, &u0, &u1);
    if (sqrt(u0 * u0 + u1 * u1) > rv) lo = mid;
    else hi = mid;
  }
  solve2_sym(a + hi, b, c + hi, -g0, -g1, &u0, &u1);
  const double yn = sqrt(u0 * u0 + u1 * u1);
  if (yn > rv) {
    const double sc = radius / fmax(yn, 2.220446049250313e-16);
    u0 *= sc;
    u1 *= sc;
  }
  *y0 = u0;
  *y1 = u1;
}

// The point where the path sd -> gn (sd = -alpha g) leaves the trust region
// (linear.py:578-583): step = sd + tau (gn - sd) = tau gn + (tau - 1) alpha g,
// tau the positive root of |sd + tau diff|^2 = radius^2.
__device__ __forceinline__ void dogleg_interpolation(double alpha, double g2, double n2, double ggn, double radius,
                                                     double* cgn, double* cg) {
  constexpr double eps = 2.220446049250313e-16;
  const double a = n2 + 2.0 * alpha * ggn + alpha * alpha * g2;   // |gn - sd|^2
  const double b = -2.0 * (alpha * ggn + alpha * alpha * g2);     // 2 sd . (gn - sd)
  const double c = alpha * alpha * g2 - radius * radius;          // |sd|^2 - radius^2
  const double tau = (-b + sqrt(fmax(b * b - 4.0 * a * c, 0.0))) / fmax(2.0 * a, eps);
  *cgn = tau;
  *cg = (tau - 1.0) * alpha;
}

// DOGLEG step coefficients (linear.py:554-608): the step is
// st->dl_cgn * gn + st->dl_cg * g, with gn the Gauss-Newton step and g the
// gradient.  d = {|g|^2, |gn|^2, g.gn, |J g|^2, |J gn|^2, (J g).(J gn)}, and
// r.(J g) = |g|^2, r.(J gn) = g.gn give the subspace problem's gradient.
__device__ __forceinline__ void lm_dogleg_coefficients(LMState* st, const double* d) {
  constexpr double eps = 2.220446049250313e-16;
  const double radius = st->radius;
  const double g2 = d[0], n2 = d[1], gn = d[2], jg2 = d[3], jn2 = d[4], jgjn = d[5];
  double cgn = 0.0, cg = 0.0;
  if (radius > 0.0) {
    const double gn_norm = sqrt(n2);
    if (gn_norm <= radius) {
      cgn = 1.0;
    } else {
      // steepest descent sd = -alpha g (linear.py:611-616)
      const double alpha = g2 / fmax(jg2, eps);
      const double sd_norm = alpha * sqrt(g2);
      if (sd_norm >= radius) {
        cg = -alpha * radius / fmax(sd_norm, eps);
      } else if (st->dogleg_type != CT_SUBSPACE_DOGLEG) {
        dogleg_interpolation(alpha, g2, n2, gn, radius, &cgn, &cg);
      } else {
        // orthonormal basis of (-g, gn) (linear.py:619-633), then the 2-D
        // trust-region problem in that basis (linear.py:600-608)
        const double G = sqrt(g2);
        const bool have1 = G > 100.0 * eps * fmax(G, 1.0);
        double s = 0.0, c2n2 = n2;
        if (have1) {
          s = -gn / G;                       // q1 . gn with q1 = -g / G
          c2n2 = fmax(n2 - s * s, 0.0);      // |gn - s q1|^2
        }
        const double c2n = sqrt(c2n2);
        const bool have2 = c2n > 100.0 * eps * fmax(sqrt(n2), 1.0);
        if (!have1 && !have2) {
          cgn = cg = 0.0;
        } else if (!(have1 && have2)) {
          // fewer than two basis vectors: the traditional dogleg (linear.py:599-601)
          dogleg_interpolation(alpha, g2, n2, gn, radius, &cgn, &cg);
        } else {
          // q2 = (gn + (s / G) g) / c2n; J q1 = -J g / G; J q2 = (J gn + (s / G) J g) / c2n
          const double k = s / G;
          const double h11 = jg2 / g2;
          const double h12 = -(jgjn + k * jg2) / (G * c2n);
          const double h22 = (jn2 + 2.0 * k * jgjn + k * k * jg2) / c2n2;
          // reduced gradient (J B)^T r: q1 . g = -G, q2 . g = (gn + k g2) / c2n
          const double r1 = -G, r2 = (gn + k * g2) / c2n;
          double y1, y2;
          boundary_step2(h11, h12, h22, r1, r2, radius, &y1, &y2);
          // step = y1 q1 + y2 q2
          cgn = y2 / c2n;
          cg = -y1 / G + y2 * k / c2n;
        }
      }
    }
  }
  st->dl_cgn = cgn;
  st->dl_cg = cg;
}

__device__ __forceinline__ void lm_state_init(LMState* st, const ct_lm_options& o) {
  st->max_radius = o.max_trust_region_radius;
  st->min_radius = o.min_trust_region_radius;
  st->min_rel_dec = o.min_relative_decrease;
  st->min_diag = o.min_lm_diagonal;
  st->max_diag = o.max_lm_diagonal;
  st->ftol = o.function_tolerance;
  st->gtol = o.gradient_tolerance;
  st->ptol = o.parameter_tolerance;
  st->eta = o.eta;
  st->max_iter = o.max_num_iterations;
  st->max_invalid = o.max_num_consecutive_invalid_steps;
  st->jacobi_scaling = o.jacobi_scaling;
  st->min_lin_iter = o.min_linear_solver_iterations;
  st->max_lin_iter = o.max_linear_solver_iterations;
  st->refine_iters = o.max_num_refinement_iterations;
  st->strategy = o.trust_region_strategy;
  st->dogleg_type = o.dogleg_type;
  st->precond = o.preconditioner;
  st->max_time = o.max_solver_time_in_seconds;
  st->pending_time = 0;
  st->poison_left = o.debug_invalid_steps > 0 ? o.debug_invalid_steps : 0;
  st->poison = 0;
  st->split_blocks = 0;
  st->dl_cgn = 1.0;
  st->dl_cg = 0.0;
  st->radius = o.initial_trust_region_radius;
  st->decrease_factor = 2.0;
  st->cost = st->initial_cost = st->best_cost = st->final_cost = st->prev_cost = 0.0;
  st->ev_current = 0.0;
  st->gnorm2 = st->gmax = 0.0;
  st->cand_cost = st->model_decrease = st->step_norm2 = st->state_norm2 = 0.0;
  st->iteration = 0;
  st->done = 0;
  st->skip_step = 0;
  st->cur = 0;
  st->pending_grad = st->pending_radius = 0;
  st->consecutive_invalid = 0;
  st->has_prev = 0;
  st->termination = CT_TERM_NO_CONVERGENCE;
  st->reason = CT_REASON_NONE;
  st->num_records = st->num_successful = st->num_unsuccessful = 0;
  st->num_residual_evals = st->num_jacobian_evals = st->num_linear_solves = 0;
  st->step_invalid = 0;
  st->linear_iterations = 0;
  st->t0 = wall_clock64();
  st->t_end = st->t0;
}

// The evaluation at the top of iteration st->iteration: cost and gradient
// norms at x.  records (may be null) is the [max_iter+2][CT_REC_FIELDS] array.
__device__ __forceinline__ void lm_begin_update(LMState* st, double cost, double gnorm2, double gmax,
                                                double* records) {
  const int t = st->iteration;
  st->step_invalid = 0;
  st->linear_iterations = 0;
  if (st->pending_grad) {
    // the evaluation after an accepted step (solver.py:325-336)
    st->pending_grad = 0;
    st->num_jacobian_evals += 1;
    if (gmax <= st->gtol) {
      st->final_cost = cost;
      lm_terminate(st, CT_TERM_CONVERGENCE, CT_REASON_GRADIENT);
      return;
    }
  }
  if (st->pending_radius) {
    st->pending_radius = 0;
    if (st->radius <= st->min_radius) {  // solver.py:338-341
      lm_terminate(st, CT_TERM_CONVERGENCE, CT_REASON_MIN_RADIUS);
      return;
    }
  }
  if (st->pending_time) {
    st->pending_time = 0;
    if (lm_now(st) >= st->max_time) {  // solver.py:342-345
      lm_terminate(st, CT_TERM_NO_CONVERGENCE, CT_REASON_MAX_TIME);
      return;
    }
  }
  if (t > st->max_iter) {  // for-else of solver.py:77, :346-348
    lm_terminate(st, CT_TERM_NO_CONVERGENCE, CT_REASON_MAX_ITERATIONS);
    return;
  }
  st->num_jacobian_evals += 1;
  st->cost = cost;
  st->gnorm2 = gnorm2;
  st->gmax = gmax;
  if (records) {
    double* rec = records + (size_t)t * CT_REC_FIELDS;
    for (int k = 0; k < CT_REC_FIELDS; ++k) rec[k] = 0.0;
    rec[CT_REC_ITERATION] = t;
    rec[CT_REC_COST] = cost;
    rec[CT_REC_GRADIENT_MAX_NORM] = gmax;
    rec[CT_REC_GRADIENT_NORM] = sqrt(gnorm2);
    rec[CT_REC_TRUST_REGION_RADIUS] = st->radius;
    rec[CT_REC_STEP_IS_VALID] = 1.0;
    rec[CT_REC_STEP_IS_SUCCESSFUL] = 1.0;
    rec[CT_REC_TIMESTAMP] = lm_now(st);
  }
  if (t == 0) {
    // solver.py:64-73 and :106-121
    st->initial_cost = st->best_cost = st->final_cost = st->ev_current = cost;
    st->num_records = 1;
    if (gmax <= st->gtol) {
      lm_terminate(st, CT_TERM_CONVERGENCE, CT_REASON_GRADIENT);
    } else if (st->max_iter == 0) {
      lm_terminate(st, CT_TERM_NO_CONVERGENCE, CT_REASON_MAX_ITERATIONS);
    } else {
      st->skip_step = 1;
    }
    return;
  }
  st->skip_step = 0;
  if (st->strategy != CT_DOGLEG) st->num_linear_solves += 1;  // solver.py:129-133 counts no linear solve
  st->poison = st->poison_left > 0;
  if (st->poison) st->poison_left -= 1;
}

// End of the evaluation-only iteration 0.
__device__ __forceinline__ void lm_skip_update(LMState* st) {
  st->skip_step = 0;
  st->iteration += 1;
}

// After the candidate evaluation: gain ratio, radius, accept/reject and the
// parameter / function tests (solver.py:168-294, 315-324).  nonfinite: the
// step had a non-finite entry.  On acceptance st->cur flips to the candidate.
__device__ __forceinline__ void lm_end_update(LMState* st, double cand, double md, double dx2, double x2,
                                              bool nonfinite, double* records) {
  const int t = st->iteration;
  double* rec = records ? records + (size_t)t * CT_REC_FIELDS : nullptr;
  const bool finite = !nonfinite && !st->step_invalid;
  if (rec) rec[CT_REC_LINEAR_SOLVER_ITERATIONS] = st->linear_iterations;
  if (!finite) {
    // solver.py:168-190
    st->consecutive_invalid += 1;
    const double radius = lm_reject_radius(st);
    if (rec) {
      rec[CT_REC_STEP_IS_VALID] = 0.0;
      rec[CT_REC_STEP_IS_SUCCESSFUL] = 0.0;
      rec[CT_REC_TRUST_REGION_RADIUS] = radius;
      rec[CT_REC_TIMESTAMP] = lm_now(st);
    }
    st->num_unsuccessful += 1;
    if (st->consecutive_invalid >= st->max_invalid) {
      lm_terminate(st, CT_TERM_FAILURE, CT_REASON_INVALID_STEPS);
      return;
    }
    st->num_records += 1;
    st->iteration += 1;
    return;
  }
  st->num_residual_evals += 1;
  st->cand_cost = cand;
  st->model_decrease = md > 0.0 ? md : 0.0;  // solver.py:798-802
  st->step_norm2 = dx2;
  st->state_norm2 = x2;
  const double mdc = st->model_decrease;
  const bool valid = mdc > 0.0;
  // step evaluator with no non-monotonic steps: reference cost == current
  const double rho = valid ? (st->ev_current - cand) / mdc : 0.0;
  const bool accepted = valid && rho > st->min_rel_dec;
  const double step_norm = sqrt(dx2);
  if (rec) {
    rec[CT_REC_STEP_NORM] = step_norm;
    rec[CT_REC_COST_CHANGE] = st->cost - cand;
    rec[CT_REC_RELATIVE_DECREASE] = rho;
    rec[CT_REC_STEP_IS_VALID] = valid ? 1.0 : 0.0;
  }
  double radius;
  if (accepted) {
    st->consecutive_invalid = 0;
    st->num_successful += 1;
    st->ev_current = cand;
    st->cur = 1 - st->cur;
    if (cand < st->best_cost) st->best_cost = cand;
    st->final_cost = st->best_cost;
    radius = lm_accepted_radius(st, rho, step_norm);
  } else {
    st->num_unsuccessful += 1;
    radius = lm_reject_radius(st);
  }
  if (rec) {
    rec[CT_REC_TRUST_REGION_RADIUS] = radius;
    rec[CT_REC_STEP_IS_SUCCESSFUL] = accepted ? 1.0 : 0.0;
    rec[CT_REC_TIMESTAMP] = lm_now(st);
  }
  st->num_records += 1;
  if (accepted) {
    // solver.py:315-324
    if (step_norm <= st->ptol * (sqrt(x2) + st->ptol)) {
      lm_terminate(st, CT_TERM_CONVERGENCE, CT_REASON_PARAMETER);
      return;
    }
    if (st->has_prev && fabs(st->prev_cost - cand) <= st->ftol * fmax(st->prev_cost, 1.0)) {
      lm_terminate(st, CT_TERM_CONVERGENCE, CT_REASON_FUNCTION);
      return;
    }
    st->prev_cost = cand;
    st->has_prev = 1;
    st->pending_grad = 1;
    st->pending_radius = 1;
    st->pending_time = 1;
  } else if (st->radius <= st->min_radius) {
    lm_terminate(st, CT_TERM_CONVERGENCE, CT_REASON_MIN_RADIUS);
    return;
  } else if (lm_now(st) >= st->max_time) {  // solver.py:342-345
    lm_terminate(st, CT_TERM_NO_CONVERGENCE, CT_REASON_MAX_TIME);
    return;
  }
  st->iteration += 1;
}

}  // namespace ct
