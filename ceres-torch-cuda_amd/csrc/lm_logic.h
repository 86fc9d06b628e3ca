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
  // solver.py:816-819
  st->radius = st->radius / st->decrease_factor;
  st->decrease_factor *= 2.0;
  return st->radius;
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
  st->num_linear_solves += 1;
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
    radius = lm_accept_radius(st, rho);
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
  } else if (st->radius <= st->min_radius) {
    lm_terminate(st, CT_TERM_CONVERGENCE, CT_REASON_MIN_RADIUS);
    return;
  }
  st->iteration += 1;
}

}  // namespace ct
