// Device-resident Levenberg-Marquardt bookkeeping (solver.py:35-357).
//
// An LM iteration t is the kernel sequence
//   eval (observations, points) -> camera_blocks -> lm_begin -> RCS solve
//   -> step (points) -> candidate (observations) -> lm_end
// lm_begin reduces the cost/gradient partials of the evaluation at x and runs
// the checks that solver.py performs at the top of an iteration and, deferred
// by one evaluation, the gradient and radius checks that follow an accepted
// step (solver.py:315-341) - the evaluation after an accepted step is the
// evaluation at the top of the next iteration, so it is computed once.
// lm_end reduces the candidate partials and applies solver.py:168-294 and the
// parameter/function tests of :315-324.  Every kernel returns immediately
// once state->done is set, so a graph of several iterations can be replayed
// past convergence.
#include "ba_kernels.h"

namespace ct {

namespace {

constexpr int kLMThreads = 256;

__device__ double now_seconds(const LMState* st) {
  // wall_clock64 ticks at a constant 100 MHz on CDNA parts
  return (double)(wall_clock64() - st->t0) * 1e-8;
}

__device__ void terminate(LMState* st, int type, int reason) {
  st->done = 1;
  st->termination = type;
  st->reason = reason;
  st->t_end = wall_clock64();
}

__device__ double lm_accept_radius(LMState* st, double rho) {
  // solver.py:810-814
  const double t = 2.0 * rho - 1.0;
  const double denom = fmax(1.0 / 3.0, 1.0 - t * t * t);
  st->radius = fmin(st->max_radius, st->radius / denom);
  st->decrease_factor = 2.0;
  return st->radius;
}

__device__ double lm_reject_radius(LMState* st) {
  // solver.py:816-819
  st->radius = st->radius / st->decrease_factor;
  st->decrease_factor *= 2.0;
  return st->radius;
}

__global__ void k_lm_init(LMInit init, BAWork w, const double* __restrict__ ucams, const double* __restrict__ upts) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int stride = gridDim.x * blockDim.x;
  for (int i = t; i < init.n_cam_params; i += stride) w.xcam[i] = ucams[i];
  for (int i = t; i < init.n_pt_params; i += stride) w.xpt[i] = upts[i];
  if (t == 0) {
    LMState* st = w.state;
    const ct_lm_options& o = init.opt;
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
}

__global__ __launch_bounds__(kLMThreads) void k_lm_begin(LMInit init, BAWork w) {
  __shared__ double red[kLMThreads / kWave];
  LMState* st = w.state;
  if (st->done) return;
  double c = 0.0, g2 = 0.0, gm = 0.0;
  for (int i = threadIdx.x; i < init.num_obs_blocks; i += kLMThreads) c += w.part_cost[i];
  for (int i = threadIdx.x; i < init.num_eval_blocks; i += kLMThreads) {
    g2 += w.part_eval[2 * i];
    gm = fmax(gm, w.part_eval[2 * i + 1]);
  }
  for (int i = threadIdx.x; i < init.num_cameras; i += kLMThreads) {
    g2 += w.part_cam[2 * i];
    gm = fmax(gm, w.part_cam[2 * i + 1]);
  }
  const double cost = block_sum<kLMThreads>(c, red);
  const double gnorm2 = block_sum<kLMThreads>(g2, red);
  const double gmax = block_max<kLMThreads>(gm, red);
  if (threadIdx.x != 0) return;
  if (w.pcg_bar) *w.pcg_bar = 0u;  // the PCG kernel of this iteration counts arrivals from 0
  const int t = st->iteration;
  st->step_invalid = 0;
  st->linear_iterations = 0;
  if (st->pending_grad) {
    // the evaluation after an accepted step (solver.py:325-336)
    st->pending_grad = 0;
    st->num_jacobian_evals += 1;
    if (gmax <= st->gtol) {
      st->final_cost = cost;
      terminate(st, CT_TERM_CONVERGENCE, CT_REASON_GRADIENT);
      return;
    }
  }
  if (st->pending_radius) {
    st->pending_radius = 0;
    if (st->radius <= st->min_radius) {  // solver.py:338-341
      terminate(st, CT_TERM_CONVERGENCE, CT_REASON_MIN_RADIUS);
      return;
    }
  }
  if (t > st->max_iter) {  // for-else of solver.py:77, :346-348
    terminate(st, CT_TERM_NO_CONVERGENCE, CT_REASON_MAX_ITERATIONS);
    return;
  }
  st->num_jacobian_evals += 1;
  st->cost = cost;
  st->gnorm2 = gnorm2;
  st->gmax = gmax;
  double* rec = w.records + (size_t)t * CT_REC_FIELDS;
  for (int k = 0; k < CT_REC_FIELDS; ++k) rec[k] = 0.0;
  rec[CT_REC_ITERATION] = t;
  rec[CT_REC_COST] = cost;
  rec[CT_REC_GRADIENT_MAX_NORM] = gmax;
  rec[CT_REC_GRADIENT_NORM] = sqrt(gnorm2);
  rec[CT_REC_TRUST_REGION_RADIUS] = st->radius;
  rec[CT_REC_STEP_IS_VALID] = 1.0;
  rec[CT_REC_STEP_IS_SUCCESSFUL] = 1.0;
  rec[CT_REC_TIMESTAMP] = now_seconds(st);
  if (t == 0) {
    // solver.py:64-73 and :106-121
    st->initial_cost = st->best_cost = st->final_cost = st->ev_current = cost;
    st->num_records = 1;
    if (gmax <= st->gtol) {
      terminate(st, CT_TERM_CONVERGENCE, CT_REASON_GRADIENT);
    } else if (st->max_iter == 0) {
      terminate(st, CT_TERM_NO_CONVERGENCE, CT_REASON_MAX_ITERATIONS);
    } else {
      st->skip_step = 1;
    }
    return;
  }
  st->skip_step = 0;
  st->num_linear_solves += 1;
}

__global__ __launch_bounds__(kLMThreads) void k_lm_end(LMInit init, BAWork w) {
  __shared__ double red[kLMThreads / kWave];
  LMState* st = w.state;
  // every thread reads the flags before thread 0 may change them
  const int done = st->done;
  const int skip = st->skip_step;
  __syncthreads();
  if (done) return;
  if (skip) {
    if (threadIdx.x == 0) {
      st->skip_step = 0;
      st->iteration += 1;
    }
    return;
  }
  double a[5] = {0, 0, 0, 0, 0};
  for (int i = threadIdx.x; i < init.num_obs_blocks; i += kLMThreads) {
    a[0] += w.part_cand[2 * i];
    a[1] += w.part_cand[2 * i + 1];
  }
  for (int i = threadIdx.x; i < init.num_step_blocks; i += kLMThreads) {
    a[2] += w.part_step[3 * i];
    a[3] += w.part_step[3 * i + 1];
    a[4] += w.part_step[3 * i + 2];
  }
  double s[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) s[k] = block_sum<kLMThreads>(a[k], red);
  if (threadIdx.x != 0) return;
  const int t = st->iteration;
  double* rec = w.records + (size_t)t * CT_REC_FIELDS;
  const double cand = s[0], md = s[1], dx2 = s[2], x2 = s[3];
  const bool finite = s[4] == 0.0 && !st->step_invalid;
  rec[CT_REC_LINEAR_SOLVER_ITERATIONS] = st->linear_iterations;
  if (!finite) {
    // solver.py:168-190
    st->consecutive_invalid += 1;
    const double radius = lm_reject_radius(st);
    rec[CT_REC_STEP_IS_VALID] = 0.0;
    rec[CT_REC_STEP_IS_SUCCESSFUL] = 0.0;
    rec[CT_REC_TRUST_REGION_RADIUS] = radius;
    rec[CT_REC_TIMESTAMP] = now_seconds(st);
    st->num_unsuccessful += 1;
    if (st->consecutive_invalid >= st->max_invalid) {
      terminate(st, CT_TERM_FAILURE, CT_REASON_INVALID_STEPS);
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
  rec[CT_REC_STEP_NORM] = step_norm;
  rec[CT_REC_COST_CHANGE] = st->cost - cand;
  rec[CT_REC_RELATIVE_DECREASE] = rho;
  rec[CT_REC_STEP_IS_VALID] = valid ? 1.0 : 0.0;
  if (accepted) {
    st->consecutive_invalid = 0;
    st->num_successful += 1;
    st->ev_current = cand;
    st->cur = 1 - st->cur;
    if (cand < st->best_cost) st->best_cost = cand;
    st->final_cost = st->best_cost;
    rec[CT_REC_TRUST_REGION_RADIUS] = lm_accept_radius(st, rho);
    rec[CT_REC_STEP_IS_SUCCESSFUL] = 1.0;
  } else {
    st->num_unsuccessful += 1;
    rec[CT_REC_TRUST_REGION_RADIUS] = lm_reject_radius(st);
    rec[CT_REC_STEP_IS_SUCCESSFUL] = 0.0;
  }
  rec[CT_REC_TIMESTAMP] = now_seconds(st);
  st->num_records += 1;
  if (accepted) {
    // solver.py:315-324
    if (step_norm <= st->ptol * (sqrt(x2) + st->ptol)) {
      terminate(st, CT_TERM_CONVERGENCE, CT_REASON_PARAMETER);
      return;
    }
    if (st->has_prev && fabs(st->prev_cost - cand) <= st->ftol * fmax(st->prev_cost, 1.0)) {
      terminate(st, CT_TERM_CONVERGENCE, CT_REASON_FUNCTION);
      return;
    }
    st->prev_cost = cand;
    st->has_prev = 1;
    st->pending_grad = 1;
    st->pending_radius = 1;
  } else if (st->radius <= st->min_radius) {
    terminate(st, CT_TERM_CONVERGENCE, CT_REASON_MIN_RADIUS);
    return;
  }
  st->iteration += 1;
}

__global__ void k_lm_finalize(LMInit init, BAWork w, double* __restrict__ ucams, double* __restrict__ upts) {
  const LMState* st = w.state;
  const int cur = st->cur;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int stride = gridDim.x * blockDim.x;
  for (int i = t; i < init.n_cam_params; i += stride) ucams[i] = w.xcam[(size_t)cur * init.n_cam_params + i];
  for (int i = t; i < init.n_pt_params; i += stride) upts[i] = w.xpt[(size_t)cur * init.n_pt_params + i];
}

}  // namespace

int launch_lm_init(const LMInit& init, const BAWork& w, const double* ucams, const double* upts, hipStream_t s) {
  const int work = init.n_pt_params > init.n_cam_params ? init.n_pt_params : init.n_cam_params;
  int grid = (work + 255) / 256;
  if (grid < 1) grid = 1;
  if (grid > 1024) grid = 1024;
  hipLaunchKernelGGL(k_lm_init, dim3(grid), dim3(256), 0, s, init, w, ucams, upts);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int launch_lm_begin(const LMInit& init, const BAWork& w, hipStream_t s) {
  hipLaunchKernelGGL(k_lm_begin, dim3(1), dim3(kLMThreads), 0, s, init, w);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int launch_lm_end(const LMInit& init, const BAWork& w, hipStream_t s) {
  hipLaunchKernelGGL(k_lm_end, dim3(1), dim3(kLMThreads), 0, s, init, w);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int launch_lm_finalize(const LMInit& init, const BAWork& w, double* ucams, double* upts, hipStream_t s) {
  const int work = init.n_pt_params > init.n_cam_params ? init.n_pt_params : init.n_cam_params;
  int grid = (work + 255) / 256;
  if (grid < 1) grid = 1;
  if (grid > 1024) grid = 1024;
  hipLaunchKernelGGL(k_lm_finalize, dim3(grid), dim3(256), 0, s, init, w, ucams, upts);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

}  // namespace ct
