// Device-side structures of the bundle-adjustment LM solver.
#pragma once

#include "ct_common.h"

namespace ct {

// Levenberg-Marquardt state kept in device memory for the whole solve.  The
// fields follow solver.py:35-357 (radius state :805-820, step evaluator
// :823-861 with max_consecutive_nonmonotonic_steps = 0, counters of
// SolverSummary types.py:360-387).
struct LMState {
  // options
  double max_radius, min_radius, min_rel_dec, min_diag, max_diag;
  double ftol, gtol, ptol, eta;
  int32_t max_iter, max_invalid, jacobi_scaling, min_lin_iter, max_lin_iter, refine_iters;
  // trust region
  double radius, decrease_factor;
  // costs
  double cost;          // cost at x (this iteration)
  double initial_cost, best_cost, final_cost, prev_cost;
  double ev_current;    // step evaluator current cost (== reference cost, monotonic)
  double gnorm2, gmax;
  double cand_cost, model_decrease, step_norm2, state_norm2;
  // progress
  int32_t iteration;    // iteration being executed (0 = initial evaluation)
  int32_t done;
  int32_t skip_step;    // iteration 0: evaluation only
  int32_t cur;          // ping-pong index of the current parameters
  int32_t pending_grad, pending_radius;
  int32_t consecutive_invalid, has_prev;
  int32_t termination, reason;
  int32_t num_records, num_successful, num_unsuccessful;
  int32_t num_residual_evals, num_jacobian_evals, num_linear_solves;
  int32_t step_invalid;     // set when dx is not finite / factorisation failed
  int32_t linear_iterations;
  int32_t strategy;         // ct_trust_region_strategy
  int32_t dogleg_type;      // ct_dogleg_type
  int32_t precond;          // ct_preconditioner
  double dl_cgn, dl_cg;     // DOGLEG step = dl_cgn * (Gauss-Newton step) + dl_cg * g
  double max_time;          // seconds of device clock before NO_CONVERGENCE
  int32_t pending_time;     // the time test after an accepted step runs with the deferred checks
  int32_t poison_left;      // test hook: steps still to be overwritten with NaN (debug_invalid_steps)
  int32_t poison;           // this iteration's step is overwritten with NaN
  int32_t split_blocks;     // SCHUR_JACOBI blocks of a camera split into rotation (3) and the rest (6)
  uint64_t t0;              // wall clock at start
  uint64_t t_end;           // wall clock at termination
};

// Read-only view of the problem arrays (see ct_ba_problem).
struct BAView {
  int32_t C, P, N, B;
  int64_t T;
  int32_t cam_model;   // ct_camera_model
  int32_t cam_stride;  // stored values per camera: 9 or 10
  const int32_t* __restrict__ obs_camera;
  const int32_t* __restrict__ obs_point;
  const int32_t* __restrict__ point_offsets;
  const double* __restrict__ obs_xy;
  const int32_t* __restrict__ camera_offsets;
  const int32_t* __restrict__ camera_obs;
  const int32_t* __restrict__ block_cameras;
  const int32_t* __restrict__ block_offsets;
  const int32_t* __restrict__ block_terms;
  const uint8_t* __restrict__ camera_mask;  // may be null: all free
  const uint8_t* __restrict__ point_mask;
  ct_loss loss;
};

// One unit of camera-block assembly work: a bounded slice of one camera-pair
// block.  kind 0: Schur terms [begin, end) of block_terms; kind 1: camera
// observations [begin, end) of camera_obs (diagonal blocks: U, rhs, g_c and
// the column norms).
struct BlockItem {
  int32_t block, kind, begin, end;
};
constexpr int kItemMaxWork = 224;  // terms or observations per item (7 lane groups x 32)

// One record per observation holds what the camera-pair assembly gathers for
// both observations of a term (camera Jacobian and P), so a term touches two
// records instead of four separate rows.
constexpr int kJPStride = 24;

// Per-observation and per-point buffers written by the evaluation kernel.
template <typename real>
struct BABuffers {
  real* JP;        // [N][kJPStride] per observation: corrected camera Jacobian [2][9] (masked),
                   //                then P = Jp L^-T [2][3] with L the Cholesky factor of damped V_p
  real* Jp;        // [N][2][3] corrected point Jacobian (masked)
  real* e;         // [N][2]    P q - r
  double* r;       // [N][2]    corrected residual (FP64)
  real* Linv;      // [P][6]    L^-1, lower triangle row-major
  real* q;         // [P][3]    L^-1 g_p
  double* gp;      // [P][3]    gradient point block (FP64 accumulation)
  real* item_S;    // [items][81] partial 9x9 camera blocks
};

struct BAWork {
  double* xcam;       // [2][C*9] ping-pong parameters
  double* xpt;        // [2][P*3]
  double* S;          // [n*n], n = 9C
  double* rhs;        // [n]
  double* dxc;        // [n]
  double* dxp;        // [P*3]
  double* part_cost;  // [Gobs]   cost at x
  double* part_eval;  // [Gpt][2] |g_p|^2, max|g_p|
  double* part_step;  // [Gstep][3] |dx|^2, |x'|^2, non-finite count
  double* part_cand;  // [Gobs][2] candidate cost, model decrease
  double* part_dogleg;  // [Gpt][6] DOGLEG dot products (see k_dogleg_partials)
  double* records;    // [(max_iter+2)][CT_REC_FIELDS]
  LMState* state;
  // damping and gradient kept for the refinement residual
  double* Dp;         // [P*3] point damping
  double* Dc;         // [n]   camera damping
  double* gc;         // [n]   camera gradient
  // mixed-precision refinement buffers
  double* res_c;      // [n]   reduced right-hand side of the correction
  double* res_p;      // [P*3] L^-1 rho_p
  double* u_obs;      // [N*4] J dx and P L^-1 rho_p per observation
  double* dxc_corr;   // [n]   camera correction
  // camera-block work items (built at plan creation)
  const BlockItem* items;            // [num_items], grouped by block
  const int32_t* block_item_offsets; // [B+1]
  const int32_t* item_perm;          // [num_items] launch slot -> item (locality order)
  double* item_vec;   // [items][27] rhs, g_c, column norms (kind 1 items)
  // whole points packed into waves of at most 64 observations (fused evaluation)
  const int32_t* point_groups;       // [num_groups + 1] first point of each group
  // dense solve scratch
  double* solve_tmp;  // [2n]
  unsigned int* pcg_bar;  // PCG solve epoch (null for dense solves), advanced by lm_begin
  int32_t* status;    // [4] device status words
};

constexpr int kPointThreads = 256;

__device__ __forceinline__ double loss_rho(const ct_loss& L, double s, double* rho1, double* rho2) {
  double r0, r1, r2;
  switch (L.type) {
    case CT_LOSS_HUBER: {
      const double a = L.a, b = a * a;
      if (s <= b) {
        r0 = s; r1 = 1.0; r2 = 0.0;
      } else {
        const double ss = sqrt(fmax(s, 2.2250738585072014e-308));
        r0 = 2.0 * a * ss - b;
        r1 = a / ss;
        r2 = -0.5 * a / (ss * s);
      }
      break;
    }
    case CT_LOSS_SOFTLONE: {
      const double b = L.a * L.a, z = 1.0 + s / b, sz = sqrt(z);
      r0 = 2.0 * b * (sz - 1.0); r1 = 1.0 / sz; r2 = -0.5 / (b * z * sz);
      break;
    }
    case CT_LOSS_CAUCHY: {
      const double b = L.a * L.a, z = 1.0 + s / b;
      r0 = b * log(z); r1 = 1.0 / z; r2 = -1.0 / (b * z * z);
      break;
    }
    case CT_LOSS_ARCTAN: {
      const double a = L.a, z = s / a, d = 1.0 + z * z;
      r0 = a * atan(z); r1 = 1.0 / d; r2 = -2.0 * z / (a * d * d);
      break;
    }
    case CT_LOSS_TOLERANT: {
      const double a = L.a, b = L.b, z = (s - a) / b;
      const double x = -a / b;
      const double spx = x > 20.0 ? x : log1p(exp(x));
      const double spz = z > 20.0 ? z : log1p(exp(z));
      r1 = 1.0 / (1.0 + exp(-z));
      r0 = b * spz - b * spx;
      r2 = r1 * (1.0 - r1) / b;
      break;
    }
    case CT_LOSS_TUKEY: {
      const double a2 = L.a * L.a;
      if (s <= a2) {
        const double in = 1.0 - s / a2;
        r0 = (a2 / 3.0) * (1.0 - in * in * in); r1 = in * in; r2 = -2.0 * in / a2;
      } else {
        r0 = a2 / 3.0; r1 = 0.0; r2 = 0.0;
      }
      break;
    }
    default:
      r0 = s; r1 = 1.0; r2 = 0.0;
  }
  *rho1 = L.scale * r1;
  *rho2 = L.scale * r2;
  return L.scale * r0;
}

}  // namespace ct
