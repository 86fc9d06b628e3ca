// Launchers of the bundle-adjustment, Cholesky and LM-state kernels.
#pragma once

#include "ba_device.h"

namespace ct {

template <typename real>
int launch_eval_obs(const BAView& v, const BABuffers<real>& b, const BAWork& w, hipStream_t s);
template <typename real>
int launch_eval_points(const BAView& v, const BABuffers<real>& b, const BAWork& w, hipStream_t s);
template <typename real>
int launch_camera_items(const BAView& v, const BABuffers<real>& b, const BAWork& w, int num_items, hipStream_t s);
// diagonal_only: write the C diagonal blocks of S (with rhs, g_c, D_c) and no others.
template <typename real>
int launch_camera_reduce(const BAView& v, const BABuffers<real>& b, const BAWork& w, bool diagonal_only,
                         hipStream_t s);
template <typename real>
int launch_step(const BAView& v, const BABuffers<real>& b, const BAWork& w, bool fused_backsub, hipStream_t s);
template <typename real>
int launch_candidate(const BAView& v, const BABuffers<real>& b, const BAWork& w, hipStream_t s);
// Fused evaluation over point groups (one wavefront per group of whole
// points, one lane per observation); replaces k_eval_obs + k_eval_points.
constexpr int kFusedThreads = 256;
template <typename real>
int launch_eval_fused(const BAView& v, const BABuffers<real>& b, const BAWork& w, int num_groups,
                      bool unobserved_points, hipStream_t s);
// Fused step + candidate over the same point groups (replaces k_step_points +
// k_candidate_obs); backsub = false takes dx_p from the refinement.
template <typename real>
int launch_step_candidate(const BAView& v, const BABuffers<real>& b, const BAWork& w, int num_groups, bool backsub,
                          hipStream_t s);
template <typename real>
int launch_backsub(const BAView& v, const BABuffers<real>& b, const BAWork& w, hipStream_t s);
// DOGLEG step (back-substitution of the Gauss-Newton step, dot products,
// coefficients, dx <- cgn gn + cg g); partial rows: dogleg_blocks(P).
template <typename real>
int launch_dogleg_step(const BAView& v, const BABuffers<real>& b, const BAWork& w, hipStream_t s);
inline int dogleg_blocks(int num_points) { return num_points > 0 ? (num_points + kPointThreads - 1) / kPointThreads : 1; }
// Test hook: NaN over the whole step when state->poison (ct_lm_options.debug_invalid_steps).
int launch_poison_step(const BAView& v, const BAWork& w, hipStream_t s);
// Mixed-precision refinement: `steps` refinement steps after the solve of
// dx_c, the back-substitution of dx_p included (steps <= 0: that alone).
template <typename real>
int launch_refinement(const BAView& v, const BABuffers<real>& b, const BAWork& w, int steps, hipStream_t s);
int launch_evaluate_only(const BAView& v, const double* cams, const double* pts, double* residuals, double* part,
                         hipStream_t s);
int launch_evaluate_jacobian(const BAView& v, const double* cams, const double* pts, const int32_t* row_of_obs,
                             const int64_t* cam_cols, const int64_t* pt_cols, int64_t ncols, int apply_loss,
                             double* jac, double* residuals, double* part, hipStream_t s);

// Covariance blocks from the factored reduced camera system (cov_kernels.hip):
// S^-1 = L^-T L^-1 formed densely in `work` (covariance_work_doubles(n)), then
// one block of (J^T J)^-1 per request; state->step_invalid is set when H is
// singular (a failed pivot of S or a non-finite point factor).
constexpr int kCovMaxN = 8192;  // the column solves keep one column of L^-1 in LDS
size_t covariance_work_doubles(int n);
struct CovLaunch {
  const double* L;               // [n][n] factor of S (lower triangle)
  int n;
  const double* JP;              // FP64 evaluation buffers
  const double* Linv;
  const int32_t* obs_camera;
  const int32_t* point_offsets;
  const uint8_t* camera_mask;    // may be null
  const uint8_t* point_mask;
  int num_points;
  const int32_t* requests;       // [R][4] device
  int64_t num_requests;
  double pivot_rtol;
  double* out;                   // [R][81]
  double* work;                  // covariance_work_doubles(n)
  int32_t* counts;               // [2] deficient point pivots, deficient pivots of S
  LMState* st;
};
int launch_covariance(const CovLaunch& c, hipStream_t s);
// Null directions of S (rebuilt unfactored) by pivoted Cholesky: counts[1].
int launch_covariance_rank(double* S, int n, const uint8_t* camera_mask, double pivot_rtol, int32_t* counts,
                           hipStream_t s);

// LM state machine (lm_kernels.hip)
struct LMInit {
  ct_lm_options opt;
  int num_obs_blocks;    // candidate partial rows
  int num_cost_blocks;   // cost partial rows written by the evaluation
  int num_eval_blocks;   // gradient partial rows written by the evaluation
  int num_step_blocks;   // step partial rows
  int num_cameras;
  int n_cam_params;
  int n_pt_params;
  int split_blocks;      // LMState::split_blocks
};
int launch_lm_init(const LMInit& init, const BAWork& w, const double* user_cams, const double* user_pts,
                   hipStream_t s);
int launch_lm_begin(const LMInit& init, const BAWork& w, hipStream_t s);
int launch_lm_end(const LMInit& init, const BAWork& w, hipStream_t s);
int launch_lm_finalize(const LMInit& init, const BAWork& w, double* user_cams, double* user_pts, hipStream_t s);

// Dense SPD solves of the reduced camera system (chol_kernels.hip).
//   factor S (n x n row-major, lower triangle) in place and solve S x = rhs
//   (rhs == nullptr: factor only, keeping what later triangular solves need);
//   skip when state->done or state->skip_step; flag state->step_invalid on a
//   non-positive or non-finite pivot.
// max_groups: persistent workgroups of the tiled factorisation (0: up to two per compute unit).
int launch_rcs_cholesky_solve(double* S, const double* rhs, double* x, int n, LMState* state, double* tmp,
                              hipStream_t s, int max_groups = 0);
// Triangular solves only, with the factor left in S by the Cholesky.
int launch_rcs_tri_solve(const double* L, const double* rhs, double* x, int n, LMState* state, double* tmp,
                         hipStream_t s);
// Explicit-S PCG with 9x9 block-Jacobi (SCHUR_JACOBI) preconditioning.
int launch_rcs_pcg(const double* S, const double* rhs, double* x, int n, LMState* state, double* work,
                   size_t work_doubles, hipStream_t s);
size_t rcs_pcg_work_doubles(int n);
unsigned int* rcs_pcg_barrier(double* work, int n);  // the PCG solve-epoch word inside its workspace
int rcs_pcg_prepare(double* work, int n, hipStream_t s);
// Shared by the bundle-adjustment entry points (ba_api.hip).
int ba_validate_problem(const ct_ba_problem* p);
int ba_check_tables(const ct_ba_problem& p);  // point and camera CSR tables
BAView ba_make_view(const ct_ba_problem* p);

// Test hook: every compute unit's LDS filled with `value` (ct_lm_options::debug_fill_lds).
int launch_fill_lds(double value, hipStream_t s);  // zero the epoch and barrier flags (once per plan)
// Implicit-S PCG (pcg_implicit.hip): S p from the JP records, block-Jacobi
// factors of the explicitly assembled diagonal blocks of S.
constexpr int kImplicitMaxCameras = 300;  // p, r and z (9C entries each) in LDS beside the y staging
constexpr int kImpMaxGroups = 1024;
struct ImplicitPcg {
  const int32_t* obs_camera;     // [N]
  const int32_t* obs_point;      // [N]
  const int32_t* point_offsets;  // [P+1]
  const int32_t* camera_offsets; // [C+1] slot ranges in camera-major order
  const int32_t* point_groups;   // [num_groups+1] first point of each wavefront's group
  const int32_t* camera_obs;     // [N] each camera's observations, camera by camera
  const double* JP;              // [N][kJPStride]
  const double* S;               // n x n storage; only the diagonal blocks are read
  const double* Dc;              // [n] camera damping
  const double* rhs;             // [n]
  double* x;                     // [n] solution
  double* y;                     // [N][9] Jc^T w per observation
  double* Ap;                    // [n]
  double* part;                  // [C] p_i . (S p)_i
  double* fac;                   // [C][81] inverse of each preconditioner block
  int32_t* kind;                 // [C] preconditioner kind (< 0: singular block)
  unsigned int* epoch;           // solves so far (advanced by the LM begin kernel)
  unsigned long long* flags;     // [kImpMaxGroups] grid-barrier slots
  int C, n, num_groups;
  LMState* st;
};
size_t implicit_pcg_doubles(int num_cameras, int num_observations);
ImplicitPcg implicit_pcg_layout(double* base, int num_cameras, int num_observations);
int implicit_pcg_prepare(const ImplicitPcg& a, hipStream_t s);  // zero the barrier words (once per plan)
int launch_pcg_implicit(const ImplicitPcg& a, hipStream_t s);
size_t rcs_cholesky_tmp_doubles(int n);
// Clears the synchronisation words inside the solve scratch; once per plan, before the first solve.
int rcs_cholesky_prepare(double* tmp, int n, hipStream_t s);

}  // namespace ct
