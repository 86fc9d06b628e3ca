/*
 * ceres_mi355x C ABI: the MI355X (gfx950) solver core behind the ceres_torch
 * Problem / SolverOptions / solve API.
 *
 * Conventions for every entry point:
 *   - returns int status: CT_OK (0) or a negative ct_status code; a message for
 *     the last failure on the calling thread is available from ct_last_error();
 *   - no C++ exception crosses the boundary;
 *   - all array arguments are caller-owned device pointers (unless marked
 *     "host"), with explicit sizes; the library never allocates device memory
 *     except where a function says so (plans keep host-side graph objects);
 *   - work is enqueued on the caller's hipStream_t (0 = the null stream) and
 *     functions are re-entrant per stream.
 *
 * The reference exposes its native tier as a pybind11 module with two
 * functions (native/cuda/ceres_torch_cuda.cpp:3-16, bodies in
 * native/cuda/ceres_torch_cuda_kernel.cu:48-132) plus a Python linear-solver
 * plugin registry (src/ceres_torch/linear.py:82-103).  This header replaces
 * that module; the per-function comments cite what each entry point replaces.
 */
#ifndef CERES_MI355X_H_
#define CERES_MI355X_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* ct_stream_t; /* == hipStream_t */

#define CT_ABI_VERSION 14

typedef enum ct_status {
  CT_OK = 0,
  CT_ERR_ARGUMENT = -1,    /* bad argument: message names it */
  CT_ERR_HIP = -2,         /* a HIP runtime call failed */
  CT_ERR_NOT_SPD = -3,     /* factorisation met a non-positive pivot */
  CT_ERR_WORKSPACE = -4,   /* workspace smaller than *_workspace_bytes() */
  CT_ERR_UNSUPPORTED = -5  /* valid request the device path does not cover */
} ct_status;

/* Version of this ABI (CT_ABI_VERSION).  No reference counterpart. */
int ct_abi_version(void);

/* Message for the last failed call on this thread ("" if none). */
const char* ct_last_error(void);

/* ------------------------------------------------------------------------ */
/* Robust losses: losses.py:21-122 (Trivial, Huber, SoftLOne, Cauchy, Arctan,
 * Tolerant, Tukey, each optionally scaled as ScaledLoss).                    */
/* ------------------------------------------------------------------------ */
typedef enum ct_loss_type {
  CT_LOSS_TRIVIAL = 0,
  CT_LOSS_HUBER = 1,
  CT_LOSS_SOFTLONE = 2,
  CT_LOSS_CAUCHY = 3,
  CT_LOSS_ARCTAN = 4,
  CT_LOSS_TOLERANT = 5,
  CT_LOSS_TUKEY = 6
} ct_loss_type;

typedef struct ct_loss {
  int32_t type;  /* ct_loss_type */
  double a;      /* scale parameter */
  double b;      /* second parameter (Tolerant) */
  double scale;  /* ScaledLoss factor; 1.0 for none */
} ct_loss;

/* ------------------------------------------------------------------------ */
/* Trust-region (Levenberg-Marquardt) options: the fields of SolverOptions
 * (src/ceres_torch/types.py:164-229) that solver.py:35-357 reads on the LM
 * path.                                                                     */
/* ------------------------------------------------------------------------ */
typedef enum ct_linear_solver {
  CT_DENSE_SCHUR = 0,      /* Schur complement + FP64 Cholesky of S */
  CT_ITERATIVE_SCHUR = 1   /* explicit S + SCHUR_JACOBI PCG */
} ct_linear_solver;

typedef enum ct_precision {
  CT_PLAIN_FP64 = 0,  /* FP64 everywhere */
  CT_MIXED = 1        /* FP32 J, J^T J and Schur assembly; FP64 Cholesky,
                         cost, gain ratio and iterative refinement */
} ct_precision;

typedef enum ct_loop_mode {
  CT_LOOP_HOST = 0,   /* host-driven LM loop, one host sync per iteration */
  CT_LOOP_DEVICE = 1  /* device-resident LM state, graph-captured iterations */
} ct_loop_mode;

/* Trust-region strategy (TrustRegionStrategyType, types.py:88-91) and dogleg
 * variant (DoglegType, types.py:93-95).  DOGLEG takes the step of
 * dogleg_step (linear.py:554-664) from the undamped Gauss-Newton step (the
 * Schur complement with no damping, factored by the FP64 Cholesky whatever
 * the linear solver) and the gradient, and updates the radius by
 * solver.py:864-883, 1051-1061.  J^T J must be nonsingular (for bundle
 * adjustment: the gauge fixed, e.g. by constant cameras); a factorisation that
 * meets a non-positive pivot gives an invalid step.                         */
typedef enum ct_trust_region_strategy {
  CT_LEVENBERG_MARQUARDT = 0,
  CT_DOGLEG = 1
} ct_trust_region_strategy;

typedef enum ct_dogleg_type {
  CT_TRADITIONAL_DOGLEG = 0,
  CT_SUBSPACE_DOGLEG = 1
} ct_dogleg_type;

/* ITERATIVE_SCHUR preconditioner (build_schur_complement_preconditioner,
 * linear.py:462-537): the Cholesky (LU if not SPD) of the 9x9 diagonal camera
 * blocks of S, the inverse diagonal 1 / max(S_kk, eps), or none.            */
typedef enum ct_preconditioner {
  CT_PRECONDITIONER_SCHUR_JACOBI = 0,
  CT_PRECONDITIONER_JACOBI = 1,
  CT_PRECONDITIONER_IDENTITY = 2
} ct_preconditioner;

typedef struct ct_lm_options {
  int32_t max_num_iterations;
  int32_t max_num_consecutive_invalid_steps;
  int32_t linear_solver;                /* ct_linear_solver */
  int32_t precision;                    /* ct_precision */
  int32_t loop_mode;                    /* ct_loop_mode */
  int32_t min_linear_solver_iterations;
  int32_t max_linear_solver_iterations;
  int32_t max_num_refinement_iterations;
  int32_t jacobi_scaling;               /* bool */
  int32_t iterations_per_graph;         /* device loop: iterations per graph */
  double initial_trust_region_radius;
  double max_trust_region_radius;
  double min_trust_region_radius;
  double min_relative_decrease;
  double min_lm_diagonal;
  double max_lm_diagonal;
  double function_tolerance;
  double gradient_tolerance;
  double parameter_tolerance;
  double eta;
  int32_t trust_region_strategy;        /* ct_trust_region_strategy (bundle adjustment only) */
  int32_t dogleg_type;                  /* ct_dogleg_type */
  int32_t preconditioner;               /* ct_preconditioner (ITERATIVE_SCHUR) */
  double max_solver_time_in_seconds;    /* device clock from the start of the solve (solver.py:342-345);
                                           SolverOptions' default is 1e9, and 0 stops after the first step */
  int32_t debug_invalid_steps;          /* test hook (bundle adjustment; 0 in normal use): the first k
                                           trust-region steps are overwritten with NaN on the device before
                                           they are checked, so the invalid-step branch of solver.py:168-190
                                           runs as after a linear solver returning NaN.  Curve fits and
                                           pose graphs refuse a nonzero value. */
  int32_t schur_operator;               /* ct_schur_operator (ITERATIVE_SCHUR, plain FP64): CT_SCHUR_AUTO applies
                                           S from the evaluation's records when the plan allows it (up to 600
                                           cameras, no point seen by more than 64); CT_SCHUR_EXPLICIT always
                                           assembles S.  Same step either way, to rounding. */
  int32_t debug_fill_lds;               /* test hook (bundle adjustment; 0 in normal use): 1 fills the LDS of
                                           every compute unit with NaN right before each reduced-camera solve,
                                           so a solver kernel that reads LDS it has not written in its own
                                           launch fails its parity test.  Curve fits and pose graphs ignore it. */
  int32_t solver_workgroups;            /* persistent workgroups of a reduced-system factorisation larger than
                                           192 unknowns (bundle adjustment, pose graphs); 0 = up to two per
                                           compute unit.  A server running several solves at once gives each a
                                           share of the chip (the factorisation's column chain keeps few
                                           workgroups busy at a time).  No reference counterpart. */
} ct_lm_options;

/* How ITERATIVE_SCHUR forms S p (no reference counterpart: the reference
 * always builds S densely, linear.py:794-839). */
typedef enum ct_schur_operator { CT_SCHUR_AUTO = 0, CT_SCHUR_EXPLICIT = 1 } ct_schur_operator;

/* Termination codes match TerminationType (types.py:98-103). */
typedef enum ct_termination {
  CT_TERM_CONVERGENCE = 0,
  CT_TERM_NO_CONVERGENCE = 1,
  CT_TERM_FAILURE = 2
} ct_termination;

/* Why the loop stopped; messages of solver.py:35-357. */
typedef enum ct_reason {
  CT_REASON_NONE = 0,
  CT_REASON_GRADIENT = 1,         /* "Gradient tolerance reached." */
  CT_REASON_PARAMETER = 2,        /* "Parameter tolerance reached." */
  CT_REASON_FUNCTION = 3,         /* "Function tolerance reached." */
  CT_REASON_MIN_RADIUS = 4,       /* "Minimum trust-region radius reached." */
  CT_REASON_MAX_ITERATIONS = 5,   /* "Maximum iterations reached." */
  CT_REASON_INVALID_STEPS = 6,    /* "Too many invalid trust-region steps." */
  CT_REASON_NO_PARAMETERS = 7,    /* "No active parameters." */
  CT_REASON_INVALID_INPUT = 8,    /* batched problems: malformed sample range */
  CT_REASON_MAX_TIME = 9,         /* "Maximum solver time reached." */
  CT_REASON_PROJECTED_GRADIENT = 10,  /* line search: "Projected gradient tolerance reached." */
  CT_REASON_LINE_SEARCH_FAILED = 11   /* line search: "Line search failed to find a decreasing step." */
} ct_reason;

/* Per-iteration record layout (doubles), mirroring IterationSummary
 * (types.py:134-161).  ct_lm_report.num_records rows are valid.            */
enum {
  CT_REC_ITERATION = 0,
  CT_REC_COST = 1,
  CT_REC_COST_CHANGE = 2,
  CT_REC_GRADIENT_MAX_NORM = 3,
  CT_REC_GRADIENT_NORM = 4,
  CT_REC_STEP_NORM = 5,
  CT_REC_RELATIVE_DECREASE = 6,
  CT_REC_TRUST_REGION_RADIUS = 7,
  CT_REC_STEP_IS_VALID = 8,
  CT_REC_STEP_IS_SUCCESSFUL = 9,
  CT_REC_LINEAR_SOLVER_ITERATIONS = 10,
  CT_REC_TIMESTAMP = 11,          /* device wall clock, seconds since start */
  CT_REC_FIELDS = 12
};

/* Host-side result of a solve (SolverSummary counters, types.py:360-387). */
typedef struct ct_lm_report {
  int32_t termination;       /* ct_termination */
  int32_t reason;            /* ct_reason */
  int32_t num_records;       /* rows of the iteration record table */
  int32_t num_successful_steps;
  int32_t num_unsuccessful_steps;
  int32_t num_residual_evaluations;
  int32_t num_jacobian_evaluations;
  int32_t num_linear_solves;
  double initial_cost;
  double final_cost;
  double device_seconds;     /* device wall clock from start to finish */
} ct_lm_report;

/* ------------------------------------------------------------------------ */
/* Bundle adjustment (Snavely/BAL cameras: angle-axis, translation, f, k1, k2;
 * residual of examples/simple_bundle_adjuster.py:73-79).  Observations are
 * grouped by point.  The camera-pair tables come from ct_ba_* helpers in the
 * Python layer (built once per problem structure).                          */
/* ------------------------------------------------------------------------ */
/* Camera parameterisations of a bundle-adjustment problem.  Every camera has a
 * 9-wide tangent space [rotation(3), translation(3), f, k1, k2] (the order of
 * camera_mask, camera columns and the reduced camera system); the stored
 * values are:
 *   CT_CAMERA_ANGLE_AXIS       [angle-axis(3), t(3), f, k1, k2]: 9 values, Euclidean
 *                              Plus (examples/simple_bundle_adjuster.py:73-79);
 *   CT_CAMERA_QUATERNION       [q(4) in w, x, y, z order, t(3), f, k1, k2]: 10 values; the
 *                              rotation is quaternion_rotate_point(q, X) (rotation.py:68-69,
 *                              211-218), q moves by QuaternionManifold.plus (manifolds.py:
 *                              127-158) and t, f, k1, k2 by Euclidean Plus: in the reference,
 *                              q and the other six values are two parameter blocks, so the
 *                              SCHUR_JACOBI preconditioner has a 3x3 and a 6x6 block per camera;
 *   CT_CAMERA_EIGEN_QUATERNION as CT_CAMERA_QUATERNION with q stored in x, y, z, w order
 *                              (EigenQuaternionManifold, manifolds.py:161-171).          */
typedef enum ct_camera_model {
  CT_CAMERA_ANGLE_AXIS = 0,
  CT_CAMERA_QUATERNION = 1,
  CT_CAMERA_EIGEN_QUATERNION = 2
} ct_camera_model;

typedef struct ct_ba_problem {
  int32_t num_cameras;
  int32_t num_points;
  int32_t num_observations;
  int32_t num_blocks;              /* camera pair blocks: every pair i <= j, C (C + 1) / 2 */
  int64_t num_terms;               /* observation pairs over all blocks */
  const int32_t* obs_camera;       /* [N] camera of each observation */
  const int32_t* obs_point;        /* [N] point of each observation (non-decreasing) */
  const int32_t* point_offsets;    /* [P+1] observations of point p */
  const double* observations;      /* [N*2] observed pixels */
  const int32_t* camera_offsets;   /* [C+1] CSR of observations by camera */
  const int32_t* camera_obs;       /* [N]   observation ids by camera */
  const int32_t* block_cameras;    /* [B*2] (i, j), i <= j, row-major: (0,0), (0,1), ..., (1,1), ... */
  const int32_t* block_offsets;    /* [B+1] term ranges */
  const int32_t* block_terms;      /* [T*2] (obs a of camera i, obs b of camera j) */
  const uint8_t* camera_mask;      /* [C*9] 1 = free, 0 = held constant (tangent coordinates) */
  const uint8_t* point_mask;       /* [P*3] */
  ct_loss loss;
  int32_t camera_model;            /* ct_camera_model: 9 (angle-axis) or 10 stored values per camera */
  int32_t reserved;
} ct_ba_problem;

/* Device workspace needed by ct_ba_plan_create for this problem/options. */
int ct_ba_workspace_bytes(const ct_ba_problem* problem, const ct_lm_options* options, size_t* bytes);

typedef struct ct_ba_plan ct_ba_plan; /* opaque, host memory only */

/* Bind a problem, options, parameter buffers and workspace to a plan.
 * cameras [C*9] (C*10 for the quaternion camera models) and points [P*3]
 * (device, float64) hold the initial estimate
 * and receive the final (best) state, as Problem parameter tensors do in the
 * reference (problem.py:457-460, solver.py:672).  Replaces the per-iteration
 * Python loop of solver.py:77-345 + Problem.evaluate (problem.py:377-413) +
 * solve_linear_system(DENSE_SCHUR / ITERATIVE_SCHUR) (linear.py:141-246).
 * Synchronises stream before reading the problem's tables back on the host
 * (the caller's uploads on it complete first).                             */
int ct_ba_plan_create(const ct_ba_problem* problem, const ct_lm_options* options, double* cameras,
                      double* points, void* workspace, size_t workspace_bytes, ct_stream_t stream,
                      ct_ba_plan** plan);

/* Run the LM solve; report and records (host memory, may be NULL) receive the
 * summary and (max_num_iterations + 2) * CT_REC_FIELDS doubles.            */
int ct_ba_plan_run(ct_ba_plan* plan, ct_lm_report* report, double* records);

int ct_ba_plan_destroy(ct_ba_plan* plan);

/* Kernel groups of one LM iteration, in launch order. */
enum {
  CT_KERNEL_EVAL_OBS = 0,        /* residual, Jacobian, loss correction per observation; with points of at
                                    most 64 observations also the point elimination (one fused kernel) */
  CT_KERNEL_EVAL_POINTS = 1,     /* point elimination: V, its Cholesky, q, P, e (empty when fused) */
  CT_KERNEL_CAMERA_ITEMS = 2,    /* reduced camera system: partial camera-pair blocks */
  CT_KERNEL_CAMERA_REDUCE = 3,   /* reduced camera system: block sums, damping, rhs */
  CT_KERNEL_LM_BEGIN = 4,        /* cost/gradient reduction + checks */
  CT_KERNEL_RCS_SOLVE = 5,       /* Cholesky (or PCG) of S */
  CT_KERNEL_REFINE = 6,          /* mixed: back-substitution + iterative refinement */
  CT_KERNEL_DOGLEG = 7,          /* DOGLEG: back-substitution, dot products, step coefficients, combined step */
  CT_KERNEL_STEP = 8,            /* step application (+ point back-substitution) */
  CT_KERNEL_CANDIDATE = 9,       /* candidate cost + model decrease per observation */
  CT_KERNEL_LM_END = 10,         /* gain ratio, radius, accept/reject */
  CT_KERNEL_COUNT = 11
};

/* Run up to `iterations` LM iterations (after an untimed iteration 0; fewer
 * if LM terminates) host-driven, with HIP events around each kernel group
 * on the plan's stream; kernel_ms[k] and kernel_launches[k] (host,
 * CT_KERNEL_COUNT entries) receive the summed milliseconds and the number
 * of timed launches (0 for a group that launches nothing in this plan).  The parameters end at the
 * state reached, as after ct_ba_plan_run.  No reference counterpart: the
 * reference times whole calls on the host (benchmarking.py:32-69).         */
int ct_ba_plan_profile(ct_ba_plan* plan, int iterations, double* kernel_ms, int32_t* kernel_launches);

/* Single evaluation at the current parameters (Problem.evaluate with
 * compute_jacobian=False, problem.py:377-413): cost and residuals [N*2]
 * (loss-corrected).  residuals may be NULL.                                 */
int ct_ba_evaluate(const ct_ba_problem* problem, const double* cameras, const double* points,
                   double* cost_out_host, double* residuals, void* workspace, size_t workspace_bytes,
                   ct_stream_t stream);

/* Evaluation with the Jacobian: Problem.evaluate(compute_jacobian=True)
 * (problem.py:377-413) with _evaluate_residual_block_internal
 * (problem.py:486-520) and _active_column_map (problem.py:522-530).  The
 * residual block of observation o fills rows 2*row_of_obs[o] and
 * 2*row_of_obs[o]+1 of residuals [N*2] and of jacobian [N*2][num_columns]
 * (row-major; the caller zero-fills it); camera_columns [C*9] and
 * point_columns [P*3] give each coordinate's tangent column or -1 (constant
 * block or coordinate).  apply_loss = 0 gives the uncorrected residuals and
 * the cost 1/2 |r|^2 (EvaluateOptions.apply_loss_function).  Indices are
 * checked on the host.                                                      */
int ct_ba_evaluate_jacobian(const ct_ba_problem* problem, const double* cameras, const double* points,
                            const int32_t* row_of_obs, const int64_t* camera_columns, const int64_t* point_columns,
                            int64_t num_columns, int32_t apply_loss, double* jacobian, double* residuals,
                            double* cost_out_host, void* workspace, size_t workspace_bytes, ct_stream_t stream);

/* ------------------------------------------------------------------------ */
/* The LINE_SEARCH minimizer on bundle adjustment: _line_search_solve
 * (solver.py:360-607) with its search directions (solver.py:1078-1162) and
 * step-size interpolation (line_search.py:9-81).  Every evaluation (cost and
 * gradient J^T r), the LBFGS two-loop recursion, the BFGS update and Plus run
 * on the device; the loop runs on the host and reads a few scalars back per
 * trial step.  No bounds, callbacks or non-monotonic steps (the Python layer
 * refuses them).                                                            */
/* ------------------------------------------------------------------------ */
typedef enum ct_line_search_direction {
  CT_LS_STEEPEST_DESCENT = 0,
  CT_LS_NONLINEAR_CONJUGATE_GRADIENT = 1,
  CT_LS_LBFGS = 2,
  CT_LS_BFGS = 3
} ct_line_search_direction;

typedef enum ct_line_search_type { CT_LS_ARMIJO = 0, CT_LS_WOLFE = 1 } ct_line_search_type;

typedef enum ct_nonlinear_cg_type {
  CT_NCG_FLETCHER_REEVES = 0,
  CT_NCG_POLAK_RIBIERE = 1,
  CT_NCG_HESTENES_STIEFEL = 2
} ct_nonlinear_cg_type;

typedef enum ct_line_search_interpolation {
  CT_LS_BISECTION = 0,
  CT_LS_QUADRATIC = 1,
  CT_LS_CUBIC = 2
} ct_line_search_interpolation;

/* BFGS keeps a dense n x n inverse Hessian (n = 9C + 3P tangent coordinates);
 * the device path takes it up to this n. */
#define CT_LS_MAX_BFGS_N 4096

/* SolverOptions' fields of the line search (types.py:164-198). */
typedef struct ct_line_search_options {
  int32_t max_num_iterations;
  int32_t direction;                     /* ct_line_search_direction */
  int32_t line_search_type;              /* ct_line_search_type */
  int32_t nonlinear_cg_type;             /* ct_nonlinear_cg_type */
  int32_t max_lbfgs_rank;
  int32_t use_approximate_eigenvalue_bfgs_scaling;
  int32_t interpolation;                 /* ct_line_search_interpolation */
  int32_t max_num_step_size_iterations;  /* max_num_line_search_step_size_iterations */
  int32_t max_num_direction_restarts;    /* max_num_line_search_direction_restarts */
  int32_t reserved;
  double min_line_search_step_size;
  double sufficient_function_decrease;   /* line_search_sufficient_function_decrease */
  double max_step_contraction;           /* max_line_search_step_contraction */
  double min_step_contraction;           /* min_line_search_step_contraction */
  double sufficient_curvature_decrease;  /* line_search_sufficient_curvature_decrease */
  double function_tolerance;
  double gradient_tolerance;
  double parameter_tolerance;
  double max_solver_time_in_seconds;
} ct_line_search_options;

typedef struct ct_line_search_report {
  int32_t termination;                   /* ct_termination */
  int32_t reason;                        /* ct_reason */
  int32_t num_records;                   /* iteration records written */
  int32_t num_successful_steps;
  int32_t num_unsuccessful_steps;
  int32_t num_jacobian_evaluations;
  int32_t num_line_search_steps;
  int32_t num_line_search_function_evaluations;
  int32_t num_line_search_gradient_evaluations;
  int32_t num_line_search_direction_restarts;
  double initial_cost;
  double final_cost;
  double seconds;                        /* host wall time of the solve */
} ct_line_search_report;

/* One record per iteration (IterationSummary of the line search, solver.py:387-395, 499-509). */
enum {
  CT_LS_REC_ITERATION = 0,
  CT_LS_REC_COST = 1,
  CT_LS_REC_GRADIENT_NORM = 2,
  CT_LS_REC_GRADIENT_MAX_NORM = 3,
  CT_LS_REC_STEP_SIZE = 4,
  CT_LS_REC_LINE_SEARCH_ITERATIONS = 5,
  CT_LS_REC_FUNCTION_EVALUATIONS = 6,
  CT_LS_REC_GRADIENT_EVALUATIONS = 7,
  CT_LS_REC_DIRECTION_RESTARTS = 8,
  CT_LS_REC_STEP_NORM = 9,
  CT_LS_REC_COST_CHANGE = 10,
  CT_LS_REC_STEP_IS_SUCCESSFUL = 11,
  CT_LS_REC_TIMESTAMP = 12,              /* seconds since the start of the solve */
  CT_LS_REC_FIELDS = 13
};

/* Device workspace of ct_ba_line_search_solve for this problem and options. */
int ct_ba_line_search_workspace_bytes(const ct_ba_problem* problem, const ct_line_search_options* options,
                                      size_t* bytes);

/* Solve from cameras / points (device, float64, the problem's camera model)
 * and leave the last accepted state in them (solver.py:552, 560).  report and
 * records (host; records holds (max_num_iterations + 1) * CT_LS_REC_FIELDS
 * doubles, may be NULL) receive the summary.  Synchronous on stream.       */
int ct_ba_line_search_solve(const ct_ba_problem* problem, const ct_line_search_options* options, double* cameras,
                            double* points, void* workspace, size_t workspace_bytes, ct_stream_t stream,
                            ct_line_search_report* report, double* records);

/* ------------------------------------------------------------------------ */
/* Batched curve fits: y = exp(m x + c), one residual block per sample with a
 * robust loss (examples/robust_curve_fitting.py:80-97; parameter blocks m and
 * c), many independent problems of uneven size solved by the LM loop of
 * solver.py:35-357 in one persistent launch.  For two parameters the damped
 * normal equations give the reference's DENSE_QR step (linear.py:359-371).  */
/* ------------------------------------------------------------------------ */
typedef struct ct_curve_batch {
  int32_t num_problems;
  int32_t reserved;
  int64_t num_samples;
  const int64_t* offsets;   /* [num_problems+1] device: problem p owns samples offsets[p]..offsets[p+1]-1 */
  const double* x;          /* [num_samples] device */
  const double* y;          /* [num_samples] device */
  const int32_t* order;     /* [num_problems] device processing order (a permutation), or NULL */
  ct_loss loss;
  const float* x32;         /* [num_samples] device FP32 samples, read by the mixed form only (which then  */
  const float* y32;         /* solves the FP32-rounded problem; x / y may be NULL for it)                  */
} ct_curve_batch;

/* Launch shape of a persistent batched solve (chosen by the autotuner of
 * ceres_mi355x.autotune).  No reference counterpart.                        */
typedef struct ct_batch_schedule {
  int32_t threads;          /* workgroup size: 64, 128, 256 or 512 */
  int32_t groups;           /* persistent workgroups; 0 = 8 per compute unit */
  int32_t wave_problems;    /* the last wave_problems problems of the processing order are solved by one
                               wave (64 lanes) each, the others by a whole workgroup; 0 = none */
} ct_batch_schedule;

/* Device workspace of ct_curve_solve (two work-queue counters). */
int ct_curve_workspace_bytes(const ct_curve_batch* batch, size_t* bytes);

/* Solve every problem of the batch: params [num_problems][2] (m, c) FP64
 * device, start values in, best values out; results [num_problems] device
 * (may be NULL); records [num_problems][max_iter+2][CT_REC_FIELDS] device (may
 * be NULL).  options->precision CT_MIXED evaluates residuals, Jacobians, J^T J
 * and J^T r in FP32 from x32 / y32 and every cost (the candidate cost that
 * accepts or rejects a step, the cost tests) in FP64.  Enqueues
 * two kernels on stream and returns: no host synchronisation, so a caller may
 * capture it in a graph.  Malformed sample ranges end that problem with
 * CT_REASON_INVALID_INPUT; ct_curve_validate checks them on the host.        */
int ct_curve_solve(const ct_curve_batch* batch, const ct_lm_options* options, const ct_batch_schedule* schedule,
                   double* params, ct_lm_report* results, double* records, void* workspace, size_t workspace_bytes,
                   ct_stream_t stream);

/* Host check of offsets (and order) of a batch; synchronous.  It reads the
 * arrays with a plain device-to-host copy, which does not wait for work on a
 * non-blocking stream: the caller's uploads of them must be complete.       */
int ct_curve_validate(const ct_curve_batch* batch, int32_t* first_bad_problem);

/* ------------------------------------------------------------------------ */
/* Batched small bundle adjustments: many independent problems (angle-axis
 * Snavely cameras, every parameter free, at most CT_BA_BATCH_MAX_CAMERAS
 * cameras each) solved in one persistent launch, one workgroup per problem
 * running the whole LM loop of solver.py:35-357 with the DENSE_SCHUR step of
 * schur_solve_dense (linear.py:667-699) in FP64.  The reference solves one
 * Problem per solve() call (solver.py:28); batching has no counterpart there.
 * Every index table holds problem-local values (cameras 0..C-1, points
 * 0..P-1, observations 0..N-1), laid out as ct_ba_problem's tables of that
 * problem alone; the per-problem record says where its slice of each
 * concatenated array starts.                                               */
/* ------------------------------------------------------------------------ */
#define CT_BA_BATCH_MAX_CAMERAS 21   /* 9C <= 189: the reduced camera system is factored by one workgroup */

typedef struct ct_ba_batch_problem {
  int32_t num_cameras;             /* C, 1..CT_BA_BATCH_MAX_CAMERAS */
  int32_t num_points;              /* P */
  int32_t num_observations;        /* N */
  int32_t reserved;
  int64_t num_terms;               /* T: its observation pairs in block_terms */
  int64_t camera_base;             /* first camera in cameras [.][9] */
  int64_t point_base;              /* first point in points [.][3] */
  int64_t obs_base;                /* first observation in obs_camera, obs_point, observations, camera_obs */
  int64_t point_offsets_base;      /* first of the problem's P+1 entries of point_offsets */
  int64_t camera_offsets_base;     /* first of its C+1 entries of camera_offsets */
  int64_t block_offsets_base;      /* first of its C(C+1)/2+1 entries of block_offsets */
  int64_t term_base;               /* first of its observation pairs in block_terms */
  int64_t work_base;               /* byte offset of its workspace slice (256-byte aligned, >= 256) */
  ct_loss loss;
} ct_ba_batch_problem;

typedef struct ct_ba_batch {
  int32_t num_problems;
  int32_t reserved;
  const ct_ba_batch_problem* problems;  /* [num_problems] device */
  const int32_t* obs_camera;       /* camera of each observation, grouped by point */
  const int32_t* obs_point;        /* point of each observation, non-decreasing within a problem */
  const double* observations;      /* [*][2] observed pixels */
  const int32_t* point_offsets;    /* observation range of each point */
  const int32_t* camera_offsets;   /* CSR of observations by camera */
  const int32_t* camera_obs;
  const int32_t* block_offsets;    /* term ranges of the camera pairs (i, j), i <= j, row-major */
  const int32_t* block_terms;      /* [*][2] (observation a of camera i, observation b of camera j) */
  const int32_t* order;            /* [num_problems] processing order (a permutation), or NULL */
} ct_ba_batch;

/* Host tables of a batch (host memory, synchronous).  Inputs: per problem C, P,
 * N and the problem-local obs_camera / obs_point, concatenated in problem
 * order.  Called with order == NULL it only fills num_terms [num_problems];
 * called again with every output (sizes: order, sorted_obs_camera,
 * sorted_obs_point, camera_obs [sum N]; point_offsets [sum (P+1)];
 * camera_offsets [sum (C+1)]; block_offsets [sum (C(C+1)/2+1)]; block_terms
 * [2 sum T]) it writes each problem's tables as ct_ba_problem's (observations
 * grouped by point, cameras ascending within a point, ties in input order;
 * order[k] = the input observation at sorted position k; each block's pairs
 * in ascending (a, b)), in problem order.  No reference counterpart.       */
int ct_ba_batch_tables(int32_t num_problems, const int32_t* num_cameras, const int32_t* num_points,
                       const int32_t* num_observations, const int32_t* obs_camera, const int32_t* obs_point,
                       int64_t* num_terms, int32_t* order, int32_t* sorted_obs_camera, int32_t* sorted_obs_point,
                       int32_t* point_offsets, int32_t* camera_offsets, int32_t* camera_obs, int32_t* block_offsets,
                       int32_t* block_terms);

/* Bytes of one problem's workspace slice. */
int ct_ba_batch_problem_bytes(int32_t num_cameras, int32_t num_points, int32_t num_observations, int64_t num_terms,
                              size_t* bytes);

/* Solve every problem of the batch.  cameras [*][9] and points [*][3] (device
 * FP64) hold the start values and receive each problem's best state;
 * results [num_problems] and records [num_problems][max_iter+2][CT_REC_FIELDS]
 * (device, may be NULL).  options: LEVENBERG_MARQUARDT with DENSE_SCHUR in
 * plain FP64 (other strategies, solvers and precisions are refused).  The
 * workspace's first 256 bytes hold the work-queue counter; the problems'
 * slices follow at their work_base.  Enqueues two kernels and returns (no
 * host synchronisation: a caller may capture it in a graph).  A problem whose
 * sizes are out of range ends with CT_REASON_INVALID_INPUT.                 */
int ct_ba_batch_solve(const ct_ba_batch* batch, const ct_lm_options* options, double* cameras, double* points,
                      ct_lm_report* results, double* records, void* workspace, size_t workspace_bytes,
                      ct_stream_t stream);

/* ------------------------------------------------------------------------ */
/* SE2 pose graphs (examples/pose_graph_2d.py:6-46): poses (x, y, theta),
 * residual relative_pose(a, b) - measurement per edge with a robust loss, one
 * node held constant, the LM loop of solver.py:35-357 with the exact step of
 * the damped normal equations (the reference's DENSE_QR), FP64.  Tables
 * (built by ceres_mi355x.pose2d): per-node incident edges and per-node-pair
 * edges, each entry tagged 2*edge + (the node is the edge's second end).    */
/* ------------------------------------------------------------------------ */
typedef struct ct_pose2d_problem {
  int32_t num_nodes;
  int32_t num_edges;
  int32_t num_pairs;             /* distinct unordered node pairs joined by edges */
  int32_t fixed_node;            /* the constant node, or -1 */
  const int32_t* edges;          /* [E][2] (a, b) device */
  const double* measurements;    /* [E][3] pose of b in a's frame */
  const int32_t* node_offsets;   /* [N+1] CSR into node_edges */
  const int32_t* node_edges;     /* [2E] tagged incident edges */
  const int32_t* pair_nodes;     /* [P][2] u < v */
  const int32_t* pair_offsets;   /* [P+1] CSR into pair_edges */
  const int32_t* pair_edges;     /* [E] tagged (tag bit: u is the edge's second end) */
  ct_loss loss;
} ct_pose2d_problem;

typedef struct ct_pose2d_plan ct_pose2d_plan;

/* Device workspace of a plan; reads only the sizes (num_nodes, num_edges,
 * num_pairs, fixed_node), so a caller may ask before uploading the tables. */
int ct_pose2d_workspace_bytes(const ct_pose2d_problem* problem, const ct_lm_options* options, size_t* bytes);
/* poses [N][3] FP64 device: start values in; ct_pose2d_plan_run leaves the
 * best values there.  Validates every table index on the host, after
 * synchronising stream (the caller's uploads on it complete first).        */
int ct_pose2d_plan_create(const ct_pose2d_problem* problem, const ct_lm_options* options, double* poses,
                          void* workspace, size_t workspace_bytes, ct_stream_t stream, ct_pose2d_plan** plan);
int ct_pose2d_plan_run(ct_pose2d_plan* plan, ct_lm_report* report, double* records);
/* Unknowns of the system the plan factors densely each iteration: 3 x (free
 * nodes) in the dense form, 3 x (separator nodes) in the separator form (the
 * nodes touched by a non-chain edge; the runs of consecutive nodes between
 * them joined only by chain edges (i, i + 1) are eliminated as block
 * tridiagonal systems).  The plan takes the separator form whenever some node
 * lies in such a run.  No counterpart in the reference, whose DENSE_QR step
 * (examples/pose_graph_2d.py:42-45) both forms reproduce exactly.            */
int ct_pose2d_plan_system_size(const ct_pose2d_plan* plan, int32_t* unknowns);
int ct_pose2d_plan_destroy(ct_pose2d_plan* plan);

/* Covariance blocks of a bundle-adjustment problem at the user's parameters
 * (replaces Covariance.compute / GetCovarianceBlockInTangentSpace,
 * covariance.py:86-172, 279-296, for a full-rank Jacobian: the reference's
 * SVD and QR forms both give (J^T J)^-1 then).  The plan must have been
 * created with trust_region_strategy = CT_DOGLEG and CT_PLAIN_FP64, so that
 * its system is the undamped J^T J (constant coordinates keep a unit
 * diagonal and decouple).  S = U - W^T V^-1 W is factored in FP64, S^-1 =
 * L^-T L^-1 formed densely (9C <= 8192), and each request
 * {kind_a, index_a, kind_b, index_b} (kind 0 camera, 1 point; host array)
 * writes the (9|3) x (9|3) block of (J^T J)^-1 over the two blocks' full
 * coordinates, row-major, into out[r * 81 ...] (device).
 * Rank test: a pivot of S's factor or of a point's 3x3 factor not above
 * pivot_rtol times the largest pivot of its factor counts as a null
 * direction (the reference applies 20 (m + n) eps to singular values,
 * covariance.py:374-377); their number goes to *deficient_pivots (may be
 * NULL) and the call returns CT_ERR_NOT_SPD, out left unwritten.
 * Synchronises the plan's stream.                                           */
int ct_ba_covariance_workspace_bytes(int32_t num_cameras, int64_t num_requests, size_t* bytes);
int ct_ba_plan_covariance(ct_ba_plan* plan, const int32_t* requests, int64_t num_requests, double pivot_rtol,
                          double* out, int64_t* deficient_pivots, void* workspace, size_t workspace_bytes);

/* ------------------------------------------------------------------------ */
/* Dense damped normal equations for a row-major A [m x n] (leading dimension
 * lda): x = (A^T A + diag(damping))^-1 A^T b, damping may be NULL.  Replaces
 * the reference's native normal_equations_solve / block_schur_solve
 * (ceres_torch_cuda.cpp:3-16, cuda_backends.py:267-307) behind the plugin
 * names sparse_normal_cholesky, sparse_cholesky, sparse_schur, block_schur.
 * FP64.  A^T A and A^T b are formed from row chunks of A and the chunks are
 * summed in a fixed order, so repeated calls give identical results; the
 * workspace (ct_dense_workspace_bytes) depends on n only.  Synchronises the
 * stream to report CT_ERR_NOT_SPD.                                          */
/* ------------------------------------------------------------------------ */
int ct_dense_workspace_bytes(int64_t n, size_t* bytes);
int ct_normal_equations_solve_f64(const double* A, int64_t m, int64_t n, int64_t lda, const double* b,
                                  const double* damping, double* x, void* workspace, size_t workspace_bytes,
                                  ct_stream_t stream);
/* ct_normal_equations_solve_f64 without a workspace argument: no counterpart
 * in the reference.  It is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the workspace is a static device array inside the library, one
 * copy per GPU, held until the process ends (36 MiB), so the call itself
 * allocates nothing.  n <= CT_NOALLOC_MAX_N; larger n returns CT_ERR_ARGUMENT.
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (other streams or threads) race and must be serialised by the caller.
 * Synchronises the stream to report CT_ERR_NOT_SPD.                         */
#define CT_NOALLOC_MAX_N 512
int ct_normal_equations_solve_f64_noalloc(const double* A, int64_t m, int64_t n, int64_t lda, const double* b,
                                          const double* damping, double* x, ct_stream_t stream);
/* Schur complement solve with the first num_eliminate unknowns eliminated
 * (cuda_backends.py:232-264); the solution is that of the whole system.     */
int ct_block_schur_solve_f64(const double* A, int64_t m, int64_t n, int64_t lda, const double* b,
                             int64_t num_eliminate, const double* damping, double* x, void* workspace,
                             size_t workspace_bytes, ct_stream_t stream);

/* Standalone FP64 Cholesky of an SPD S [n x n] row-major (lower triangle
 * used): potrf factors in place; potrs solves L L^T x = rhs with that factor
 * and the same workspace (torch.linalg.cholesky / cholesky_solve at
 * linear.py:1151, 1170).  Both synchronise the stream to report
 * CT_ERR_NOT_SPD.                                                            */
int ct_rcs_workspace_bytes(int64_t n, size_t* bytes);
int ct_rcs_potrf_f64(double* S, int64_t n, void* workspace, size_t workspace_bytes, ct_stream_t stream);
int ct_rcs_potrs_f64(const double* L, int64_t n, const double* rhs, double* x, void* workspace,
                     size_t workspace_bytes, ct_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* CERES_MI355X_H_ */
