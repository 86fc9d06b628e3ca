// C ABI of the bundle-adjustment solver: workspace sizing, plans, and the
// host-driven or graph-captured device-resident LM loop.
#include <cstring>
#include <memory>
#include <vector>

#include "ba_kernels.h"
#include "lm_plan.h"

namespace ct {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

namespace {

int validate_problem(const ct_ba_problem* p) {
  CT_CHECK_ARG(p != nullptr, "problem is null");
  CT_CHECK_ARG(p->num_cameras >= 0 && p->num_points >= 0 && p->num_observations >= 0, "negative problem size");
  CT_CHECK_ARG(p->num_blocks >= 0 && p->num_terms >= 0, "negative block/term count");
  if (p->num_observations > 0) {
    CT_CHECK_ARG(p->obs_camera && p->obs_point && p->point_offsets && p->observations, "observation arrays are null");
    CT_CHECK_ARG(p->camera_offsets && p->camera_obs, "camera CSR arrays are null");
  }
  if (p->num_blocks > 0) {
    CT_CHECK_ARG(p->block_cameras && p->block_offsets, "block arrays are null");
    CT_CHECK_ARG(p->num_terms == 0 || p->block_terms, "block_terms is null");
  }
  CT_CHECK_ARG(p->loss.type >= CT_LOSS_TRIVIAL && p->loss.type <= CT_LOSS_TUKEY, "unknown loss type");
  CT_CHECK_ARG((int64_t)p->num_cameras * 9 <= (int64_t)1 << 20, "too many cameras for the dense reduced system");
  CT_CHECK_ARG(p->camera_model >= CT_CAMERA_ANGLE_AXIS && p->camera_model <= CT_CAMERA_EIGEN_QUATERNION,
               "unknown camera model");
  return CT_OK;
}

int camera_stride(const ct_ba_problem* p) { return p->camera_model == CT_CAMERA_ANGLE_AXIS ? 9 : 10; }

int validate_options(const ct_lm_options* o) {
  CT_CHECK_ARG(o != nullptr, "options is null");
  CT_CHECK_ARG(o->max_num_iterations >= 0, "max_num_iterations must be >= 0");
  CT_CHECK_ARG(o->linear_solver == CT_DENSE_SCHUR || o->linear_solver == CT_ITERATIVE_SCHUR, "unknown linear solver");
  CT_CHECK_ARG(o->precision == CT_PLAIN_FP64 || o->precision == CT_MIXED, "unknown precision");
  CT_CHECK_ARG(o->loop_mode == CT_LOOP_HOST || o->loop_mode == CT_LOOP_DEVICE, "unknown loop mode");
  CT_CHECK_ARG(o->initial_trust_region_radius > 0.0, "initial_trust_region_radius must be > 0");
  CT_CHECK_ARG(o->max_solver_time_in_seconds >= 0.0, "max_solver_time_in_seconds must be >= 0");
  CT_CHECK_ARG(o->trust_region_strategy == CT_LEVENBERG_MARQUARDT || o->trust_region_strategy == CT_DOGLEG,
               "unknown trust-region strategy");
  CT_CHECK_ARG(o->dogleg_type == CT_TRADITIONAL_DOGLEG || o->dogleg_type == CT_SUBSPACE_DOGLEG, "unknown dogleg type");
  CT_CHECK_ARG(o->preconditioner >= CT_PRECONDITIONER_SCHUR_JACOBI && o->preconditioner <= CT_PRECONDITIONER_IDENTITY,
               "unknown preconditioner");
  CT_CHECK_ARG(o->debug_invalid_steps >= 0, "debug_invalid_steps must be >= 0");
  CT_CHECK_ARG(o->schur_operator == CT_SCHUR_AUTO || o->schur_operator == CT_SCHUR_EXPLICIT, "unknown schur_operator");
  CT_CHECK_ARG(o->debug_fill_lds == 0 || o->debug_fill_lds == 1, "debug_fill_lds must be 0 or 1");
  return CT_OK;
}

// DOGLEG solves the undamped Gauss-Newton system exactly (FP64 Cholesky),
// whatever the linear solver (solver.py:129-133 never calls the linear solvers).
bool dogleg_options(const ct_lm_options* o) { return o->trust_region_strategy == CT_DOGLEG; }
bool uses_cholesky(const ct_lm_options* o) { return o->linear_solver == CT_DENSE_SCHUR || dogleg_options(o); }

int obs_blocks(const ct_ba_problem* p) { return (p->num_observations + kPointThreads - 1) / kPointThreads; }
int eval_blocks(const ct_ba_problem* p) { return (p->num_points + kPointThreads - 1) / kPointThreads; }
int step_blocks(const ct_ba_problem* p) {
  const int work = p->num_points > 9 * p->num_cameras ? p->num_points : 9 * p->num_cameras;
  return (work + kPointThreads - 1) / kPointThreads;
}

// Upper bound of the camera-block work items: a block (or a camera's
// observation list) of t entries is split into ceil(t / kItemMaxWork) items.
size_t max_items(const ct_ba_problem* p) {
  const size_t W = kItemMaxWork;
  return (size_t)p->num_blocks + (size_t)p->num_cameras + ((size_t)p->num_terms + W - 1) / W +
         ((size_t)p->num_observations + W - 1) / W + 1;
}

// The implicit-S PCG needs plain FP64 records, the reduced system's vectors
// in LDS and the fused point groups (checked at plan creation).
bool implicit_possible(const ct_ba_problem* p, const ct_lm_options* o) {
  return o->linear_solver == CT_ITERATIVE_SCHUR && !dogleg_options(o) && o->precision == CT_PLAIN_FP64 &&
         o->schur_operator == CT_SCHUR_AUTO && p->num_cameras > 0 && p->num_cameras <= kImplicitMaxCameras;
}

struct Layout {
  BAWork w{};
  BlockItem* items = nullptr;
  int32_t* block_item_offsets = nullptr;
  int32_t* item_perm = nullptr;
  int32_t* point_groups = nullptr;
  BABuffers<float> bf{};
  BABuffers<double> bd{};
  double* pcg_work = nullptr;
  size_t pcg_doubles = 0;
  double* implicit_work = nullptr;  // ITERATIVE_SCHUR: the implicit-S PCG's buffers
  size_t bytes = 0;
};

Layout carve(const ct_ba_problem* p, const ct_lm_options* o, void* base) {
  Layout L;
  Carver cv(base);
  const size_t C = p->num_cameras, P = p->num_points, N = p->num_observations;
  const size_t n = 9 * C;
  L.w.xcam = cv.take<double>(2 * (size_t)camera_stride(p) * C);
  L.w.xpt = cv.take<double>(2 * 3 * P);
  L.w.S = cv.take<double>(n * n);
  L.w.rhs = cv.take<double>(n);
  L.w.dxc = cv.take<double>(n);
  L.w.dxp = cv.take<double>(3 * P);
  // cost and gradient partials: per observation / point block, or per fused-evaluation workgroup
  const size_t fused_blocks = ((size_t)P + kFusedThreads / kWave - 1) / (kFusedThreads / kWave);
  L.w.part_cost = cv.take<double>(std::max((size_t)obs_blocks(p), fused_blocks) + 1);
  L.w.part_eval = cv.take<double>(2 * std::max((size_t)eval_blocks(p), fused_blocks) + 2);
  L.w.part_step = cv.take<double>(3 * std::max((size_t)step_blocks(p), fused_blocks) + 3);
  L.w.part_cand = cv.take<double>(2 * std::max((size_t)obs_blocks(p), fused_blocks) + 2);
  L.w.part_dogleg = cv.take<double>(6 * (size_t)dogleg_blocks((int)P) + 6);
  L.w.records = cv.take<double>((size_t)(o->max_num_iterations + 2) * CT_REC_FIELDS);
  L.w.state = cv.take<LMState>(1);
  L.w.Dp = cv.take<double>(3 * P);
  L.w.Dc = cv.take<double>(n);
  L.w.gc = cv.take<double>(n);
  L.w.res_c = cv.take<double>(n);
  L.w.res_p = cv.take<double>(3 * P);
  L.w.u_obs = cv.take<double>(4 * N);
  L.w.dxc_corr = cv.take<double>(n);
  L.w.solve_tmp = cv.take<double>(rcs_cholesky_tmp_doubles((int)n) + 1);
  L.w.status = cv.take<int32_t>(4);
  const size_t I = max_items(p);
  L.items = cv.take<BlockItem>(I);
  L.block_item_offsets = cv.take<int32_t>((size_t)p->num_blocks + 1);
  L.item_perm = cv.take<int32_t>(I);
  L.w.items = L.items;
  L.w.block_item_offsets = L.block_item_offsets;
  L.w.item_perm = L.item_perm;
  L.w.item_vec = cv.take<double>(27 * I);
  L.point_groups = cv.take<int32_t>(P + 1);
  L.w.point_groups = L.point_groups;
  if (o->precision == CT_MIXED) {
    L.bf.JP = cv.take<float>(kJPStride * N);
    L.bf.Jp = cv.take<float>(6 * N);
    L.bf.e = cv.take<float>(2 * N);
    L.bf.r = cv.take<double>(2 * N);
    L.bf.Linv = cv.take<float>(6 * P);
    L.bf.q = cv.take<float>(3 * P);
    L.bf.gp = cv.take<double>(3 * P);
    L.bf.item_S = cv.take<float>(81 * I);
  } else {
    L.bd.JP = cv.take<double>(kJPStride * N);
    L.bd.Jp = cv.take<double>(6 * N);
    L.bd.e = cv.take<double>(2 * N);
    L.bd.r = cv.take<double>(2 * N);
    L.bd.Linv = cv.take<double>(6 * P);
    L.bd.q = cv.take<double>(3 * P);
    L.bd.gp = cv.take<double>(3 * P);
    L.bd.item_S = cv.take<double>(81 * I);
  }
  if (o->linear_solver == CT_ITERATIVE_SCHUR && !dogleg_options(o)) {
    L.pcg_doubles = rcs_pcg_work_doubles((int)n);
    L.pcg_work = cv.take<double>(L.pcg_doubles);
    L.w.pcg_bar = base ? rcs_pcg_barrier(L.pcg_work, (int)n) : nullptr;
    if (implicit_possible(p, o)) {
      L.implicit_work = cv.take<double>(implicit_pcg_doubles((int)C, (int)N));
    }
  }
  L.bytes = cv.offset + 256;
  return L;
}

BAView make_view(const ct_ba_problem* p) {
  BAView v;
  v.C = p->num_cameras;
  v.P = p->num_points;
  v.N = p->num_observations;
  v.B = p->num_blocks;
  v.T = p->num_terms;
  v.obs_camera = p->obs_camera;
  v.obs_point = p->obs_point;
  v.point_offsets = p->point_offsets;
  v.obs_xy = p->observations;
  v.camera_offsets = p->camera_offsets;
  v.camera_obs = p->camera_obs;
  v.block_cameras = p->block_cameras;
  v.block_offsets = p->block_offsets;
  v.block_terms = p->block_terms;
  v.camera_mask = p->camera_mask;
  v.point_mask = p->point_mask;
  v.loss = p->loss;
  if (v.loss.scale == 0.0) v.loss.scale = 1.0;
  v.cam_model = p->camera_model;
  v.cam_stride = camera_stride(p);
  return v;
}

}  // namespace

// Helpers for the library's other bundle-adjustment entry points.
int ba_validate_problem(const ct_ba_problem* p) { return validate_problem(p); }

BAView ba_make_view(const ct_ba_problem* p) { return make_view(p); }

}  // namespace ct

using namespace ct;

struct ct_ba_plan {
  ct_ba_problem problem;
  ct_lm_options options;
  BAView view;
  Layout layout;
  LMInit init;
  double* user_cams = nullptr;
  double* user_pts = nullptr;
  hipStream_t stream = nullptr;
  LMDriver driver;
  int num_items = 0;
  int num_groups = 0;       // point groups of the fused evaluation (0: two-kernel evaluation)
  bool unobserved_points = false;  // some point has no observation (the fused evaluation then covers it)
  bool implicit = false;           // ITERATIVE_SCHUR applies S from the records (no off-diagonal blocks)
  ImplicitPcg imp{};
};

namespace {

bool refining(const ct_ba_plan* pl) {
  return pl->options.precision == CT_MIXED && pl->options.linear_solver == CT_DENSE_SCHUR && !dogleg_options(&pl->options) &&
         pl->options.max_num_refinement_iterations > 0;
}

// Kernel group k of one LM iteration (CT_KERNEL_* order).
int enqueue_group(ct_ba_plan* pl, int group, hipStream_t s) {
  const BAView& v = pl->view;
  Layout& L = pl->layout;
  const int n = 9 * v.C;
  const bool mixed = pl->options.precision == CT_MIXED;
  switch (group) {
    case CT_KERNEL_EVAL_OBS:
      if (pl->num_groups > 0)
        return mixed ? launch_eval_fused<float>(v, L.bf, L.w, pl->num_groups, pl->unobserved_points, s)
                     : launch_eval_fused<double>(v, L.bd, L.w, pl->num_groups, pl->unobserved_points, s);
      return mixed ? launch_eval_obs<float>(v, L.bf, L.w, s) : launch_eval_obs<double>(v, L.bd, L.w, s);
    case CT_KERNEL_EVAL_POINTS:
      if (pl->num_groups > 0) return CT_OK;  // the fused evaluation covered the points
      return mixed ? launch_eval_points<float>(v, L.bf, L.w, s) : launch_eval_points<double>(v, L.bd, L.w, s);
    case CT_KERNEL_CAMERA_ITEMS:
      return mixed ? launch_camera_items<float>(v, L.bf, L.w, pl->num_items, s)
                   : launch_camera_items<double>(v, L.bd, L.w, pl->num_items, s);
    case CT_KERNEL_CAMERA_REDUCE:
      return mixed ? launch_camera_reduce<float>(v, L.bf, L.w, pl->implicit, s)
                   : launch_camera_reduce<double>(v, L.bd, L.w, pl->implicit, s);
    case CT_KERNEL_LM_BEGIN:
      return launch_lm_begin(pl->init, L.w, s);
    case CT_KERNEL_RCS_SOLVE:
      if (pl->options.debug_fill_lds) CT_TRY(launch_fill_lds(__builtin_nan(""), s));
      if (pl->implicit) return launch_pcg_implicit(pl->imp, s);
      if (!uses_cholesky(&pl->options))
        return launch_rcs_pcg(L.w.S, L.w.rhs, L.w.dxc, n, L.w.state, L.pcg_work, L.pcg_doubles, s);
      return launch_rcs_cholesky_solve(L.w.S, L.w.rhs, L.w.dxc, n, L.w.state, L.w.solve_tmp, s,
                                       pl->options.solver_workgroups);
    case CT_KERNEL_REFINE:
      if (!refining(pl)) return CT_OK;
      return launch_refinement<float>(v, L.bf, L.w, pl->options.max_num_refinement_iterations, s);
    case CT_KERNEL_DOGLEG:
      if (!dogleg_options(&pl->options)) return CT_OK;
      return mixed ? launch_dogleg_step<float>(v, L.bf, L.w, s) : launch_dogleg_step<double>(v, L.bd, L.w, s);
    case CT_KERNEL_STEP: {
      if (pl->options.debug_invalid_steps > 0) CT_TRY(launch_poison_step(v, L.w, s));
      // the refinement and the dogleg leave the whole step in dxc / dxp
      const bool backsub = !refining(pl) && !dogleg_options(&pl->options);
      if (pl->num_groups > 0)
        return mixed ? launch_step_candidate<float>(v, L.bf, L.w, pl->num_groups, backsub, s)
                     : launch_step_candidate<double>(v, L.bd, L.w, pl->num_groups, backsub, s);
      return mixed ? launch_step<float>(v, L.bf, L.w, backsub, s) : launch_step<double>(v, L.bd, L.w, backsub, s);
    }
    case CT_KERNEL_CANDIDATE:
      if (pl->num_groups > 0) return CT_OK;  // the fused step computed the candidate
      return mixed ? launch_candidate<float>(v, L.bf, L.w, s) : launch_candidate<double>(v, L.bd, L.w, s);
    case CT_KERNEL_LM_END:
      return launch_lm_end(pl->init, L.w, s);
    default:
      return CT_OK;
  }
}

// Whether a group launches any kernel in this plan (the point evaluation is
// part of the fused evaluation; the refinement only runs in the mixed form).
bool group_launches(const ct_ba_plan* pl, int group) {
  if (group == CT_KERNEL_EVAL_POINTS || group == CT_KERNEL_CANDIDATE) return pl->num_groups == 0;
  if (group == CT_KERNEL_REFINE) return refining(pl);
  if (group == CT_KERNEL_DOGLEG) return dogleg_options(&pl->options);
  return true;
}

// Groups that compute and test a step; iteration 0 only evaluates.
bool step_group(int group) {
  return group == CT_KERNEL_RCS_SOLVE || group == CT_KERNEL_REFINE || group == CT_KERNEL_DOGLEG ||
         group == CT_KERNEL_STEP || group == CT_KERNEL_CANDIDATE;
}

// One LM iteration.  evaluation_only (host-driven iteration 0) leaves out the
// step groups, which would return at once; a captured graph keeps them.
int enqueue_iteration(ct_ba_plan* pl, hipStream_t s, bool evaluation_only = false) {
  for (int g = 0; g < CT_KERNEL_COUNT; ++g) {
    if (evaluation_only && step_group(g)) continue;
    CT_TRY(enqueue_group(pl, g, s));
  }
  return CT_OK;
}

int read_state(ct_ba_plan* pl, hipStream_t s) { return lm_driver_read_state(pl->driver, s); }

template <typename T>
int copy_back(std::vector<T>& h, const T* d, size_t count) {
  h.resize(count);
  if (count) CT_HIP(hipMemcpy(h.data(), d, sizeof(T) * count, hipMemcpyDeviceToHost));
  return CT_OK;
}

// Every index the kernels follow is checked once on the host, so a malformed
// table is an argument error instead of an out-of-bounds access.
int validate_point_indices(const ct_ba_problem& p) {
  const int64_t C = p.num_cameras, P = p.num_points, N = p.num_observations;
  if (N == 0) return CT_OK;
  std::vector<int32_t> oc, op, po;
  CT_TRY(copy_back(oc, p.obs_camera, N));
  CT_TRY(copy_back(op, p.obs_point, N));
  CT_TRY(copy_back(po, p.point_offsets, P + 1));
  CT_CHECK_ARG(po[0] == 0 && po[P] == N, "point_offsets must start at 0 and end at num_observations");
  for (int64_t i = 0; i < P; ++i) CT_CHECK_ARG(po[i] <= po[i + 1], "point_offsets must be non-decreasing");
  for (int64_t o = 0; o < N; ++o) {
    CT_CHECK_ARG(oc[o] >= 0 && oc[o] < C, "obs_camera out of range");
    CT_CHECK_ARG(op[o] >= 0 && op[o] < P && o >= po[op[o]] && o < po[op[o] + 1],
                 "obs_point must agree with point_offsets (observations grouped by point)");
  }
  return CT_OK;
}

int validate_camera_csr(const ct_ba_problem& p, const std::vector<int32_t>& camera_offsets) {
  const int64_t C = p.num_cameras, N = p.num_observations;
  if (N == 0) return CT_OK;
  std::vector<int32_t> oc, cobs;
  CT_TRY(copy_back(oc, p.obs_camera, N));
  CT_TRY(copy_back(cobs, p.camera_obs, N));
  CT_CHECK_ARG(camera_offsets[0] == 0 && camera_offsets[C] == N,
               "camera_offsets must start at 0 and end at num_observations");
  for (int64_t c = 0; c < C; ++c)
    CT_CHECK_ARG(camera_offsets[c] <= camera_offsets[c + 1], "camera_offsets must be non-decreasing");
  for (int64_t o = 0; o < N; ++o) CT_CHECK_ARG(cobs[o] >= 0 && cobs[o] < N, "camera_obs out of range");
  for (int64_t c = 0; c < C; ++c)
    for (int32_t t = camera_offsets[c]; t < camera_offsets[c + 1]; ++t)
      CT_CHECK_ARG(oc[cobs[t]] == c, "camera_obs must list each camera's observations");
  return CT_OK;
}

int validate_indices(const ct_ba_problem& p, const std::vector<int32_t>& camera_offsets) {
  CT_TRY(validate_point_indices(p));
  CT_TRY(validate_camera_csr(p, camera_offsets));
  const int64_t N = p.num_observations, T = p.num_terms;
  if (N == 0) return CT_OK;
  std::vector<int32_t> terms;
  CT_TRY(copy_back(terms, p.block_terms, 2 * T));
  for (int64_t t = 0; t < 2 * T; ++t) CT_CHECK_ARG(terms[t] >= 0 && terms[t] < N, "block_terms out of range");
  return CT_OK;
}

// Launch order of the camera-block items.  Workgroups are dispatched to the
// chip's L2 domains (XCDs) round-robin by id, so slot k runs on domain
// k % kDomains.  Block rows are cut into kDomains bands of about equal work;
// a band's items go to one domain, row panels of kPanelRows rows at a time
// with the column camera outermost, so the row cameras' Jacobians stay in that
// domain's L2 and each column camera's data serves the whole panel.  Results
// do not depend on the order (every item writes its own partial slot).
constexpr int kDomains = 8;
constexpr int kPanelRows = 4;

std::vector<int32_t> item_order(const std::vector<BlockItem>& items, const std::vector<int32_t>& offsets,
                                const std::vector<int32_t>& bc, int C) {
  const int B = (int)offsets.size() - 1;
  std::vector<int32_t> block_of(C * (size_t)C, -1);
  std::vector<double> row_work(C, 0.0);
  for (int blk = 0; blk < B; ++blk) {
    const int i = bc[2 * (size_t)blk], j = bc[2 * (size_t)blk + 1];
    block_of[(size_t)i * C + j] = blk;
    for (int k = offsets[blk]; k < offsets[blk + 1]; ++k) row_work[i] += items[k].end - items[k].begin + 1;
  }
  double total = 0.0;
  for (double v : row_work) total += v;
  // bands of consecutive rows with about total / kDomains work each
  std::vector<std::vector<int32_t>> seq(kDomains);
  int row = 0;
  double done = 0.0;
  for (int d = 0; d < kDomains && row < C; ++d) {
    const double target = total * (d + 1) / kDomains;
    int r1 = row;
    while (r1 < C && (done + row_work[r1] <= target || r1 == row || d == kDomains - 1)) done += row_work[r1++];
    for (int p0 = row; p0 < r1; p0 += kPanelRows) {
      const int p1 = p0 + kPanelRows < r1 ? p0 + kPanelRows : r1;
      for (int j = p0; j < C; ++j)
        for (int i = p0; i < p1 && i <= j; ++i) {
          const int blk = block_of[(size_t)i * C + j];
          if (blk < 0) continue;
          for (int k = offsets[blk]; k < offsets[blk + 1]; ++k) seq[d].push_back(k);
        }
    }
    row = r1;
  }
  // interleave: slot k takes the next item of domain k % kDomains (or of the
  // domain with the most items left once that one is exhausted)
  std::vector<int32_t> perm;
  perm.reserve(items.size());
  std::vector<size_t> next(kDomains, 0);
  while (perm.size() < items.size()) {
    int d = (int)(perm.size() % kDomains);
    if (next[d] >= seq[d].size()) {
      size_t best = 0;
      for (int e = 0; e < kDomains; ++e)
        if (seq[e].size() - next[e] > best) {
          best = seq[e].size() - next[e];
          d = e;
        }
    }
    perm.push_back(seq[d][next[d]++]);
  }
  return perm;
}

void split_range(std::vector<BlockItem>& out, int block, int kind, int begin, int end) {
  const int count = end - begin;
  if (count <= 0) return;
  const int pieces = (count + kItemMaxWork - 1) / kItemMaxWork;
  const int chunk = (count + pieces - 1) / pieces;
  for (int b = begin; b < end; b += chunk) out.push_back(BlockItem{block, kind, b, end < b + chunk ? end : b + chunk});
}

// Split every camera-pair block into work items of at most kItemMaxWork terms
// or observations (balanced within a block) and upload the list.  Reads the
// block and camera offsets back from the device once.
// Pack whole points, in order, into groups of at most kWave observations for
// the fused evaluation (one wavefront per group).  A point with more
// observations than a wavefront has lanes keeps the plan on the two-kernel
// evaluation (num_groups = 0).
int build_point_groups(ct_ba_plan* pl) {
  const ct_ba_problem& p = pl->problem;
  pl->num_groups = 0;
  if (p.num_points <= 0 || p.num_observations <= 0) return CT_OK;
  std::vector<int32_t> po((size_t)p.num_points + 1);
  CT_HIP(hipMemcpy(po.data(), p.point_offsets, sizeof(int32_t) * po.size(), hipMemcpyDeviceToHost));
  std::vector<int32_t> starts;
  int used = 0;
  for (int q = 0; q < p.num_points; ++q) {
    const int k = po[q + 1] - po[q];
    if (k == 0) pl->unobserved_points = true;
    if (k > kWave) return CT_OK;
    if (starts.empty() || used + k > kWave) {
      starts.push_back(q);
      used = 0;
    }
    used += k;
  }
  starts.push_back(p.num_points);
  CT_HIP(hipMemcpy(pl->layout.point_groups, starts.data(), sizeof(int32_t) * starts.size(), hipMemcpyHostToDevice));
  pl->num_groups = (int)starts.size() - 1;
  return CT_OK;
}

int build_items(ct_ba_plan* pl, bool diagonal_only) {
  const ct_ba_problem& p = pl->problem;
  const int B = p.num_blocks, C = p.num_cameras;
  std::vector<int32_t> bo(B + 1, 0), co(C + 1, 0), bc(2 * (size_t)B + 1, 0);
  if (B > 0) {
    CT_HIP(hipMemcpy(bo.data(), p.block_offsets, sizeof(int32_t) * (B + 1), hipMemcpyDeviceToHost));
    CT_HIP(hipMemcpy(bc.data(), p.block_cameras, sizeof(int32_t) * 2 * (size_t)B, hipMemcpyDeviceToHost));
  }
  if (C > 0 && p.num_observations > 0)
    CT_HIP(hipMemcpy(co.data(), p.camera_offsets, sizeof(int32_t) * (C + 1), hipMemcpyDeviceToHost));
  CT_TRY(validate_indices(p, co));
  // the assembly indexes blocks by camera pair: every pair i <= j, row-major
  if ((long long)B != (long long)C * (C + 1) / 2) {
    set_error("num_blocks must be C (C + 1) / 2: one block per camera pair i <= j");
    return CT_ERR_ARGUMENT;
  }
  std::vector<BlockItem> items;
  std::vector<int32_t> offsets(B + 1, 0);
  int expect_i = 0, expect_j = 0;
  for (int blk = 0; blk < B; ++blk) {
    offsets[blk] = (int32_t)items.size();
    const int ci = bc[2 * (size_t)blk], cj = bc[2 * (size_t)blk + 1];
    if (ci != expect_i || cj != expect_j) {
      set_error("block_cameras must list every camera pair i <= j in row-major order");
      return CT_ERR_ARGUMENT;
    }
    if (++expect_j == C) expect_j = ++expect_i;
    if (bo[blk + 1] < bo[blk] || bo[blk] < 0 || bo[blk + 1] > p.num_terms) {
      set_error("block_offsets must be non-decreasing within [0, num_terms]");
      return CT_ERR_ARGUMENT;
    }
    if (ci == cj) split_range(items, blk, 1, co[ci], co[ci + 1]);
    if (ci == cj || !diagonal_only) split_range(items, blk, 0, bo[blk], bo[blk + 1]);
  }
  offsets[B] = (int32_t)items.size();
  if (items.size() > max_items(&p)) {
    set_error("camera-block item bound exceeded");
    return CT_ERR_ARGUMENT;
  }
  pl->num_items = (int)items.size();
  if (!items.empty()) {
    CT_HIP(hipMemcpy(pl->layout.items, items.data(), sizeof(BlockItem) * items.size(), hipMemcpyHostToDevice));
    const std::vector<int32_t> perm = item_order(items, offsets, bc, C);
    std::vector<char> seen(items.size(), 0);
    for (int32_t k : perm) {
      if (k < 0 || k >= (int32_t)items.size() || seen[k]) {
        set_error("camera-block item order is not a permutation");
        return CT_ERR_ARGUMENT;
      }
      seen[k] = 1;
    }
    CT_HIP(hipMemcpy(pl->layout.item_perm, perm.data(), sizeof(int32_t) * perm.size(), hipMemcpyHostToDevice));
  }
  CT_HIP(hipMemcpy(pl->layout.block_item_offsets, offsets.data(), sizeof(int32_t) * (B + 1), hipMemcpyHostToDevice));
  return CT_OK;
}

}  // namespace

namespace ct {

// The point and camera CSR tables of a problem (no camera-pair tables), checked on the host.
int ba_check_tables(const ct_ba_problem& p) {
  CT_TRY(validate_point_indices(p));
  if (p.num_observations == 0) return CT_OK;
  std::vector<int32_t> co;
  CT_TRY(copy_back(co, p.camera_offsets, (size_t)p.num_cameras + 1));
  return validate_camera_csr(p, co);
}

}  // namespace ct

extern "C" {

int ct_abi_version(void) { return CT_ABI_VERSION; }

const char* ct_last_error(void) { return g_last_error.c_str(); }

int ct_ba_workspace_bytes(const ct_ba_problem* problem, const ct_lm_options* options, size_t* bytes) {
  clear_error();
  CT_TRY(validate_problem(problem));
  CT_TRY(validate_options(options));
  CT_CHECK_ARG(bytes != nullptr, "bytes is null");
  *bytes = carve(problem, options, nullptr).bytes;
  return CT_OK;
}

int ct_ba_plan_create(const ct_ba_problem* problem, const ct_lm_options* options, double* cameras, double* points,
                      void* workspace, size_t workspace_bytes, ct_stream_t stream, ct_ba_plan** plan) {
  clear_error();
  CT_TRY(validate_problem(problem));
  CT_TRY(validate_options(options));
  CT_CHECK_ARG(plan != nullptr, "plan is null");
  CT_CHECK_ARG(workspace != nullptr, "workspace is null");
  CT_CHECK_ARG(problem->num_cameras == 0 || cameras != nullptr, "cameras is null");
  CT_CHECK_ARG(problem->num_points == 0 || points != nullptr, "points is null");
  const size_t need = carve(problem, options, nullptr).bytes;
  if (workspace_bytes < need) {
    set_error("workspace too small: need " + std::to_string(need) + " bytes");
    return CT_ERR_WORKSPACE;
  }
  auto pl = std::make_unique<ct_ba_plan>();
  pl->problem = *problem;
  pl->options = *options;
  pl->view = make_view(problem);
  pl->layout = carve(problem, options, workspace);
  pl->user_cams = cameras;
  pl->user_pts = points;
  pl->stream = reinterpret_cast<hipStream_t>(stream);
  // the tables are read back below with synchronous copies, which do not wait
  // for a non-blocking stream: the caller's uploads on `stream` finish first
  CT_HIP(hipStreamSynchronize(pl->stream));
  pl->init.opt = *options;
  pl->init.num_obs_blocks = obs_blocks(problem);
  pl->init.num_cost_blocks = obs_blocks(problem);
  pl->init.num_eval_blocks = eval_blocks(problem);
  pl->init.num_step_blocks = step_blocks(problem);
  pl->init.num_cameras = problem->num_cameras;
  pl->init.n_cam_params = camera_stride(problem) * problem->num_cameras;
  pl->init.split_blocks = problem->camera_model != CT_CAMERA_ANGLE_AXIS;
  pl->init.n_pt_params = 3 * problem->num_points;
  CT_TRY(build_point_groups(pl.get()));
  // ITERATIVE_SCHUR applies S from the records when the plan has point groups
  pl->implicit = implicit_possible(problem, options) && pl->num_groups > 0;
  CT_TRY(build_items(pl.get(), pl->implicit));
  if (pl->implicit) {
    ImplicitPcg& a = pl->imp;
    a = implicit_pcg_layout(pl->layout.implicit_work, problem->num_cameras, problem->num_observations);
    a.obs_camera = problem->obs_camera;
    a.obs_point = problem->obs_point;
    a.point_offsets = problem->point_offsets;
    a.camera_offsets = problem->camera_offsets;
    a.point_groups = pl->layout.point_groups;
    a.camera_obs = problem->camera_obs;
    a.JP = pl->layout.bd.JP;
    a.S = pl->layout.w.S;
    a.Dc = pl->layout.w.Dc;
    a.rhs = pl->layout.w.rhs;
    a.x = pl->layout.w.dxc;
    a.C = problem->num_cameras;
    a.n = 9 * problem->num_cameras;
    a.num_groups = pl->num_groups;
    a.st = pl->layout.w.state;
    pl->layout.w.pcg_bar = a.epoch;  // the LM begin kernel advances this solve epoch
    CT_TRY(implicit_pcg_prepare(a, pl->stream));
    CT_HIP(hipStreamSynchronize(pl->stream));
  }
  if (pl->num_groups > 0) {
    const int fused_blocks = (pl->num_groups + kFusedThreads / kWave - 1) / (kFusedThreads / kWave);
    pl->init.num_cost_blocks = fused_blocks;
    pl->init.num_eval_blocks = fused_blocks;
    pl->init.num_step_blocks = fused_blocks;
    pl->init.num_obs_blocks = fused_blocks;
  }
  if (uses_cholesky(options) && problem->num_cameras > 0) {
    CT_TRY(rcs_cholesky_prepare(pl->layout.w.solve_tmp, 9 * problem->num_cameras, pl->stream));
    CT_HIP(hipStreamSynchronize(pl->stream));
  }
  if (!uses_cholesky(options) && problem->num_cameras > 0) {
    CT_TRY(rcs_pcg_prepare(pl->layout.pcg_work, 9 * problem->num_cameras, pl->stream));
    CT_HIP(hipStreamSynchronize(pl->stream));
  }
  ct_ba_plan* raw = pl.get();
  LMDriver& d = pl->driver;
  d.enqueue_iteration = [raw](hipStream_t s, bool evaluation_only) {
    return enqueue_iteration(raw, s, evaluation_only);
  };
  d.launch_init = [raw](hipStream_t s) {
    return launch_lm_init(raw->init, raw->layout.w, raw->user_cams, raw->user_pts, s);
  };
  d.launch_finalize = [raw](hipStream_t s) {
    return launch_lm_finalize(raw->init, raw->layout.w, raw->user_cams, raw->user_pts, s);
  };
  const int st = lm_driver_setup(d, *options, pl->stream, pl->layout.w.state, pl->layout.w.records);
  if (st != CT_OK) {
    lm_driver_release(d);
    return st;
  }
  *plan = pl.release();
  return CT_OK;
}

int ct_ba_plan_destroy(ct_ba_plan* plan) {
  clear_error();
  if (!plan) return CT_OK;
  lm_driver_release(plan->driver);
  delete plan;
  return CT_OK;
}

int ct_ba_plan_run(ct_ba_plan* pl, ct_lm_report* report, double* records) {
  clear_error();
  CT_CHECK_ARG(pl != nullptr, "plan is null");
  return lm_driver_run(pl->driver, report, records);
}

int ct_ba_plan_profile(ct_ba_plan* pl, int iterations, double* kernel_ms, int32_t* kernel_launches) {
  clear_error();
  CT_CHECK_ARG(pl != nullptr, "plan is null");
  CT_CHECK_ARG(iterations > 0 && iterations <= 4096, "iterations must be in [1, 4096]");
  CT_CHECK_ARG(kernel_ms != nullptr && kernel_launches != nullptr, "output arrays are null");
  hipStream_t s = pl->stream;
  Layout& L = pl->layout;
  for (int k = 0; k < CT_KERNEL_COUNT; ++k) {
    kernel_ms[k] = 0.0;
    kernel_launches[k] = 0;
  }
  // every iteration's events and kernels are enqueued before one synchronisation,
  // so no event waits on the host; iterations after LM finished are not counted
  const int per = CT_KERNEL_COUNT + 1;
  std::vector<hipEvent_t> ev((size_t)iterations * per, nullptr);
  int status = CT_OK;
  for (auto& e : ev)
    if (status == CT_OK && hipEventCreate(&e) != hipSuccess) {
      set_error("hipEventCreate failed");
      status = CT_ERR_HIP;
    }
  if (status == CT_OK) status = launch_lm_init(pl->init, L.w, pl->user_cams, pl->user_pts, s);
  if (status == CT_OK) status = enqueue_iteration(pl, s, true);  // iteration 0, not timed
  for (int it = 0; it < iterations && status == CT_OK; ++it) {
    hipEvent_t* e = &ev[(size_t)it * per];
    for (int g = 0; g < CT_KERNEL_COUNT && status == CT_OK; ++g) {
      hipEventRecord(e[g], s);
      status = enqueue_group(pl, g, s);
    }
    hipEventRecord(e[CT_KERNEL_COUNT], s);
  }
  if (status == CT_OK) status = read_state(pl, s);
  if (status == CT_OK) {
    // LM iterations 1..completed ran every group with work (a termination in
    // the begin kernel leaves that iteration's later groups without work)
    int completed = pl->driver.host_state->iteration - 1;
    if (completed > iterations) completed = iterations;
    for (int it = 0; it < completed; ++it) {
      hipEvent_t* e = &ev[(size_t)it * per];
      for (int k = 0; k < CT_KERNEL_COUNT; ++k) {
        if (!group_launches(pl, k)) continue;  // an empty group only times its events
        float ms = 0.0f;
        hipEventElapsedTime(&ms, e[k], e[k + 1]);
        kernel_ms[k] += ms;
        kernel_launches[k] += 1;
      }
    }
  }
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (status == CT_OK) status = launch_lm_finalize(pl->init, L.w, pl->user_cams, pl->user_pts, s);
  if (status == CT_OK) status = read_state(pl, s);
  return status;
}

int ct_ba_covariance_workspace_bytes(int32_t num_cameras, int64_t num_requests, size_t* bytes) {
  clear_error();
  CT_CHECK_ARG(bytes != nullptr && num_cameras >= 0 && num_requests >= 0, "bad argument");
  CT_CHECK_ARG((int64_t)num_cameras * 9 <= kCovMaxN, "too many cameras for the dense covariance of S");
  Carver cv(nullptr);
  cv.take<double>(covariance_work_doubles(9 * num_cameras));
  cv.take<int32_t>(4 * (size_t)num_requests);
  cv.take<int32_t>(2);
  *bytes = cv.offset + 256;
  return CT_OK;
}

int ct_ba_plan_covariance(ct_ba_plan* pl, const int32_t* requests, int64_t num_requests, double pivot_rtol,
                          double* out, int64_t* deficient_pivots, void* workspace, size_t workspace_bytes) {
  clear_error();
  CT_CHECK_ARG(pl != nullptr, "plan is null");
  CT_CHECK_ARG(dogleg_options(&pl->options) && pl->options.precision == CT_PLAIN_FP64,
               "covariance needs a plan created with the DOGLEG strategy and plain FP64 (undamped J^T J)");
  const ct_ba_problem& p = pl->problem;
  const int n = 9 * p.num_cameras;
  CT_CHECK_ARG(n <= kCovMaxN, "too many cameras for the dense covariance of S");
  CT_CHECK_ARG(pivot_rtol >= 0.0 && pivot_rtol < 1.0, "pivot_rtol must be in [0, 1)");
  CT_CHECK_ARG(num_requests >= 0 && (num_requests == 0 || (requests != nullptr && out != nullptr)),
               "requests / out are null");
  CT_CHECK_ARG(num_requests <= INT32_MAX, "too many requests");
  for (int64_t r = 0; r < num_requests; ++r) {
    const int32_t* q = requests + 4 * r;
    for (int side = 0; side < 2; ++side) {
      const int kind = q[2 * side], idx = q[2 * side + 1];
      CT_CHECK_ARG(kind == 0 || kind == 1, "request kind must be 0 (camera) or 1 (point)");
      CT_CHECK_ARG(idx >= 0 && idx < (kind == 0 ? p.num_cameras : p.num_points), "request index out of range");
    }
  }
  size_t need = 0;
  CT_TRY(ct_ba_covariance_workspace_bytes(p.num_cameras, num_requests, &need));
  CT_CHECK_ARG(workspace != nullptr && workspace_bytes >= need, "workspace too small (ct_ba_covariance_workspace_bytes)");
  Carver cv(workspace);
  double* work = cv.take<double>(covariance_work_doubles(n));
  int32_t* dreq = cv.take<int32_t>(4 * (size_t)num_requests);
  int32_t* counts = cv.take<int32_t>(2);
  hipStream_t s = pl->stream;
  Layout& L = pl->layout;
  if (num_requests > 0)
    CT_HIP(hipMemcpyAsync(dreq, requests, sizeof(int32_t) * 4 * (size_t)num_requests, hipMemcpyHostToDevice, s));
  // the undamped system at the user's parameters: evaluation, point factors, S
  CT_TRY(launch_lm_init(pl->init, L.w, pl->user_cams, pl->user_pts, s));
  for (int g : {CT_KERNEL_EVAL_OBS, CT_KERNEL_EVAL_POINTS, CT_KERNEL_CAMERA_ITEMS, CT_KERNEL_CAMERA_REDUCE})
    CT_TRY(enqueue_group(pl, g, s));
  if (n > 0) CT_TRY(launch_rcs_cholesky_solve(L.w.S, nullptr, nullptr, n, L.w.state, L.w.solve_tmp, s));
  CovLaunch c{};
  c.L = L.w.S;
  c.n = n;
  c.JP = L.bd.JP;
  c.Linv = L.bd.Linv;
  c.obs_camera = p.obs_camera;
  c.point_offsets = p.point_offsets;
  c.camera_mask = p.camera_mask;
  c.point_mask = p.point_mask;
  c.num_points = p.num_points;
  c.requests = dreq;
  c.num_requests = num_requests;
  c.pivot_rtol = pivot_rtol;
  c.out = out;
  c.work = work;
  c.counts = counts;
  c.st = L.w.state;
  CT_TRY(launch_covariance(c, s));
  int32_t flags[3] = {0, 0, 0};  // singular, deficient point pivots, deficient pivots of S
  CT_HIP(hipMemcpyAsync(&flags[0], &L.w.state->step_invalid, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  CT_HIP(hipMemcpyAsync(&flags[1], counts, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  CT_HIP(hipStreamSynchronize(s));
  if (!(flags[0] || flags[1] || flags[2])) {
    if (deficient_pivots) *deficient_pivots = 0;
    return CT_OK;
  }
  if (flags[1] == 0 && n > 0) {
    // the points are fine and S is singular: count its null directions on S
    // rebuilt from the assembly's partial blocks, by pivoted Cholesky
    CT_TRY(enqueue_group(pl, CT_KERNEL_CAMERA_REDUCE, s));
    CT_TRY(launch_covariance_rank(L.w.S, n, p.camera_mask, pivot_rtol, counts, s));
    CT_HIP(hipMemcpyAsync(&flags[2], counts + 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    CT_HIP(hipStreamSynchronize(s));
  } else {
    flags[2] = 0;  // S is built on singular point blocks: only the points' null directions are counted
  }
  const int64_t def = (int64_t)flags[1] + flags[2];
  if (deficient_pivots) *deficient_pivots = def;
  set_error("J^T J is singular: " + std::to_string(def) + " null direction(s) below the rank threshold");
  return CT_ERR_NOT_SPD;
}

int ct_ba_evaluate(const ct_ba_problem* problem, const double* cameras, const double* points, double* cost_out_host,
                   double* residuals, void* workspace, size_t workspace_bytes, ct_stream_t stream) {
  clear_error();
  CT_TRY(validate_problem(problem));
  CT_CHECK_ARG(cost_out_host != nullptr, "cost_out_host is null");
  CT_CHECK_ARG(workspace != nullptr, "workspace is null");
  const int grid = eval_blocks(problem);
  CT_CHECK_ARG(workspace_bytes >= sizeof(double) * (size_t)(grid + 1), "workspace too small for partials");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  CT_TRY(validate_point_indices(*problem));
  BAView v = make_view(problem);
  double* part = static_cast<double*>(workspace);
  CT_TRY(launch_evaluate_only(v, cameras, points, residuals, part, s));
  std::vector<double> h(grid > 0 ? grid : 1, 0.0);
  if (grid > 0) {
    CT_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * grid, hipMemcpyDeviceToHost, s));
    CT_HIP(hipStreamSynchronize(s));
  }
  double c = 0.0;
  for (int i = 0; i < grid; ++i) c += h[i];
  *cost_out_host = c;
  return CT_OK;
}

int ct_ba_evaluate_jacobian(const ct_ba_problem* problem, const double* cameras, const double* points,
                            const int32_t* row_of_obs, const int64_t* camera_columns, const int64_t* point_columns,
                            int64_t num_columns, int32_t apply_loss, double* jacobian, double* residuals,
                            double* cost_out_host, void* workspace, size_t workspace_bytes, ct_stream_t stream) {
  clear_error();
  CT_TRY(validate_problem(problem));
  CT_CHECK_ARG(cost_out_host != nullptr && workspace != nullptr, "null argument");
  CT_CHECK_ARG(num_columns >= 0, "num_columns must be >= 0");
  const int64_t N = problem->num_observations;
  CT_CHECK_ARG(N == 0 || (row_of_obs && residuals && camera_columns && point_columns),
               "row_of_obs, residuals and the column maps are required");
  CT_CHECK_ARG(N == 0 || num_columns == 0 || jacobian != nullptr, "jacobian is null");
  const int grid = obs_blocks(problem);
  CT_CHECK_ARG(workspace_bytes >= sizeof(double) * (size_t)(grid + 1), "workspace too small for partials");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  CT_HIP(hipStreamSynchronize(s));  // the caller's uploads on `stream` precede the host reads below
  CT_TRY(validate_point_indices(*problem));
  // every row index, and every column index the kernel follows, is checked once here
  std::vector<int32_t> rows;
  CT_TRY(copy_back(rows, row_of_obs, N));
  std::vector<char> seen(N, 0);
  for (int64_t o = 0; o < N; ++o) {
    CT_CHECK_ARG(rows[o] >= 0 && rows[o] < N && !seen[rows[o]], "row_of_obs must be a permutation of 0..N-1");
    seen[rows[o]] = 1;
  }
  std::vector<int64_t> cc, pc;
  CT_TRY(copy_back(cc, camera_columns, 9 * (size_t)problem->num_cameras));
  CT_TRY(copy_back(pc, point_columns, 3 * (size_t)problem->num_points));
  for (int64_t c : cc) CT_CHECK_ARG(c >= -1 && c < num_columns, "camera column out of range");
  for (int64_t c : pc) CT_CHECK_ARG(c >= -1 && c < num_columns, "point column out of range");
  BAView v = make_view(problem);
  double* part = static_cast<double*>(workspace);
  CT_TRY(launch_evaluate_jacobian(v, cameras, points, row_of_obs, camera_columns, point_columns, num_columns,
                                  apply_loss, jacobian, residuals, part, s));
  std::vector<double> h(grid > 0 ? grid : 1, 0.0);
  if (grid > 0) {
    CT_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * grid, hipMemcpyDeviceToHost, s));
    CT_HIP(hipStreamSynchronize(s));
  }
  double c = 0.0;
  for (int i = 0; i < grid; ++i) c += h[i];
  *cost_out_host = c;
  return CT_OK;
}

}  // extern "C"
