// Batched curve fits y = exp(m x + c) on the device: many independent problems
// of uneven size, each the Levenberg-Marquardt solve of solver.py:35-357 over
// one robust residual block per sample (examples/robust_curve_fitting.py:80-97),
// all in one persistent launch.
//
// Problems are taken from two work queues in the caller's order (largest first
// balances uneven sizes): the leading slots are solved by a whole workgroup
// each, the trailing ones (the small problems, when the order is largest
// first) by one wave each, so the workgroup size no longer has to suit the
// largest and the smallest problem at once.  A team (workgroup or wave) runs
// the whole LM loop of its problem with the LM state in LDS: per pass one
// evaluation at x (residual, corrected Jacobian, J^T J, J^T r), the damped 2x2
// normal equations, one candidate cost at x + dx.  For two parameters the
// damped normal equations have the solution of the reference's DENSE_QR
// least-squares step (linear.py:359-371).
//
// FP64 form: everything in FP64 from FP64 samples.  Mixed form: the samples
// are stored in FP32 and the evaluation at x (exp, residual, robust
// correction, Jacobian, J^T J and J^T r) runs in FP32; every cost is FP64 --
// the candidate cost that decides acceptance is evaluated in FP64 from the
// FP32 samples, and the cost at x is the FP64 cost of the last accepted
// candidate (evaluated once at the start), so the gain ratio and every cost
// test compare FP64 costs of the FP32-rounded problem.
#include <string>
#include <vector>

#include "ba_kernels.h"
#include "lm_logic.h"
#include "pcg_device.h"

namespace ct {

namespace {

struct CurveArgs {
  int32_t B;
  int32_t group_end;       // processing slots [0, group_end) per workgroup, [group_end, B) per wave
  int64_t num_samples;
  const int64_t* __restrict__ offsets;
  const double* __restrict__ x;     // FP64 form
  const double* __restrict__ y;
  const float* __restrict__ x32;    // mixed form
  const float* __restrict__ y32;
  const int32_t* __restrict__ order;
  ct_loss loss;
  ct_lm_options opt;
  double* params;          // [B][2]
  ct_lm_report* results;   // [B] or null
  double* records;         // [B][max_iter+2][CT_REC_FIELDS] or null
  unsigned int* queue;     // [2] work-queue counters: workgroup slots, wave slots
};

// Corrector of a 1-residual block (losses.py:158-193); returns 1/2 rho0 and
// rescales r, j0, j1 in place.
__device__ __forceinline__ double robustify1(const ct_loss& L, double& r, double& j0, double& j1) {
  const double s = r * r;
  if (L.type == CT_LOSS_TRIVIAL && L.scale == 1.0) return 0.5 * s;
  double rho1, rho2;
  const double rho0 = loss_rho(L, s, &rho1, &rho2);
  const double sq = sqrt(fmax(rho1, 0.0));
  double scale = sq, alpha_s = 0.0;
  if (s > 0.0 && rho2 > 0.0) {
    const double alpha = 1.0 - sqrt(1.0 + 2.0 * s * rho2 / rho1);
    scale = sq / (1.0 - alpha);
    alpha_s = alpha / s;
  }
  j0 = sq * (j0 - r * (alpha_s * r * j0));
  j1 = sq * (j1 - r * (alpha_s * r * j1));
  r *= scale;
  return 0.5 * rho0;
}

__device__ __forceinline__ double robust_cost1(const ct_loss& L, double r) {
  const double s = r * r;
  if (L.type == CT_LOSS_TRIVIAL && L.scale == 1.0) return 0.5 * s;
  double rho1, rho2;
  return 0.5 * loss_rho(L, s, &rho1, &rho2);
}

// rho'(s) and rho''(s) of the losses (losses.py:21-133) in FP32, for the FP32
// corrector of the mixed form (rho itself is only needed for costs, which are
// FP64).
__device__ __forceinline__ void loss_rho12_f32(const ct_loss& L, float s, float& rho1, float& rho2) {
  float r1, r2;
  switch (L.type) {
    case CT_LOSS_HUBER: {
      const float a = (float)L.a;
      if (s <= a * a) {
        r1 = 1.0f; r2 = 0.0f;
      } else {
        const float ss = sqrtf(fmaxf(s, 1.17549435e-38f));
        r1 = a / ss;
        r2 = -0.5f * a / (ss * s);
      }
      break;
    }
    case CT_LOSS_SOFTLONE: {
      const float b = (float)(L.a * L.a), z = 1.0f + s / b, sz = sqrtf(z);
      r1 = 1.0f / sz; r2 = -0.5f / (b * z * sz);
      break;
    }
    case CT_LOSS_CAUCHY: {
      const float b = (float)(L.a * L.a), z = 1.0f + s / b;
      r1 = 1.0f / z; r2 = -1.0f / (b * z * z);
      break;
    }
    case CT_LOSS_ARCTAN: {
      const float a = (float)L.a, z = s / a, d = 1.0f + z * z;
      r1 = 1.0f / d; r2 = -2.0f * z / (a * d * d);
      break;
    }
    case CT_LOSS_TOLERANT: {
      const float b = (float)L.b, z = (s - (float)L.a) / b;
      r1 = 1.0f / (1.0f + expf(-z));
      r2 = r1 * (1.0f - r1) / b;
      break;
    }
    case CT_LOSS_TUKEY: {
      const float a2 = (float)(L.a * L.a);
      if (s <= a2) {
        const float in = 1.0f - s / a2;
        r1 = in * in; r2 = -2.0f * in / a2;
      } else {
        r1 = 0.0f; r2 = 0.0f;
      }
      break;
    }
    default:
      r1 = 1.0f; r2 = 0.0f;
  }
  rho1 = (float)L.scale * r1;
  rho2 = (float)L.scale * r2;
}

// FP32 corrector (the steps of robustify1 without the cost).
__device__ __forceinline__ void robustify1_f32(const ct_loss& L, float& r, float& j0, float& j1) {
  if (L.type == CT_LOSS_TRIVIAL && L.scale == 1.0) return;
  const float s = r * r;
  float rho1, rho2;
  loss_rho12_f32(L, s, rho1, rho2);
  const float sq = sqrtf(fmaxf(rho1, 0.0f));
  float scale = sq, alpha_s = 0.0f;
  if (s > 0.0f && rho2 > 0.0f) {
    const float alpha = 1.0f - sqrtf(1.0f + 2.0f * s * rho2 / rho1);
    scale = sq / (1.0f - alpha);
    alpha_s = alpha / s;
  }
  j0 = sq * (j0 - r * (alpha_s * r * j0));
  j1 = sq * (j1 - r * (alpha_s * r * j1));
  r *= scale;
}

// Deterministic sum of K values over the workgroup (wave trees, then waves in
// index order); results valid in every thread.
template <int T, int K>
__device__ __forceinline__ void block_sum_k(double (&v)[K], double (*scratch)[T / kWave], double* out) {
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double t = wave_sum(v[k]);
    if (lane == 0) scratch[k][wid] = t;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    double t = 0.0;
    for (int w = 0; w < T / kWave; ++w) t += scratch[threadIdx.x][w];
    out[threadIdx.x] = t;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = out[k];
  __syncthreads();
}

// A team solves one problem: a whole workgroup of T threads (N == T) or one
// wave (N == kWave).  team_sum leaves the K sums in every lane (the xor
// butterfly of wave_sum adds the same pairs in every lane, so all lanes hold
// the same bits); team_sync makes the LDS writes of lane 0 visible to the team.
template <int T, int N, int K>
__device__ __forceinline__ void team_sum(double (&v)[K], double (*scratch)[T / kWave], double* out) {
  if constexpr (N == kWave) {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
  } else {
    block_sum_k<T, K>(v, scratch, out);
  }
}

template <int N>
__device__ __forceinline__ void team_sync() {
  if constexpr (N == kWave) {
    lds_wave_sync();
  } else {
    __syncthreads();
  }
}

__device__ __forceinline__ double damping_1(double colsq, const LMState& st) {
  const double dg = st.jacobi_scaling ? colsq : 1.0;
  return fmin(fmax(dg, st.min_diag), st.max_diag) / fmax(st.radius, 2.220446049250313e-16);
}

// FP64 robust cost of samples [lo, hi) at (m, c), this lane's share.
template <bool kMixed, int N>
__device__ __forceinline__ double lane_cost(const CurveArgs& a, int64_t lo, int64_t hi, int rank, double m,
                                            double c) {
  double cost = 0.0;
  for (int64_t i = lo + rank; i < hi; i += N) {
    const double xi = kMixed ? (double)a.x32[i] : a.x[i];
    const double yi = kMixed ? (double)a.y32[i] : a.y[i];
    cost += robust_cost1(a.loss, yi - exp(m * xi + c));
  }
  return cost;
}

// The whole LM loop of problem p by one team; rank is the lane's index in it.
template <bool kMixed, int T, int N>
__device__ void solve_problem(const CurveArgs& a, int p, int rank, LMState& st, double (&xs)[2][2], double* step,
                              double (*scratch)[T / kWave], double* sums) {
  const int passes = a.opt.max_num_iterations + 2;
  const int64_t lo = a.offsets[p], hi = a.offsets[p + 1];
  if (!(lo >= 0 && hi >= lo && hi <= a.num_samples)) {
    if (rank == 0 && a.results) {
      ct_lm_report r{};
      r.termination = CT_TERM_FAILURE;
      r.reason = CT_REASON_INVALID_INPUT;
      a.results[p] = r;
    }
    return;
  }
  double* rec = a.records ? a.records + (size_t)p * passes * CT_REC_FIELDS : nullptr;
  if (rank == 0) {
    lm_state_init(&st, a.opt);
    xs[0][0] = a.params[2 * (size_t)p];
    xs[0][1] = a.params[2 * (size_t)p + 1];
  }
  team_sync<N>();
  double cost0 = 0.0;  // mixed form: FP64 cost at the start values
  if constexpr (kMixed) {
    double c0[1] = {lane_cost<true, N>(a, lo, hi, rank, xs[0][0], xs[0][1])};
    team_sum<T, N, 1>(c0, scratch, sums);
    cost0 = c0[0];
  }
  for (int pass = 0; pass < passes; ++pass) {
    // evaluation at x: g = J^T r, J^T J (FP64 form: and the cost)
    const double m = xs[st.cur][0], c = xs[st.cur][1];
    double v[6];
    if constexpr (kMixed) {
      const float mf = (float)m, cf = (float)c;
      float g0 = 0.0f, g1 = 0.0f, a00 = 0.0f, a01 = 0.0f, a11 = 0.0f;
      for (int64_t i = lo + rank; i < hi; i += N) {
        const float xi = a.x32[i];
        const float e = __expf(mf * xi + cf);
        float r = a.y32[i] - e;
        float j0 = -xi * e, j1 = -e;
        robustify1_f32(a.loss, r, j0, j1);
        a00 += j0 * j0;
        a01 += j0 * j1;
        a11 += j1 * j1;
        g0 += j0 * r;
        g1 += j1 * r;
      }
      v[0] = 0.0;
      v[1] = g0;
      v[2] = g1;
      v[3] = a00;
      v[4] = a01;
      v[5] = a11;
    } else {
      double cost = 0.0, g0 = 0.0, g1 = 0.0, a00 = 0.0, a01 = 0.0, a11 = 0.0;
      for (int64_t i = lo + rank; i < hi; i += N) {
        const double xi = a.x[i];
        const double e = exp(m * xi + c);
        double r = a.y[i] - e;
        double j0 = -xi * e, j1 = -e;
        cost += robustify1(a.loss, r, j0, j1);
        a00 += j0 * j0;
        a01 += j0 * j1;
        a11 += j1 * j1;
        g0 += j0 * r;
        g1 += j1 * r;
      }
      v[0] = cost;
      v[1] = g0;
      v[2] = g1;
      v[3] = a00;
      v[4] = a01;
      v[5] = a11;
    }
    team_sum<T, N, 6>(v, scratch, sums);
    if (rank == 0) {
      // mixed form: the FP64 cost at x is that of the last accepted candidate
      const double cost = kMixed ? (pass == 0 ? cost0 : st.ev_current) : v[0];
      lm_begin_update(&st, cost, v[1] * v[1] + v[2] * v[2], fmax(fabs(v[1]), fabs(v[2])), rec);
    }
    team_sync<N>();
    if (st.done) break;
    if (st.skip_step) {
      team_sync<N>();
      if (rank == 0) lm_skip_update(&st);
      team_sync<N>();
      continue;
    }
    if (rank == 0) {
      // (J^T J + D) dx = -g by a 2x2 Cholesky (FP64)
      const double h00 = v[3] + damping_1(v[3], st), h01 = v[4], h11 = v[5] + damping_1(v[5], st);
      const double l00 = sqrt(h00);
      const double l10 = h01 / l00;
      const double l11 = sqrt(h11 - l10 * l10);
      const double z0 = -v[1] / l00;
      const double z1 = (-v[2] - l10 * z0) / l11;
      const double d1 = z1 / l11;
      const double d0 = (z0 - l10 * d1) / l00;
      const double n0 = xs[st.cur][0] + d0, n1 = xs[st.cur][1] + d1;
      xs[1 - st.cur][0] = n0;
      xs[1 - st.cur][1] = n1;
      // model decrease -r.u - u.u/2 with u = J dx, from the reduced sums
      step[0] = -(v[1] * d0 + v[2] * d1) - 0.5 * (d0 * d0 * v[3] + 2.0 * d0 * d1 * v[4] + d1 * d1 * v[5]);
      step[1] = d0 * d0 + d1 * d1;
      step[2] = n0 * n0 + n1 * n1;
      step[3] = (isfinite(d0) && isfinite(d1)) ? 0.0 : 1.0;
      st.linear_iterations = 1;
    }
    team_sync<N>();
    // candidate cost at x + dx (FP64)
    double cc[1] = {0.0};
    if (step[3] == 0.0) cc[0] = lane_cost<kMixed, N>(a, lo, hi, rank, xs[1 - st.cur][0], xs[1 - st.cur][1]);
    team_sum<T, N, 1>(cc, scratch, sums);
    if (rank == 0) lm_end_update(&st, cc[0], step[0], step[1], step[2], step[3] != 0.0, rec);
    team_sync<N>();
    if (st.done) break;
  }
  if (rank == 0) {
    a.params[2 * (size_t)p] = xs[st.cur][0];
    a.params[2 * (size_t)p + 1] = xs[st.cur][1];
    if (a.results) {
      ct_lm_report r;
      r.termination = st.termination;
      r.reason = st.reason;
      r.num_records = st.num_records;
      r.num_successful_steps = st.num_successful;
      r.num_unsuccessful_steps = st.num_unsuccessful;
      r.num_residual_evaluations = st.num_residual_evals;
      r.num_jacobian_evaluations = st.num_jacobian_evals;
      r.num_linear_solves = st.num_linear_solves;
      r.initial_cost = st.initial_cost;
      r.final_cost = st.best_cost;
      r.device_seconds = (double)(st.t_end - st.t0) * 1e-8;
      a.results[p] = r;
    }
  }
  team_sync<N>();
}

template <bool kMixed, int T>
__global__ __launch_bounds__(T) void k_curve_lm(CurveArgs a) {
  constexpr int W = T / kWave;
  __shared__ LMState st[W];
  __shared__ double scratch[6][W];
  __shared__ double sums[6];
  __shared__ double xs[W][2][2];
  __shared__ double step[W][4];  // model decrease, |dx|^2, |x'|^2, nonfinite
  __shared__ int slot;
  const int tid = threadIdx.x, lane = tid % kWave, wid = tid / kWave;
  // leading slots: one problem per workgroup
  for (;;) {
    if (tid == 0) slot = (int)atomicAdd(&a.queue[0], 1u);
    __syncthreads();
    const int q = slot;
    __syncthreads();
    if (q >= a.group_end) break;
    const int p = a.order ? a.order[q] : q;
    if (p < 0 || p >= a.B) continue;
    solve_problem<kMixed, T, T>(a, p, tid, st[0], xs[0], step[0], scratch, sums);
  }
  // trailing slots: one problem per wave, each wave pulling its own
  for (;;) {
    unsigned int got = 0;
    if (lane == 0) got = atomicAdd(&a.queue[1], 1u);
    const int q = a.group_end + (int)__shfl(got, 0, kWave);
    if (q >= a.B) break;
    const int p = a.order ? a.order[q] : q;
    if (p < 0 || p >= a.B) continue;
    solve_problem<kMixed, T, kWave>(a, p, lane, st[wid], xs[wid], step[wid], scratch, sums);
  }
}

__global__ void k_reset_counters(unsigned int* c) {
  if (threadIdx.x < 2) c[threadIdx.x] = 0u;
}

template <bool kMixed>
int launch_curve(const CurveArgs& a, int threads, int groups, hipStream_t s) {
  switch (threads) {
    case 64:
      hipLaunchKernelGGL((k_curve_lm<kMixed, 64>), dim3(groups), dim3(64), 0, s, a);
      break;
    case 128:
      hipLaunchKernelGGL((k_curve_lm<kMixed, 128>), dim3(groups), dim3(128), 0, s, a);
      break;
    case 256:
      hipLaunchKernelGGL((k_curve_lm<kMixed, 256>), dim3(groups), dim3(256), 0, s, a);
      break;
    case 512:
      hipLaunchKernelGGL((k_curve_lm<kMixed, 512>), dim3(groups), dim3(512), 0, s, a);
      break;
    default:
      set_error("schedule.threads must be 64, 128, 256 or 512");
      return CT_ERR_ARGUMENT;
  }
  CT_LAUNCH_CHECK();
  return CT_OK;
}

}  // namespace

}  // namespace ct

using namespace ct;

extern "C" {

int ct_curve_workspace_bytes(const ct_curve_batch* batch, size_t* bytes) {
  clear_error();
  CT_CHECK_ARG(batch != nullptr && bytes != nullptr, "null argument");
  *bytes = 256;
  return CT_OK;
}

int ct_curve_solve(const ct_curve_batch* batch, const ct_lm_options* options, const ct_batch_schedule* schedule,
                   double* params, ct_lm_report* results, double* records, void* workspace, size_t workspace_bytes,
                   ct_stream_t stream) {
  clear_error();
  CT_CHECK_ARG(batch != nullptr && options != nullptr, "null argument");
  CT_CHECK_ARG(batch->num_problems >= 0 && batch->num_samples >= 0, "negative batch size");
  CT_CHECK_ARG(batch->loss.type >= CT_LOSS_TRIVIAL && batch->loss.type <= CT_LOSS_TUKEY, "unknown loss type");
  CT_CHECK_ARG(options->max_num_iterations >= 0, "max_num_iterations must be >= 0");
  CT_CHECK_ARG(options->precision == CT_PLAIN_FP64 || options->precision == CT_MIXED, "unknown precision");
  CT_CHECK_ARG(options->initial_trust_region_radius > 0.0, "initial_trust_region_radius must be > 0");
  if (options->trust_region_strategy != CT_LEVENBERG_MARQUARDT) {
    ::ct::set_error("DOGLEG is on the device path for bundle adjustment only");
    return CT_ERR_UNSUPPORTED;
  }
  if (options->debug_invalid_steps != 0) {
    ::ct::set_error("debug_invalid_steps is a bundle-adjustment test hook");
    return CT_ERR_UNSUPPORTED;
  }
  if (batch->num_problems == 0) return CT_OK;
  CT_CHECK_ARG(batch->offsets != nullptr && params != nullptr, "offsets / params are null");
  const bool mixed = options->precision == CT_MIXED;
  if (mixed) {
    CT_CHECK_ARG(batch->num_samples == 0 || (batch->x32 != nullptr && batch->y32 != nullptr),
                 "the mixed form reads the FP32 samples: x32 / y32 are null");
  } else {
    CT_CHECK_ARG(batch->num_samples == 0 || (batch->x != nullptr && batch->y != nullptr), "x / y are null");
  }
  CT_CHECK_ARG(workspace != nullptr && workspace_bytes >= 256, "workspace too small (ct_curve_workspace_bytes)");
  int threads = 256, groups = 0, wave_problems = 0;
  if (schedule) {
    threads = schedule->threads;
    groups = schedule->groups;
    wave_problems = schedule->wave_problems;
  }
  CT_CHECK_ARG(groups >= 0, "schedule.groups must be >= 0");
  CT_CHECK_ARG(wave_problems >= 0, "schedule.wave_problems must be >= 0");
  if (wave_problems > batch->num_problems) wave_problems = batch->num_problems;
  if (groups == 0) {
    int dev = 0, cus = 0;
    CT_HIP(hipGetDevice(&dev));
    CT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    groups = 8 * cus;
  }
  if (groups > batch->num_problems) groups = batch->num_problems;
  CurveArgs a;
  a.B = batch->num_problems;
  a.group_end = batch->num_problems - wave_problems;
  a.num_samples = batch->num_samples;
  a.offsets = batch->offsets;
  a.x = batch->x;
  a.y = batch->y;
  a.x32 = batch->x32;
  a.y32 = batch->y32;
  a.order = batch->order;
  a.loss = batch->loss;
  a.opt = *options;
  a.params = params;
  a.results = results;
  a.records = records;
  a.queue = static_cast<unsigned int*>(workspace);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, s, a.queue);
  CT_LAUNCH_CHECK();
  return mixed ? launch_curve<true>(a, threads, groups, s) : launch_curve<false>(a, threads, groups, s);
}

int ct_curve_validate(const ct_curve_batch* batch, int32_t* first_bad_problem) {
  clear_error();
  CT_CHECK_ARG(batch != nullptr && first_bad_problem != nullptr, "null argument");
  *first_bad_problem = -1;
  const int B = batch->num_problems;
  if (B <= 0) return CT_OK;
  std::vector<int64_t> off(B + 1);
  CT_HIP(hipMemcpy(off.data(), batch->offsets, sizeof(int64_t) * (B + 1), hipMemcpyDeviceToHost));
  for (int p = 0; p < B; ++p) {
    if (off[p] < 0 || off[p + 1] < off[p] || off[p + 1] > batch->num_samples) {
      *first_bad_problem = p;
      set_error("offsets must be non-decreasing within [0, num_samples]; first bad problem " + std::to_string(p));
      return CT_ERR_ARGUMENT;
    }
  }
  if (batch->order) {
    std::vector<int32_t> ord(B);
    std::vector<char> seen(B, 0);
    CT_HIP(hipMemcpy(ord.data(), batch->order, sizeof(int32_t) * B, hipMemcpyDeviceToHost));
    for (int q = 0; q < B; ++q) {
      if (ord[q] < 0 || ord[q] >= B || seen[ord[q]]) {
        set_error("order must be a permutation of 0..num_problems-1");
        return CT_ERR_ARGUMENT;
      }
      seen[ord[q]] = 1;
    }
  }
  return CT_OK;
}

}  // extern "C"
