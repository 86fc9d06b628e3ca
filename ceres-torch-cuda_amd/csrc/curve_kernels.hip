// Batched curve fits y = exp(m x + c) on the device: many independent problems
// of uneven size, each the Levenberg-Marquardt solve of solver.py:35-357 over
// one robust residual block per sample (examples/robust_curve_fitting.py:80-97),
// all in one persistent launch.
//
// Persistent workgroups take problems from a work queue in the caller's order
// (largest first balances uneven sizes).  A workgroup runs the whole LM loop
// for its problem with the LM state in LDS: per pass one evaluation at x
// (residual, corrected Jacobian, J^T J, J^T r, FP64 cost), the damped 2x2
// normal equations, one candidate evaluation at x + dx.  For two parameters
// the damped normal equations have the solution of the reference's DENSE_QR
// least-squares step (linear.py:359-371).  Mixed precision accumulates J^T J in
// FP32; residuals, the gradient, costs and the gain ratio are always FP64.
#include <string>
#include <vector>

#include "ba_kernels.h"
#include "lm_logic.h"

namespace ct {

namespace {

struct CurveArgs {
  int32_t B;
  int64_t num_samples;
  const int64_t* __restrict__ offsets;
  const double* __restrict__ x;
  const double* __restrict__ y;
  const int32_t* __restrict__ order;
  ct_loss loss;
  ct_lm_options opt;
  double* params;          // [B][2]
  ct_lm_report* results;   // [B] or null
  double* records;         // [B][max_iter+2][CT_REC_FIELDS] or null
  unsigned int* queue;     // work-queue counter
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

__device__ __forceinline__ double damping_1(double colsq, const LMState& st) {
  const double dg = st.jacobi_scaling ? colsq : 1.0;
  return fmin(fmax(dg, st.min_diag), st.max_diag) / fmax(st.radius, 2.220446049250313e-16);
}

template <typename real, int T>
__global__ __launch_bounds__(T) void k_curve_lm(CurveArgs a) {
  __shared__ LMState st;
  __shared__ double scratch[6][T / kWave];
  __shared__ double sums[6];
  __shared__ double xs[2][2];
  __shared__ double step[5];  // model decrease, |dx|^2, |x'|^2, nonfinite, (unused)
  __shared__ int slot;
  const int tid = threadIdx.x;
  const int passes = a.opt.max_num_iterations + 2;
  for (;;) {
    if (tid == 0) slot = (int)atomicAdd(a.queue, 1u);
    __syncthreads();
    const int q = slot;
    __syncthreads();
    if (q >= a.B) break;
    const int p = a.order ? a.order[q] : q;
    if (p < 0 || p >= a.B) continue;
    const int64_t lo = a.offsets[p], hi = a.offsets[p + 1];
    if (!(lo >= 0 && hi >= lo && hi <= a.num_samples)) {
      if (tid == 0 && a.results) {
        ct_lm_report r{};
        r.termination = CT_TERM_FAILURE;
        r.reason = CT_REASON_INVALID_INPUT;
        a.results[p] = r;
      }
      continue;
    }
    double* rec = a.records ? a.records + (size_t)p * passes * CT_REC_FIELDS : nullptr;
    if (tid == 0) {
      lm_state_init(&st, a.opt);
      xs[0][0] = a.params[2 * (size_t)p];
      xs[0][1] = a.params[2 * (size_t)p + 1];
    }
    __syncthreads();
    for (int pass = 0; pass < passes; ++pass) {
      // evaluation at x: cost, g = J^T r (FP64), J^T J (real)
      const double m = xs[st.cur][0], c = xs[st.cur][1];
      double cost = 0.0, g0 = 0.0, g1 = 0.0;
      real a00 = 0, a01 = 0, a11 = 0;
      for (int64_t i = lo + tid; i < hi; i += T) {
        const double xi = a.x[i];
        const double e = exp(m * xi + c);
        double r = a.y[i] - e;
        double j0 = -xi * e, j1 = -e;
        cost += robustify1(a.loss, r, j0, j1);
        const real f0 = (real)j0, f1 = (real)j1;
        a00 += f0 * f0;
        a01 += f0 * f1;
        a11 += f1 * f1;
        g0 += j0 * r;
        g1 += j1 * r;
      }
      double v[6] = {cost, g0, g1, (double)a00, (double)a01, (double)a11};
      block_sum_k<T, 6>(v, scratch, sums);
      if (tid == 0) lm_begin_update(&st, v[0], v[1] * v[1] + v[2] * v[2], fmax(fabs(v[1]), fabs(v[2])), rec);
      __syncthreads();
      if (st.done) break;
      if (st.skip_step) {
        __syncthreads();
        if (tid == 0) lm_skip_update(&st);
        __syncthreads();
        continue;
      }
      if (tid == 0) {
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
      __syncthreads();
      // candidate cost at x + dx (FP64)
      const double mn = xs[1 - st.cur][0], cn = xs[1 - st.cur][1];
      double cc[1] = {0.0};
      if (step[3] == 0.0) {
        for (int64_t i = lo + tid; i < hi; i += T) cc[0] += robust_cost1(a.loss, a.y[i] - exp(mn * a.x[i] + cn));
      }
      block_sum_k<T, 1>(cc, scratch, sums);
      if (tid == 0) lm_end_update(&st, cc[0], step[0], step[1], step[2], step[3] != 0.0, rec);
      __syncthreads();
      if (st.done) break;
    }
    if (tid == 0) {
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
    __syncthreads();
  }
}

__global__ void k_reset_counter(unsigned int* c) {
  if (threadIdx.x == 0) *c = 0u;
}

template <typename real>
int launch_curve(const CurveArgs& a, int threads, int groups, hipStream_t s) {
  switch (threads) {
    case 64:
      hipLaunchKernelGGL((k_curve_lm<real, 64>), dim3(groups), dim3(64), 0, s, a);
      break;
    case 128:
      hipLaunchKernelGGL((k_curve_lm<real, 128>), dim3(groups), dim3(128), 0, s, a);
      break;
    case 256:
      hipLaunchKernelGGL((k_curve_lm<real, 256>), dim3(groups), dim3(256), 0, s, a);
      break;
    case 512:
      hipLaunchKernelGGL((k_curve_lm<real, 512>), dim3(groups), dim3(512), 0, s, a);
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
  CT_CHECK_ARG(batch->num_samples == 0 || (batch->x != nullptr && batch->y != nullptr), "x / y are null");
  CT_CHECK_ARG(workspace != nullptr && workspace_bytes >= 256, "workspace too small (ct_curve_workspace_bytes)");
  int threads = 256, groups = 0;
  if (schedule) {
    threads = schedule->threads;
    groups = schedule->groups;
  }
  CT_CHECK_ARG(groups >= 0, "schedule.groups must be >= 0");
  if (groups == 0) {
    int dev = 0, cus = 0;
    CT_HIP(hipGetDevice(&dev));
    CT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    groups = 8 * cus;
  }
  if (groups > batch->num_problems) groups = batch->num_problems;
  CurveArgs a;
  a.B = batch->num_problems;
  a.num_samples = batch->num_samples;
  a.offsets = batch->offsets;
  a.x = batch->x;
  a.y = batch->y;
  a.order = batch->order;
  a.loss = batch->loss;
  a.opt = *options;
  a.params = params;
  a.results = results;
  a.records = records;
  a.queue = static_cast<unsigned int*>(workspace);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, s, a.queue);
  CT_LAUNCH_CHECK();
  return options->precision == CT_MIXED ? launch_curve<float>(a, threads, groups, s)
                                        : launch_curve<double>(a, threads, groups, s);
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
