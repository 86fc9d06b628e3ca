// Device-resident Levenberg-Marquardt bookkeeping kernels of the bundle-
// adjustment solve (solver.py:35-357).
//
// An LM iteration t is the kernel sequence
//   eval (observations, points) -> camera blocks -> lm_begin -> RCS solve
//   -> step (points) -> candidate (observations) -> lm_end
// lm_begin reduces the cost/gradient partials of the evaluation at x and
// lm_end the candidate partials; the trust-region arithmetic itself is in
// lm_logic.h.  Every kernel returns immediately once state->done is set, so
// a graph of several iterations can be replayed past convergence.
#include "ba_kernels.h"
#include "lm_logic.h"

namespace ct {

namespace {

constexpr int kLMThreads = 1024;
constexpr int kLMBatch = 4;  // partials per thread whose loads are issued before their adds

// Fixed-order partial sums over a strided range: thread t takes i = t, t + T, ...
// in batches of kLMBatch (all loads of a batch in flight together).
template <int kWidth>
__device__ __forceinline__ void strided_sums(const double* __restrict__ part, int count, double (&acc)[kWidth]) {
  for (int i0 = threadIdx.x; i0 < count; i0 += kLMBatch * kLMThreads) {
    double v[kLMBatch][kWidth];
#pragma unroll
    for (int q = 0; q < kLMBatch; ++q) {
      const int i = i0 + q * kLMThreads;
#pragma unroll
      for (int k = 0; k < kWidth; ++k) v[q][k] = i < count ? part[(size_t)kWidth * i + k] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < kLMBatch; ++q)
#pragma unroll
      for (int k = 0; k < kWidth; ++k) acc[k] += v[q][k];
  }
}

// Thread 0 updates an LDS copy of the state, loaded word-parallel at kernel
// start (in flight during the partial sums; visible to thread 0 after the
// block reductions' barriers) and stored back once: on the global state the
// update's branches would wait for their fields one load at a time.
constexpr int kStateWords = (int)(sizeof(LMState) / sizeof(uint32_t));
static_assert(sizeof(LMState) % sizeof(uint32_t) == 0 && kStateWords <= kLMThreads, "state copy");

__device__ __forceinline__ void load_state(LMState& dst, const LMState* src) {
  if (threadIdx.x < kStateWords)
    reinterpret_cast<uint32_t*>(&dst)[threadIdx.x] = reinterpret_cast<const uint32_t*>(src)[threadIdx.x];
}

__device__ __forceinline__ void store_state(LMState* dst, const LMState& src) {
  for (int i = 0; i < kStateWords; ++i) reinterpret_cast<uint32_t*>(dst)[i] = reinterpret_cast<const uint32_t*>(&src)[i];
}

__global__ void k_lm_init(LMInit init, BAWork w, const double* __restrict__ ucams, const double* __restrict__ upts) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int stride = gridDim.x * blockDim.x;
  // both ping-pong copies: a point without observations is never stepped (the
  // fused step moves a point from its first observation's lane), so it must
  // read the same value from either copy
  for (int i = t; i < init.n_cam_params; i += stride) {
    const double v = ucams[i];
    w.xcam[i] = v;
    w.xcam[init.n_cam_params + i] = v;
  }
  for (int i = t; i < init.n_pt_params; i += stride) {
    const double v = upts[i];
    w.xpt[i] = v;
    w.xpt[init.n_pt_params + i] = v;
  }
  if (t == 0) {
    lm_state_init(w.state, init.opt);
    w.state->split_blocks = init.split_blocks;
  }
}

__global__ __launch_bounds__(kLMThreads) void k_lm_begin(LMInit init, BAWork w) {
  __shared__ double red[kLMThreads / kWave];
  LMState* st = w.state;
  if (st->done) return;
  __shared__ LMState ls;
  load_state(ls, st);
  double c[1] = {0.0}, g2 = 0.0, gm = 0.0;
  strided_sums<1>(w.part_cost, init.num_cost_blocks, c);
  for (int i0 = threadIdx.x; i0 < init.num_eval_blocks; i0 += kLMBatch * kLMThreads) {
    double v[kLMBatch][2];
#pragma unroll
    for (int q = 0; q < kLMBatch; ++q) {
      const int i = i0 + q * kLMThreads;
      v[q][0] = i < init.num_eval_blocks ? w.part_eval[2 * i] : 0.0;
      v[q][1] = i < init.num_eval_blocks ? w.part_eval[2 * i + 1] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < kLMBatch; ++q) {
      g2 += v[q][0];
      gm = fmax(gm, v[q][1]);
    }
  }
  // camera gradient blocks: |g_c|^2 and max|g_c| per camera, then across cameras
  for (int i = threadIdx.x; i < init.num_cameras; i += kLMThreads) {
    double gv[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) gv[j] = w.gc[9 * (size_t)i + j];
    double c2 = 0.0, cm = 0.0;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
      c2 += gv[j] * gv[j];
      cm = fmax(cm, fabs(gv[j]));
    }
    g2 += c2;
    gm = fmax(gm, cm);
  }
  const double cost = block_sum<kLMThreads>(c[0], red);
  const double gnorm2 = block_sum<kLMThreads>(g2, red);
  const double gmax = block_max<kLMThreads>(gm, red);
  if (threadIdx.x != 0) return;
  if (w.pcg_bar) *w.pcg_bar += 1u;  // a new epoch for the barrier numbers of this iteration's PCG
  lm_begin_update(&ls, cost, gnorm2, gmax, w.records);
  store_state(st, ls);
}

__global__ __launch_bounds__(kLMThreads) void k_lm_end(LMInit init, BAWork w) {
  __shared__ double red[kLMThreads / kWave];
  LMState* st = w.state;
  // every thread reads the flags before thread 0 may change them
  const int done = st->done;
  const int skip = st->skip_step;
  __shared__ LMState ls;
  load_state(ls, st);
  __syncthreads();
  if (done) return;
  if (skip) {
    if (threadIdx.x == 0) {
      lm_skip_update(&ls);
      store_state(st, ls);
    }
    return;
  }
  double cand[2] = {0.0, 0.0}, step[3] = {0.0, 0.0, 0.0};
  strided_sums<2>(w.part_cand, init.num_obs_blocks, cand);
  strided_sums<3>(w.part_step, init.num_step_blocks, step);
  const double a[5] = {cand[0], cand[1], step[0], step[1], step[2]};
  double s[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) s[k] = block_sum<kLMThreads>(a[k], red);
  if (threadIdx.x != 0) return;
  lm_end_update(&ls, s[0], s[1], s[2], s[3], s[4] != 0.0, w.records);
  store_state(st, ls);
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
